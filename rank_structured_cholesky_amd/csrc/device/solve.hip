// Preconditioner apply M^{-1} = P^T L^{-T} L^{-1} P on the device
// (FactorPreconditioner::solve, src/rschol/solver/FactorPreconditioner.cpp:
// 135-162; FactorManager::forwardSolve / backwardSolve, FactorManager.cpp:
// 6245-6341; Supernode::forwardSolve / backwardSolve, Supernode.cpp:
// 4571-5250; DiagonalBlock::BlockNode::diagonalSolve, DiagonalBlock.cpp:
// 1306-1378) and the CSR SpMV of the PCG operator.
//
// Level-scheduled and task-parallel: per elimination-tree level and
// direction a few phases, each one launch of independent matrix-vector
// tasks (a task = up to 256 rows or 32 columns of one factor block), so the
// large blocks near the top of the tree spread over all compute units.
// Diagonal blocks are applied through stored inverses (L^{-T}); compressed
// diagonals use one leaf step plus one low-rank correction per tree height
// (precomputed W = L_B^{-1} V, X = L_A^{-T} U).  The forward sweep is
// multifrontal: a supernode's front vector receives its children's update
// vectors (no atomics).  Vectors are FP64; the factor may be FP32 (widened
// on load).
#include "solve.h"

#include <algorithm>

namespace rsc {
namespace dev {

constexpr int MAXR = 1024;  // largest rank handled in LDS by the solves
constexpr int kXlds = 2048;  // x entries staged in LDS per ROW batch (dataflow tasks)
constexpr int kColGroup = 8;  // columns per pass of a single-RHS column task's wave
// Loads of the single-RHS task bodies: an FP32 factor is loaded
// unconditionally (a clamped in-range address, the value masked afterwards),
// since predicated narrow loads were serialized through one register; the
// FP64 bodies keep predicated loads, which need fewer live registers.

__global__ void k_gather(const double* __restrict__ b, int ldb, double* __restrict__ w,
                         const int* __restrict__ perm, int n, int k,
                         const int* __restrict__ live) {
  if (live && *live == 0) return;
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * k) return;
  int i = (int)(t % n), j = (int)(t / n);
  w[i + (int64_t)j * n] = b[perm[i] + (int64_t)j * ldb];
}

__global__ void k_scatter(const double* __restrict__ w, double* __restrict__ x, int ldx,
                          const int* __restrict__ perm, int n, int k,
                          const int* __restrict__ live) {
  if (live && *live == 0) return;
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * k) return;
  int i = (int)(t % n), j = (int)(t / n);
  x[perm[i] + (int64_t)j * ldx] = w[i + (int64_t)j * n];
}

void launch_gather(const double* b, int ldb, double* w, const int* perm, int n,
                   int k, const int* live, hipStream_t s) {
  int64_t tot = (int64_t)n * k;
  k_gather<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(b, ldb, w, perm, n, k, live);
}

void launch_scatter(const double* w, double* x, int ldx, const int* perm, int n,
                    int k, const int* live, hipStream_t s) {
  int64_t tot = (int64_t)n * k;
  k_scatter<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(w, x, ldx, perm, n, k, live);
}

// ---- vector gathers of a level -------------------------------------------

// Front vectors are processed in chunks of kSolveChunk entries (blockIdx.y),
// so a large supernode spreads over many workgroups.  In the forward assembly a
// chunk adds, child by child, the child entries whose parent position falls
// in the chunk: deterministic, no atomics.
constexpr int VCH = kSolveChunk;  // front-vector entries per assembly / gather task
// right-hand sides per LDS pass of a front-vector assembly: all sixteen of
// a matrix-core pass, so a task reads each child entry in one round trip
constexpr int kAsmRhs = 16;

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Every field is read unconditionally and the choice made by selects on the
// values: a choice between the fields' addresses would keep the argument
// block in private (scratch) memory.
__device__ inline double* vsel(const VecBufs& b, int sel, int col) {
  double* const w = b.w;
  double* const fv = b.fv;
  double* const tv = b.tv;
  double* const z = b.z;
  const int64_t n = b.n, fvs = b.fvs, tvs = b.tvs;
  double* const base = sel == 0 ? w : (sel == 1 ? fv : (sel == 2 ? tv : z));
  const int64_t st = sel == 1 ? fvs : (sel == 2 ? tvs : n);
  return base + (int64_t)col * st;
}

// Task descriptors are the same for every lane of a workgroup; copying them
// through readfirstlane puts them in scalar registers, so the compiler sees
// uniform control flow around the workgroup barriers of the task bodies.
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline int64_t uni64(int64_t v) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v & 0xffffffff));
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
// The thread index as a value the compiler cannot hoist: the task bodies of
// the persistent dataflow kernel run inside its task loop, and values
// derived from a plain threadIdx.x (LDS and vector addresses of every body)
// would be computed once before the loop and kept live across all bodies,
// beyond the register budget of four resident workgroups per compute unit.
__device__ inline int body_tid() {
  int t = (int)threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}
__device__ inline GvTask uniform_task(const GvTask& t) {
  GvTask u;
  u.mat = uni64(t.mat);
  u.ld = uni(t.ld);
  u.rows = uni(t.rows);
  u.cols = uni(t.cols);
  u.kind = uni(t.kind);
  u.op = uni(t.op);
  u.xsel = uni(t.xsel);
  u.ysel = uni(t.ysel);
  u.part = uni(t.part);
  u.xpart = uni(t.xpart);
  u.xnsub = uni(t.xnsub);
  u.xoff = uni64(t.xoff);
  u.yoff = uni64(t.yoff);
  return u;
}
__device__ inline VecNode uniform_node(const VecNode& n) {
  VecNode u;
  u.first = uni(n.first);
  u.ncols = uni(n.ncols);
  u.R = uni(n.R);
  u.nchild = uni(n.nchild);
  u.child_off = uni(n.child_off);
  u.range_off = uni(n.range_off);
  u.fv_off = uni64(n.fv_off);
  u.rows_off = uni64(n.rows_off);
  u.relmap_off = uni64(n.relmap_off);
  return u;
}

// Vector accesses of the task bodies.  COH = true (dataflow launch): data
// that other workgroups of the same launch produce or consume goes through
// device-scope loads and stores, which bypass the per-XCD caches, so no
// cache-wide writeback or invalidate is needed between tasks.
template <bool COH>
__device__ inline double vld(const double* p) {
  if constexpr (COH) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else return *p;
}
template <bool COH>
__device__ inline void vst(double* p, double v) {
  if constexpr (COH) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else *p = v;
}

// Entry j of a ROW task's x window for right-hand side col: the vector
// itself or, when the producing column task's reduction is fused here
// (xpart >= 0), the sum of its partial slots in slot order.
template <bool COH>
__device__ inline double x_entry(const VecBufs& b, const GvTask& t, int col, int j) {
  if (t.xpart < 0) return vld<COH>(vsel(b, t.xsel, col) + t.xoff + j);
  const double* pp = b.part + (int64_t)col * b.ps +
                     ((int64_t)t.xpart + (int64_t)(j >> 5) * t.xnsub) * 32 + (j & 31);
  double acc = 0;
#pragma unroll 4
  for (int s2 = 0; s2 < t.xnsub; ++s2) acc += vld<COH>(pp + (int64_t)s2 * 32);
  return acc;
}

// Forward assembly of one 1024-entry chunk of node S's front vector:
// fv_c = source[perm[cols]] (perm may be null: source indexed directly),
// fv_R = 0, then the children's update vectors added through their relmaps.
// Children whose descriptors an assembly task keeps in LDS for all its
// passes (more are read again per pass).
constexpr int kAsmChildren = 32;
struct AsmChild {
  int64_t rm_off;  // relmap offset of the child
  int64_t fc_off;  // its update vector (fv_R) offset, first right-hand side
  int a0, cr;      // its rows that land in this chunk: [a0, a0 + cr)
};

// SPAN consecutive chunks (span index `chunk`) form one task: a child's
// rows landing in them are contiguous, from the first chunk's range start
// to the last chunk's range end.
template <bool COH, int SPAN = 1>
__device__ void assemble_chunk(const VecNode& S, int chunk, const VecNode* __restrict__ nodes,
                               const int* __restrict__ children,
                               const int* __restrict__ relmap,
                               const int* __restrict__ ranges, const VecBufs& b, int k,
                               const double* __restrict__ src, int64_t lds,
                               const int* __restrict__ perm, double* __restrict__ sm) {
  // The span of kAsmRhs right-hand sides at a time is summed in LDS (sm:
  // kAsmRhs * SPAN * VCH doubles, at most what the caller's area holds):
  // its own entries, then child by child (a child's rows map to distinct
  // positions), then written once.  Each entry's sum order is the same for
  // any k.  The children's descriptors and the own entries' source indices
  // are read once for all passes.
  constexpr int WS = SPAN * VCH;  // entries of a span
  __shared__ AsmChild s_ch[kAsmChildren];
  const int c = S.ncols, R = S.R, tid = body_tid();
  const int p0 = chunk * WS, p1 = min(p0 + WS, c + R);
  if (p0 >= c + R) return;
  const int len = p1 - p0;
  const int nch = S.nchild;
  const int q0 = chunk * SPAN, q1 = (p1 - 1) / VCH;  // first and last chunk of the span
  const bool cached = nch <= kAsmChildren;
  if (cached && tid < nch) {
    const VecNode C = nodes[children[S.child_off + tid]];
    const int* rg0 = ranges + 2 * ((int64_t)S.range_off + (int64_t)q0 * nch + tid);
    const int* rg1 = ranges + 2 * ((int64_t)S.range_off + (int64_t)q1 * nch + tid);
    s_ch[tid] = AsmChild{C.relmap_off, C.fv_off + C.ncols, rg0[0], rg1[1] - rg0[0]};
  }
  constexpr int PER = WS / 256;
  int own[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int i = p0 + tid + 256 * u;
    own[u] = i < p1 && i < c ? (perm ? perm[S.first + i] : S.first + i) : -1;
  }
  for (int c0 = 0; c0 < k; c0 += kAsmRhs) {
    const int g = min(kAsmRhs, k - c0);
    if constexpr (PER == 1) {
      // the thread's entry for every right-hand side of the pass, loaded
      // together (clamped in-range addresses, masked values)
      double v[kAsmRhs];
      const int o = own[0];
#pragma unroll
      for (int cc = 0; cc < kAsmRhs; ++cc) {
        const bool ok = cc < g && o >= 0;
        const double x = src[ok ? (int64_t)(c0 + cc) * lds + o : 0];
        v[cc] = ok ? x : 0.0;
      }
      if (tid < len)
#pragma unroll
        for (int cc = 0; cc < kAsmRhs; ++cc) sm[cc * WS + tid] = v[cc];
    } else {
      for (int cc = 0; cc < g; ++cc) {
        const double* sc = src + (int64_t)(c0 + cc) * lds;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
          const int i = tid + 256 * u;
          if (i < len) sm[cc * WS + i] = own[u] >= 0 ? sc[own[u]] : 0.0;
        }
      }
    }
    __syncthreads();
    for (int q = 0; q < nch; ++q) {
      int64_t rm_off, fc_off;
      int a0, cr;
      if (cached) {
        rm_off = uni64(s_ch[q].rm_off);
        fc_off = uni64(s_ch[q].fc_off);
        a0 = uni(s_ch[q].a0);
        cr = uni(s_ch[q].cr);
      } else {
        const VecNode C = uniform_node(nodes[uni(children[S.child_off + q])]);
        const int* rg0 = ranges + 2 * ((int64_t)S.range_off + (int64_t)q0 * nch + q);
        const int* rg1 = ranges + 2 * ((int64_t)S.range_off + (int64_t)q1 * nch + q);
        rm_off = C.relmap_off;
        fc_off = C.fv_off + C.ncols;
        a0 = uni(rg0[0]);
        cr = uni(rg1[1]) - a0;
      }
      const int* rm = relmap + rm_off;
      for (int t = tid; t < cr; t += 256) {
        const int a = a0 + t;
        const int p = rm[a] - p0;
        double v[kAsmRhs];
#pragma unroll
        for (int cc = 0; cc < kAsmRhs; ++cc)
          v[cc] = cc < g ? vld<COH>(b.fv + fc_off + (int64_t)(c0 + cc) * b.fvs + a) : 0.0;
#pragma unroll
        for (int cc = 0; cc < kAsmRhs; ++cc)
          if (cc < g) sm[cc * WS + p] += v[cc];
      }
      __syncthreads();
    }
    for (int idx = tid; idx < g * len; idx += 256) {
      const int cc = idx / len, i = idx - cc * len;
      vst<COH>(b.fv + S.fv_off + (int64_t)(c0 + cc) * b.fvs + p0 + i, sm[cc * WS + i]);
    }
    __syncthreads();
  }
}

// Backward gather of one chunk: fv_R[a] = w[rows[a]] (the solution at the
// node's off-diagonal rows) and w[cols] = z[cols] (start of x_c); FLAT runs
// all right-hand sides in one pass.
template <bool COH, bool FLAT = false, int SPAN = 1>
__device__ void gather_chunk(const VecNode& S, int chunk, const int* __restrict__ rows,
                             const VecBufs& b, int k) {
  const int c = S.ncols, R = S.R, tid = body_tid();
  const int p0 = chunk * SPAN * VCH, p1 = min(p0 + SPAN * VCH, c + R);
  if (p0 >= c + R) return;
  const int* rw = rows + S.rows_off;
  if (!FLAT) {
    for (int col = 0; col < k; ++col) {
      double* f = b.fv + S.fv_off + (int64_t)col * b.fvs;
      double* wc = b.w + (int64_t)col * b.n;
      const double* zc = b.z + (int64_t)col * b.n;
      for (int i = p0 + tid; i < p1; i += 256) {
        if (i < c) vst<COH>(wc + S.first + i, vld<COH>(zc + S.first + i));
        else vst<COH>(f + i, vld<COH>(wc + rw[i - c]));
      }
    }
    return;
  }
  // a thread's entry index is loaded once for every right-hand side, and
  // the loads of eight right-hand sides go out together
  for (int i = p0 + tid; i < p1; i += 256) {
    const bool own = i < c;
    const int64_t src = own ? (int64_t)S.first + i : (int64_t)rw[i - c];
    const double* from = own ? b.z : b.w;
    double* to = own ? b.w + S.first + i : b.fv + S.fv_off + i;
    const int64_t ts = own ? b.n : b.fvs;
    int col = 0;
    for (; col + 8 <= k; col += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = vld<COH>(from + (int64_t)(col + u) * b.n + src);
#pragma unroll
      for (int u = 0; u < 8; ++u) vst<COH>(to + (int64_t)(col + u) * ts, v[u]);
    }
    for (; col < k; ++col) vst<COH>(to + (int64_t)col * ts, vld<COH>(from + (int64_t)col * b.n + src));
  }
}

__global__ __launch_bounds__(256) void k_fwd_assemble(const VecNode* __restrict__ nodes,
                                                      const int* __restrict__ lnodes,
                                                      const int* __restrict__ children,
                                                      const int* __restrict__ relmap,
                                                      const int* __restrict__ ranges,
                                                      VecBufs b, int k) {
  __shared__ double sm[kAsmRhs * VCH];
  if (b.live && *b.live == 0) return;
  const VecNode S = nodes[lnodes[blockIdx.x]];
  assemble_chunk<false>(S, blockIdx.y, nodes, children, relmap, ranges, b, k, b.w, b.n,
                        nullptr, sm);
}

__global__ __launch_bounds__(256) void k_bwd_gather(const VecNode* __restrict__ nodes,
                                                    const int* __restrict__ lnodes,
                                                    const int* __restrict__ rows,
                                                    VecBufs b, int k) {
  if (b.live && *b.live == 0) return;
  const VecNode S = nodes[lnodes[blockIdx.x]];
  gather_chunk<false>(S, blockIdx.y, rows, b, k);
}

void launch_fwd_assemble(const VecNode* nodes, const int* level_nodes, int count,
                         const int* children, const int* relmap, const int* ranges,
                         VecBufs b, int k, int max_front, hipStream_t s) {
  if (count <= 0) return;
  const int ny = (max_front + VCH - 1) / VCH;
  k_fwd_assemble<<<dim3(count, ny), 256, 0, s>>>(nodes, level_nodes, children, relmap,
                                                  ranges, b, k);
}

void launch_bwd_gather(const VecNode* nodes, const int* level_nodes, int count,
                       const int* rows, VecBufs b, int k, int max_front, hipStream_t s) {
  if (count <= 0) return;
  const int ny = (max_front + VCH - 1) / VCH;
  k_bwd_gather<<<dim3(count, ny), 256, 0, s>>>(nodes, level_nodes, rows, b, k);
}

// ---- matrix-vector tasks ---------------------------------------------------

// Small tasks: plain loops, few registers (many workgroups per CU).
template <typename T>
__global__ __launch_bounds__(256) void k_gemv_small(const T* __restrict__ F,
                                                    const GvTask* __restrict__ tasks,
                                                    VecBufs b, int k) {
  __shared__ double red[4][64];
  if (b.live && *b.live == 0) return;
  const GvTask t = tasks[blockIdx.x];
  const T* __restrict__ M = F + t.mat;
  const int tid = body_tid(), lane = tid & 63, wave = tid >> 6;
  for (int col = 0; col < k; ++col) {
    double* y = vsel(b, t.ysel, col) + t.yoff;
    if (t.kind == 2) {
      const double* pp = b.part + (int64_t)col * b.ps + (int64_t)t.part * 32;
      if (tid < t.cols) {
        double acc = 0;
        for (int s2 = 0; s2 < t.rows; ++s2) acc += pp[(int64_t)s2 * 32 + tid];
        if (t.op) y[tid] -= acc;
        else y[tid] = acc;
      }
      continue;
    }
    const double* x = vsel(b, t.xsel, col) + t.xoff;
    if (t.kind == 0) {
      double acc = 0;
      if (lane < t.rows) {
        const T* row = M + lane;
        for (int j = wave; j < t.cols; j += 4) acc += (double)row[(int64_t)j * t.ld] * x[j];
      }
      red[wave][lane] = acc;
      __syncthreads();
      if (wave == 0 && lane < t.rows) {
        const double v = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
        if (t.op) y[lane] -= v;
        else y[lane] = v;
      }
      __syncthreads();
    } else {
      const int q0 = wave * 8;
      const int nq = min(8, t.cols - q0);
      for (int u = 0; u < nq; ++u) {
        const T* cp = M + (int64_t)(q0 + u) * t.ld;
        double acc = 0;
        for (int i = lane; i < t.rows; i += 64) acc += (double)cp[i] * x[i];
        acc = wave_sum(acc);
        if (lane == 0) {
          if (t.part >= 0) b.part[(int64_t)col * b.ps + (int64_t)t.part * 32 + q0 + u] = acc;
          else if (t.op) y[q0 + u] -= acc;
          else y[q0 + u] = acc;
        }
      }
    }
  }
}

// Large tasks: every load of a block issued before its first use.
template <typename T>
__device__ void gemv_task(const T* __restrict__ F, const GvTask& t, const VecBufs& b, int k,
                          double (*red)[64]) {
  const T* __restrict__ M = F + t.mat;
  const int tid = body_tid(), lane = tid & 63, wave = tid >> 6;
  for (int col = 0; col < k; ++col) {
    double* y = vsel(b, t.ysel, col) + t.yoff;
    if (t.kind == 2) {
      // reduce: y[j] (op)= sum_s part[first + s][j]
      const double* pp = b.part + (int64_t)col * b.ps + (int64_t)t.part * 32;
      if (tid < t.cols) {
        double acc = 0;
        for (int s2 = 0; s2 < t.rows; ++s2) acc += pp[(int64_t)s2 * 32 + tid];
        if (t.op) y[tid] -= acc;
        else y[tid] = acc;
      }
      continue;
    }
    const double* x = vsel(b, t.xsel, col) + t.xoff;
    if (t.kind == 0) {
      // ROW: 64 rows per task.  Lane = row, wave g takes columns jb+g+4m
      // (m < 32) of each 128-column block; all 32 loads are issued before
      // the first use (one memory latency per block), then an LDS reduction.
      double acc = 0;
      const bool act = lane < t.rows;
      const T* row = M + lane;
      for (int jb = 0; jb < t.cols; jb += 128) {
        double mv[32];
#pragma unroll
        for (int m = 0; m < 32; ++m) {
          const int j = jb + wave + 4 * m;
          mv[m] = (act && j < t.cols) ? (double)row[(int64_t)j * t.ld] : 0.0;
        }
#pragma unroll
        for (int m = 0; m < 32; ++m) {
          const int j = jb + wave + 4 * m;
          if (j < t.cols) acc += mv[m] * x[j];
        }
      }
      red[wave][lane] = acc;
      __syncthreads();
      if (wave == 0 && act) {
        const double v = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
        if (t.op) y[lane] -= v;
        else y[lane] = v;
      }
      __syncthreads();
    } else {
      // COL: up to 512 rows x 32 columns.  Wave w owns columns q0..q0+7,
      // lane l rows l + 64m (m < 8): 64 loads issued up front, then FMAs and
      // 8 wavefront reductions.
      const int q0 = wave * 8;
      const int nq = min(8, t.cols - q0);
      double acc[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc[u] = 0;
      if (nq > 0) {
        for (int ib = 0; ib < t.rows; ib += 512) {
          double mv[8][8], xv[8];
#pragma unroll
          for (int m = 0; m < 8; ++m) {
            const int i = ib + lane + 64 * m;
            const bool ok = i < t.rows;
            xv[m] = ok ? x[i] : 0.0;
            const T* base = M + i + (int64_t)q0 * t.ld;
#pragma unroll
            for (int u = 0; u < 8; ++u)
              mv[m][u] = (ok && u < nq) ? (double)base[(int64_t)u * t.ld] : 0.0;
          }
#pragma unroll
          for (int m = 0; m < 8; ++m)
#pragma unroll
            for (int u = 0; u < 8; ++u) acc[u] += mv[m][u] * xv[m];
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const double v = wave_sum(acc[u]);
        if (lane == 0 && u < nq) {
          if (t.part >= 0) {
            b.part[(int64_t)col * b.ps + (int64_t)t.part * 32 + q0 + u] = v;
          } else if (t.op) {
            y[q0 + u] -= v;
          } else {
            y[q0 + u] = v;
          }
        }
      }
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_gemv_tasks(const T* __restrict__ F,
                                                    const GvTask* __restrict__ tasks,
                                                    VecBufs b, int k) {
  __shared__ double red[4][64];
  if (b.live && *b.live == 0) return;
  const GvTask t = tasks[blockIdx.x];
  gemv_task<T>(F, t, b, k, red);
}

template <typename T>
void launch_gemv_tasks(const T* factor, const GvTask* tasks, int count, VecBufs b,
                       int k, bool big, hipStream_t s) {
  if (count <= 0) return;
  if (big) k_gemv_tasks<T><<<count, 256, 0, s>>>(factor, tasks, b, k);
  else k_gemv_small<T><<<count, 256, 0, s>>>(factor, tasks, b, k);
}

template void launch_gemv_tasks<double>(const double*, const GvTask*, int, VecBufs, int,
                                        bool, hipStream_t);
template void launch_gemv_tasks<float>(const float*, const GvTask*, int, VecBufs, int,
                                       bool, hipStream_t);

// ---- dataflow apply -------------------------------------------------------

// Matrix-vector task with few registers (about 64 VGPRs, so several
// workgroups stay resident per compute unit while others wait on their
// dependencies).  ROW: lane = row, wave g takes columns g, g+4, ..., eight
// per batch; COL: wave w owns columns 8w..8w+7, lane l rows l + 64m, two
// row blocks per batch.
// Matrix-vector task body for any number of right-hand sides: groups of
// KRHS columns share every matrix load, and the arithmetic of each column
// does not depend on the group size, so batched and single solves agree
// bitwise.  ROW: x staged in LDS in chunks of NB entries, wave g takes
// columns g, g+4, ... sixteen at a time (even/odd into two sums); COL: wave
// w owns columns 8w..8w+7 (two halves of four), lanes take rows l and
// l + 64 of every 128-row block.  KRHS: right-hand sides per group.
template <typename T, bool COH, int KRHS>
__device__ void gemv_multi(const T* __restrict__ F, const GvTask& t, const VecBufs& b, int k,
                           double (*red)[64], double* xl) {
  const T* __restrict__ M = F + t.mat;
  const int tid = body_tid(), lane = tid & 63, wave = tid >> 6;
  constexpr int NB = kXlds / KRHS;
  for (int c0 = 0; c0 < k; c0 += KRHS) {
    const int kb = min(KRHS, k - c0);
    if (t.kind == 0) {
      const bool act = lane < t.rows;
      const T* rp = M + (act ? lane : 0);
      const int64_t ld4 = 4 * (int64_t)t.ld;
      double acc0[KRHS], acc1[KRHS];
#pragma unroll
      for (int cc = 0; cc < KRHS; ++cc) acc0[cc] = acc1[cc] = 0;
      for (int jb = 0; jb < t.cols; jb += NB) {
        const int nb = min(NB, t.cols - jb);
        for (int idx = tid; idx < kb * nb; idx += 256) {
          const int cc = idx / nb, jj = idx - cc * nb;
          xl[cc * NB + jj] = x_entry<COH>(b, t, c0 + cc, jb + jj);
        }
        __syncthreads();
        int j = wave;
        for (; j + 60 < nb; j += 64) {
          const T* p = rp + (int64_t)(jb + j) * t.ld;
          // loads kept in the factor type (each its own register), widened
          // at their products
          T m[16];
#pragma unroll
          for (int u = 0; u < 16; ++u) m[u] = p[u * ld4];
#pragma unroll
          for (int cc = 0; cc < KRHS; ++cc)
#pragma unroll
            for (int u = 0; u < 16; u += 2) {
              acc0[cc] += (double)m[u] * xl[cc * NB + j + 4 * u];
              acc1[cc] += (double)m[u + 1] * xl[cc * NB + j + 4 * u + 4];
            }
        }
        if (j < nb) {
          // the last (at most sixteen) columns of this wave: loads together,
          // products in column order
          // (clamped in-range addresses, loads in batches of eight)
#pragma unroll
          for (int h = 0; h < 16; h += 8) {
            T m[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
              m[u] = rp[(int64_t)(jb + min(j + 4 * (h + u), nb - 1)) * t.ld];
#pragma unroll
            for (int u = 0; u < 8; ++u)
              if (j + 4 * (h + u) < nb)
#pragma unroll
                for (int cc = 0; cc < KRHS; ++cc)
                  acc0[cc] += (double)m[u] * xl[cc * NB + j + 4 * (h + u)];
          }
        }
        __syncthreads();
      }
      for (int cc = 0; cc < kb; ++cc) {
        red[wave][lane] = act ? acc0[cc] + acc1[cc] : 0.0;
        __syncthreads();
        if (wave == 0 && act) {
          const double v = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
          if (t.part >= 0) {
            vst<COH>(b.part + (int64_t)(c0 + cc) * b.ps +
                         (int64_t)(t.part + (lane >> 5)) * 32 + (lane & 31), v);
          } else {
            double* y = vsel(b, t.ysel, c0 + cc) + t.yoff;
            vst<COH>(y + lane, t.op ? vld<COH>(y + lane) - v : v);
          }
        }
        __syncthreads();
      }
    } else {
      // x (at most kXlds / KRHS rows) staged in LDS once; wave w takes
      // columns 8w..8w+7 together, lane l rows l and l + 64 of every
      // 128-row block (a column's sum order is that of four at a time)
      for (int idx = tid; idx < kb * t.rows; idx += 256) {
        const int cc = idx / t.rows, i = idx - cc * t.rows;
        xl[cc * NB + i] = vld<COH>(vsel(b, t.xsel, c0 + cc) + t.xoff + i);
      }
      __syncthreads();
      for (int h = 0; h < 8; h += kColGroup) {
      const int q0 = wave * 8 + h;
      const int nq = min(kColGroup, t.cols - q0);
      double acc[kColGroup][KRHS];
#pragma unroll
      for (int u = 0; u < kColGroup; ++u)
#pragma unroll
        for (int cc = 0; cc < KRHS; ++cc) acc[u][cc] = 0;
      if (nq > 0) {
        const T* cbase = M + (int64_t)q0 * t.ld;
        for (int i = lane; i < t.rows; i += 128) {
          const int i2 = i + 64;
          const bool ok2 = i2 < t.rows;
          T m1[kColGroup], m2[kColGroup];
#pragma unroll
          for (int u = 0; u < kColGroup; ++u) {
            const T* cp = cbase + (int64_t)(u < nq ? u : 0) * t.ld;
            m1[u] = cp[i];
            if constexpr (sizeof(T) == 4) m2[u] = cp[ok2 ? i2 : i];  // unused when !ok2
            else m2[u] = ok2 ? cp[i2] : T(0);
          }
#pragma unroll
          for (int cc = 0; cc < KRHS; ++cc) {
            const double x1 = cc < kb ? xl[cc * NB + i] : 0.0;
            const double x2 = (cc < kb && ok2) ? xl[cc * NB + i2] : 0.0;
#pragma unroll
            for (int u = 0; u < kColGroup; ++u)
              acc[u][cc] += (double)m1[u] * x1 + (double)m2[u] * x2;
          }
        }
      }
#pragma unroll
      for (int u = 0; u < kColGroup; ++u)
#pragma unroll
        for (int cc = 0; cc < KRHS; ++cc) {
          const double v = wave_sum(acc[u][cc]);
          if (lane == 0 && u < nq && cc < kb) {
            if (t.part >= 0)
              vst<COH>(b.part + (int64_t)(c0 + cc) * b.ps + (int64_t)t.part * 32 + q0 + u, v);
            else {
              double* y = vsel(b, t.ysel, c0 + cc) + t.yoff + q0 + u;
              vst<COH>(y, t.op ? vld<COH>(y) - v : v);
            }
          }
        }
      }
      __syncthreads();
    }
  }
}

// Matrix-vector task bodies for several right-hand sides on the FP64 matrix
// cores: sixteen right-hand sides per pass over the matrix block, as one
// MFMA 16x16x4 product per four matrix columns (ROW) or rows (COL).  Lane l
// supplies A[l%16][l/16] and B[l/16][l%16]; accumulator element v of lane l
// is D[l/16 + 4v][l%16].  A right-hand side's sums run in the same order for
// any number of right-hand sides and any position in the group of sixteen.
//   ROW (rows <= 64): wave w owns rows 16w..16w+15; x is staged in LDS in
//     chunks of 128 columns (16 x 129 doubles).
//   COL (cols <= 32): the 64-row matrix tile is staged transposed in LDS
//     (64 x 33 doubles) with x (16 x 65); wave w takes columns 16(w&1).. and
//     rows 32(w>>1).. of the tile; the two row halves are added in order.
typedef double d4v __attribute__((ext_vector_type(4)));
constexpr int kMfmaRhs = 16;
constexpr int kRowChunk = 128;
constexpr int kMfmaLds = 64 * 33 + kMfmaRhs * 65;  // doubles of LDS scratch (COL)

__device__ inline d4v mfma64(double a, double b, d4v c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Stride between the right-hand-side columns of vector sel (see vsel).
__device__ inline int64_t vstride(const VecBufs& b, int sel) {
  const int64_t n = b.n, fvs = b.fvs, tvs = b.tvs;
  return sel == 1 ? fvs : (sel == 2 ? tvs : n);
}

// The 16 x kRowChunk window of a ROW task's x (right-hand sides c0 .. c0 +
// kb, entries jb .. jb + nb), eight entries per thread.  Every load of the
// window is issued before the first is used (one memory round trip; with a
// fused reduction, one per partial slot), outside the window a clamped
// in-range address whose value is masked.  Thread tid holds right-hand side
// 2u + (tid >> 7), entry tid & 127 in xv[u].
template <bool COH>
__device__ inline void load_row_x(const VecBufs& b, const GvTask& t, int c0, int kb, int jb,
                                  int nb, double (&xv)[8]) {
  const int tid = body_tid();
  const int jj = tid & (kRowChunk - 1);
  const int ch = uni(tid >> 7);  // the same for every lane of a wave
  const bool jok = jj < nb;
  if (t.xpart < 0) {
    // per load a wave-uniform column start (scalar registers) and the
    // lane's entry as a small offset
    const double* x0 = vsel(b, t.xsel, c0) + t.xoff + jb;
    const int64_t st = vstride(b, t.xsel);
    const unsigned jo = jok ? (unsigned)jj : 0u;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int cc = 2 * u + ch;
      const double* xc = cc < kb ? x0 + cc * st : x0;
      const double v = vld<COH>(xc + jo);
      xv[u] = (cc < kb && jok) ? v : 0.0;
    }
    return;
  }
  // x[j] = sum over s < xnsub of the producing column task's partial slots,
  // in slot order (the reduce task's order)
  const double* p0 = b.part + (int64_t)c0 * b.ps + (int64_t)t.xpart * 32;
  const int j = jb + (jok ? jj : 0);
  const unsigned lo = (unsigned)((j >> 5) * t.xnsub * 32 + (j & 31));
#pragma unroll
  for (int u = 0; u < 8; ++u) xv[u] = 0.0;
  for (int s2 = 0; s2 < t.xnsub; ++s2) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int cc = 2 * u + ch;
      const double* pc = p0 + (int64_t)(cc < kb ? cc : 0) * b.ps;
      v[u] = vld<COH>(pc + (lo + (unsigned)s2 * 32u));
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) xv[u] += v[u];
  }
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int cc = 2 * u + ch;
    if (!(cc < kb && jok)) xv[u] = 0.0;
  }
}

template <typename T, bool COH>
__device__ void gemv_mfma(const T* __restrict__ F, const GvTask& t, const VecBufs& b, int k,
                          double* sm) {
  const T* __restrict__ M = F + t.mat;
  // subtracting tasks read y's old values with their first loads (y is
  // written by no other task while this one runs), so the epilogue only
  // stores
  const bool rmw = t.op && t.part < 0;
  for (int c0 = 0; c0 < k; c0 += kMfmaRhs) {
    // lane quantities per pass (not hoisted out of the pass loop)
    const int tid = body_tid(), lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, kq = lane >> 4;
    const int kb = min(kMfmaRhs, k - c0);
    // this lane's result column (clamped in range; masked at the stores)
    double* const ycol =
        t.part < 0 ? vsel(b, t.ysel, c0 + min(li, kb - 1)) + t.yoff : nullptr;
    double yold[4] = {0, 0, 0, 0};
    if (t.kind == 0) {
      constexpr int LD = kRowChunk + 1;
      const int r = wave * 16 + li;
      const bool rok = r < t.rows;
      const bool wact = wave * 16 < t.rows;
      d4v acc = {0, 0, 0, 0};
      // columns jb + 4u + kq of this lane's row, u < 8 (a quarter chunk)
      // byte offsets from the block's (uniform) start as 32-bit values, one
      // register per load: a task's block spans less than 4 GiB
      const unsigned rlane = ((unsigned)(rok ? r : 0) + (unsigned)kq * (unsigned)t.ld) *
                             (unsigned)sizeof(T);
      const unsigned cstep = (unsigned)t.ld * (unsigned)sizeof(T);
      const char* const Mb = reinterpret_cast<const char*>(M);
      auto load_q = [&](int jb, int nb, T (&m)[8]) {
#pragma unroll
        for (int u = 0; u < 8; ++u)  // past the end: an in-range column times x = 0
          m[u] = *reinterpret_cast<const T*>(
              Mb + (rlane + (unsigned)(jb + min(4 * u, nb - 1 - kq)) * cstep));
      };
      // one accumulator, columns in increasing groups of four
      auto mul_q = [&](const T (&m)[8], const double* xr) {
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = mfma64((double)m[u], xr[4 * u], acc);
      };
      for (int jb = 0; jb < t.cols; jb += kRowChunk) {
        const int nb = min(kRowChunk, t.cols - jb);
        // the chunk's matrix block goes out with the x window (and, at the
        // first chunk, the old y): all four quarters for an FP32 factor,
        // the first two for FP64 (the next two while those are multiplied);
        // quarters past the chunk's end are skipped (their x is zero)
        constexpr bool kAll = sizeof(T) == 4;
        T m0[8], m1[8], m2[8], m3[8];
        if (wact) {
          load_q(jb, nb, m0);
          if (nb > 32) load_q(jb + 32, nb - 32, m1);
          if (kAll && nb > 64) load_q(jb + 64, nb - 64, m2);
          if (kAll && nb > 96) load_q(jb + 96, nb - 96, m3);
        }
        double xv[8];
        load_row_x<COH>(b, t, c0, kb, jb, nb, xv);
        if (jb == 0 && rmw && wact) {
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = wave * 16 + kq + 4 * v;
            yold[v] = vld<COH>(ycol + (row < t.rows ? row : 0));
          }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) sm[(2 * u + (tid >> 7)) * LD + (tid & 127)] = xv[u];
        __syncthreads();
        if (wact) {
          const double* xr = sm + li * LD + kq;
          if constexpr (kAll) {
            mul_q(m0, xr);
            if (nb > 32) mul_q(m1, xr + 32);
            if (nb > 64) mul_q(m2, xr + 64);
            if (nb > 96) mul_q(m3, xr + 96);
          } else {
            mul_q(m0, xr);
            if (nb > 64) load_q(jb + 64, nb - 64, m0);
            if (nb > 32) mul_q(m1, xr + 32);
            if (nb > 96) load_q(jb + 96, nb - 96, m1);
            if (nb > 64) mul_q(m0, xr + 64);
            if (nb > 96) mul_q(m1, xr + 96);
          }
        }
        __syncthreads();
      }
      if (wact)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = wave * 16 + kq + 4 * v;
          if (row < t.rows && li < kb) {
            const double val = acc[v];
            if (t.part >= 0) {
              vst<COH>(b.part + (int64_t)(c0 + li) * b.ps +
                           (int64_t)(t.part + (row >> 5)) * 32 + (row & 31), val);
            } else {
              vst<COH>(ycol + row, rmw ? yold[v] - val : val);
            }
          }
        }
    } else {
      double* Ms = sm;             // [64][33]: row-major transposed tile
      double* xs = sm + 64 * 33;   // [16][65]
      const int g = wave & 1, h = wave >> 1;
      const double* x0 = vsel(b, t.xsel, c0) + t.xoff;
      const int64_t xst = vstride(b, t.xsel);
      d4v acc = {0, 0, 0, 0};
      // a 64-row block: eight matrix entries and four x entries per thread,
      // loaded one block ahead (in flight while the previous block is
      // multiplied), stored to LDS after the barrier that frees it
      T mv[8];
      double xv[4];
      // entry idx = tid + 256 u of a block: row lane, column wave + 4 u (the
      // same for every lane of a wave: a scalar column start per load)
      const int wv = uni(wave);
      auto load_blk = [&](int rb) {
        const int nr = min(64, t.rows - rb);
        const unsigned ro = (unsigned)(lane < nr ? rb + lane : rb);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int cc = wv + 4 * u;
          const T* mc = M + (int64_t)(cc < t.cols ? cc : 0) * t.ld;
          mv[u] = mc[ro];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int cc = wv + 4 * u;
          const double* xc = x0 + (cc < kb ? cc : 0) * xst;
          const double v = vld<COH>(xc + ro);
          xv[u] = (lane < nr && cc < kb) ? v : 0.0;
        }
      };
      auto store_blk = [&](int rb) {
        const int nr = min(64, t.rows - rb);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int idx = tid + 256 * u, rr = idx & 63, cc = idx >> 6;
          Ms[rr * 33 + cc] = (rr < nr && cc < t.cols) ? (double)mv[u] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int idx = tid + 256 * u, rr = idx & 63, cc = idx >> 6;
          xs[cc * 65 + rr] = xv[u];
        }
      };
      load_blk(0);
      if (rmw && h == 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int col = 16 * g + kq + 4 * v;
          yold[v] = vld<COH>(ycol + (col < t.cols ? col : 0));
        }
      }
      for (int rb = 0; rb < t.rows; rb += 64) {
        store_blk(rb);
        __syncthreads();
        if (rb + 64 < t.rows) load_blk(rb + 64);
#pragma unroll
        for (int q = 0; q < 32; q += 4) {
          const int rr = 32 * h + q + kq;
          acc = mfma64(Ms[rr * 33 + 16 * g + li], xs[li * 65 + rr], acc);
        }
        __syncthreads();
      }
      // rows 32..63 of every tile (h = 1) are added to rows 0..31 (h = 0)
      if (h == 1)
#pragma unroll
        for (int v = 0; v < 4; ++v) sm[(g * 4 + v) * 64 + lane] = acc[v];
      __syncthreads();
      if (h == 0)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const double val = acc[v] + sm[(g * 4 + v) * 64 + lane];
          const int col = 16 * g + kq + 4 * v;
          if (col < t.cols && li < kb) {
            if (t.part >= 0)
              vst<COH>(b.part + (int64_t)(c0 + li) * b.ps + (int64_t)t.part * 32 + col, val);
            else
              vst<COH>(ycol + col, rmw ? yold[v] - val : val);
          }
        }
      __syncthreads();
    }
  }
}

template <typename T, bool COH, bool MULTI>
__device__ void gemv_lite(const T* __restrict__ F, const GvTask& t0, const VecBufs& b, int k,
                          double (*red)[64], double* xl) {
  const GvTask t = uniform_task(t0);
  if (t.kind != 2) {
    if constexpr (MULTI) {
      gemv_mfma<T, COH>(F, t, b, k, xl);
      return;
    }
    // chunks of NB = kXlds / KRHS (multiples of 64) keep every column's sum
    // order the same for one and four right-hand sides per group
    if (k == 1) gemv_multi<T, COH, 1>(F, t, b, k, red, xl);
    else gemv_multi<T, COH, 4>(F, t, b, k, red, xl);
    return;
  }
  const int tid = body_tid();
  // reduce: y[j] (op)= sum_s part[s][j] for every (right-hand side, j)
  // pair at once, eight partial loads in flight per item; the sum runs
  // over s in order, so the result does not depend on k
  const int64_t rs = 32 * (int64_t)(t.ld > 0 ? t.ld : 1);
  const int items = k * t.cols;
  for (int it = tid; it < items; it += 256) {
    const int col = it / t.cols, j = it - col * t.cols;
    const double* pp = b.part + (int64_t)col * b.ps + (int64_t)t.part * 32 + j;
    double acc = 0;
    int s2 = 0;
    for (; s2 + 4 <= t.rows; s2 += 4) {
      double v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = vld<COH>(pp + (int64_t)(s2 + u) * rs);
#pragma unroll
      for (int u = 0; u < 4; ++u) acc += v[u];
    }
    for (; s2 < t.rows; ++s2) acc += vld<COH>(pp + (int64_t)s2 * rs);
    double* y = vsel(b, t.ysel, col) + t.yoff + j;
    vst<COH>(y, t.op ? vld<COH>(y) - acc : acc);
  }
}

// ---- subtree tasks ----------------------------------------------------------
// One workgroup runs every supernode of a subtree in turn (children first
// forward, parents first backward); the node's vectors live in LDS, the
// factor streams once.  Per node (uncompressed panel [Lc^{-T}; Zp]):
//   forward:  fv_c = b[perm], fv_R = 0 + children's fv_R;  y = Lc^{-1} fv_c;
//             fv_R -= Zp fv_c;  z = Lc^{-T} y
//   backward: x_R = w[rows];  x_c = z - Zp^T x_R
// Only the subtree root's fv_R, z and x leave the workgroup; with COH they
// are written with device-scope stores.
constexpr int kSubCols = 4;  // columns per wave in a subtree task's column products
constexpr int kSubtreeOwn = 1024;  // subtree columns preloaded by a one-RHS task

// LDS vectors of a subtree task for KS right-hand sides per pass.
template <int KS>
struct SubtreeLdsK {
  double c[KS][kSubtreeMaxCols];
  double y[KS][kSubtreeMaxCols];
  double x[KS][kSubtreeMaxRows];
  // one right-hand side: the subtree's own entries loaded once at the start
  // (forward: b[perm[cols]]; backward: z[cols]) when its columns are
  // contiguous and at most kSubtreeOwn
  double own[KS == 1 ? kSubtreeOwn : 1];
};
using SubtreeLds = SubtreeLdsK<kSubtreeRhs>;

// Right-hand sides in groups of KS (1 or kSubtreeRhs) per pass over the
// panels; each right-hand side's sums run in the same order for any KS.
template <typename T, bool COH, int KS>
__device__ void subtree_forward_g(const T* __restrict__ F, const SubtreeArgs& a, int sb,
                                  const VecBufs& b, int k, SubtreeLdsK<KS>& L) {
  const int tid = body_tid(), lane = tid & 63, wave = tid >> 6;
  const int q0 = uni(a.ptr[sb]), q1 = uni(a.ptr[sb + 1]);
  const int rootn = uni(a.nodes[q1 - 1]);
  // one right-hand side: every own entry b[perm[col]] of the subtree at once
  const int col_lo = uni(a.cols[2 * sb]), col_n = uni(a.cols[2 * sb + 1]);
  const bool pre = KS == 1 && col_n >= 0 && col_n <= kSubtreeOwn;
  if (pre) {
    for (int i = tid; i < col_n; i += 256) L.own[i] = a.B[a.perm[col_lo + i]];
    __syncthreads();
  }
  for (int qi = q0; qi < q1; ++qi) {
    const int j = uni(a.nodes[qi]);
    const VecNode S = uniform_node(a.vn[j]);
    const int64_t P = uni64(a.panel[j]);
    const int c = S.ncols, R = S.R, ld = c + R;
    const T* __restrict__ Lt = F + P;
    for (int c0 = 0; c0 < k; c0 += KS) {
      const int kb = min(KS, k - c0);
      // fv_c in L.c and fv_R in L.x, both summed in LDS
      for (int cc = 0; cc < kb; ++cc) {
        const double* bc = a.B + (int64_t)(c0 + cc) * a.ldb;
        if (pre)
          for (int i = tid; i < c; i += 256) L.c[cc][i] = L.own[S.first - col_lo + i];
        else
          for (int i = tid; i < c; i += 256) L.c[cc][i] = bc[a.perm[S.first + i]];
        for (int i = tid; i < R; i += 256) L.x[cc][i] = 0.0;
      }
      __syncthreads();
      for (int q = 0; q < S.nchild; ++q) {
        const VecNode C = uniform_node(a.vn[uni(a.children[S.child_off + q])]);
        const int* rm = a.relmap + C.relmap_off;
        for (int cc = 0; cc < kb; ++cc) {
          const double* fc = b.fv + C.fv_off + (int64_t)(c0 + cc) * b.fvs + C.ncols;
          for (int t = tid; t < C.R; t += 256) {
            const int p = rm[t];
            const double v = fc[t];
            if (p < c) L.c[cc][p] += v;
            else L.x[cc][p - c] += v;
          }
        }
        __syncthreads();
      }
      // y_i = sum_{j <= i} Lt[j, i] fv_c[j]: wave w takes columns 8g..8g+7
      // together (lane: rows lane, lane + 64, ...; a column's order and
      // wave reduction are those of one column at a time)
      for (int i0 = wave * kSubCols; i0 < c; i0 += 4 * kSubCols) {
        const int nu = min(kSubCols, c - i0);
        double acc[kSubCols][KS];
#pragma unroll
        for (int u = 0; u < kSubCols; ++u)
#pragma unroll
          for (int cc = 0; cc < KS; ++cc) acc[u][cc] = 0;
        for (int jj = lane; jj < i0 + nu; jj += 64) {
          double cv[KS];
#pragma unroll
          for (int cc = 0; cc < KS; ++cc) cv[cc] = L.c[cc][jj];
          T m[kSubCols];
#pragma unroll
          for (int u = 0; u < kSubCols; ++u) {
            const bool in = u < nu && jj <= i0 + u;
            if constexpr (sizeof(T) == 4) m[u] = Lt[jj + (int64_t)min(i0 + u, c - 1) * ld];
            else m[u] = in ? Lt[jj + (int64_t)(i0 + u) * ld] : T(0);
          }
#pragma unroll
          for (int u = 0; u < kSubCols; ++u) {
            const double mv = (u < nu && jj <= i0 + u) ? (double)m[u] : 0.0;
#pragma unroll
            for (int cc = 0; cc < KS; ++cc) acc[u][cc] += mv * cv[cc];
          }
        }
#pragma unroll
        for (int u = 0; u < kSubCols; ++u)
#pragma unroll
          for (int cc = 0; cc < KS; ++cc) {
            const double v = wave_sum(acc[u][cc]);
            if (lane == 0 && u < nu) L.y[cc][i0 + u] = v;
          }
      }
      // fv_R -= Zp fv_c   (one thread per row), stored once; the subtree
      // root's fv_R leaves the workgroup (device-scope stores)
      for (int r = tid; r < R; r += 256) {
        const T* rp = Lt + c + r;
        double acc[KS];
#pragma unroll
        for (int cc = 0; cc < KS; ++cc) acc[cc] = 0;
        // eight loads in flight, then their products in column order
        int jj = 0;
        for (; jj + 8 <= c; jj += 8) {
          T m[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) m[u] = rp[(int64_t)(jj + u) * ld];
#pragma unroll
          for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int cc = 0; cc < KS; ++cc) acc[cc] += (double)m[u] * L.c[cc][jj + u];
        }
        for (; jj < c; ++jj) {
          const double m = (double)rp[(int64_t)jj * ld];
#pragma unroll
          for (int cc = 0; cc < KS; ++cc) acc[cc] += m * L.c[cc][jj];
        }
        for (int cc = 0; cc < kb; ++cc) {
          double* fvR = b.fv + S.fv_off + (int64_t)(c0 + cc) * b.fvs + c;
          const double v = L.x[cc][r] - acc[cc];
          if (COH && j == rootn) vst<true>(fvR + r, v);
          else fvR[r] = v;
        }
      }
      __syncthreads();
      // z_i = sum_{j >= i} Lt[i, j] y_j   (one thread per row)
      for (int i = tid; i < c; i += 256) {
        const T* rp = Lt + i;
        double acc[KS];
#pragma unroll
        for (int cc = 0; cc < KS; ++cc) acc[cc] = 0;
        int jj = i;
        for (; jj + 8 <= c; jj += 8) {
          T m[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) m[u] = rp[(int64_t)(jj + u) * ld];
#pragma unroll
          for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int cc = 0; cc < KS; ++cc) acc[cc] += (double)m[u] * L.y[cc][jj + u];
        }
        for (; jj < c; ++jj) {
          const double m = (double)rp[(int64_t)jj * ld];
#pragma unroll
          for (int cc = 0; cc < KS; ++cc) acc[cc] += m * L.y[cc][jj];
        }
        for (int cc = 0; cc < kb; ++cc)
          vst<COH>(b.z + (int64_t)(c0 + cc) * b.n + S.first + i, acc[cc]);
      }
      __syncthreads();
    }
  }
}

template <typename T, bool COH, int KS>
__device__ void subtree_backward_g(const T* __restrict__ F, const SubtreeArgs& a, int sb,
                                   const VecBufs& b, int k, SubtreeLdsK<KS>& L) {
  const int tid = body_tid(), lane = tid & 63, wave = tid >> 6;
  const int q0 = uni(a.ptr[sb]), q1 = uni(a.ptr[sb + 1]);
  // one right-hand side: z of every subtree column at once
  const int col_lo = uni(a.cols[2 * sb]), col_n = uni(a.cols[2 * sb + 1]);
  const bool pre = KS == 1 && col_n >= 0 && col_n <= kSubtreeOwn;
  if (pre) {
    for (int i = tid; i < col_n; i += 256) L.own[i] = vld<COH>(b.z + col_lo + i);
    __syncthreads();
  }
  for (int qi = q1 - 1; qi >= q0; --qi) {
    const int j = uni(a.nodes[qi]);
    const VecNode S = uniform_node(a.vn[j]);
    const int64_t P = uni64(a.panel[j]);
    const int c = S.ncols, R = S.R, ld = c + R;
    const T* __restrict__ Zp = F + P + c;
    const int* rw = a.rows + S.rows_off;
    for (int c0 = 0; c0 < k; c0 += KS) {
      const int kb = min(KS, k - c0);
      for (int cc = 0; cc < kb; ++cc) {
        const double* wc = b.w + (int64_t)(c0 + cc) * b.n;
        for (int t = tid; t < R; t += 256) L.x[cc][t] = vld<COH>(wc + rw[t]);
      }
      __syncthreads();
      // x_i = z_i - sum_r Zp[r, i] x_R[r]: wave w takes columns 8g..8g+7
      // together (lane: rows lane, lane + 64, ...; a column's order and wave
      // reduction are those of one column at a time)
      for (int i0 = wave * kSubCols; i0 < c; i0 += 4 * kSubCols) {
        const int nu = min(kSubCols, c - i0);
        double acc[kSubCols][KS];
#pragma unroll
        for (int u = 0; u < kSubCols; ++u)
#pragma unroll
          for (int cc = 0; cc < KS; ++cc) acc[u][cc] = 0;
        for (int r = lane; r < R; r += 64) {
          double xv[KS];
#pragma unroll
          for (int cc = 0; cc < KS; ++cc) xv[cc] = L.x[cc][r];
          T m[kSubCols];
#pragma unroll
          for (int u = 0; u < kSubCols; ++u) {
            if constexpr (sizeof(T) == 4) m[u] = Zp[r + (int64_t)min(i0 + u, c - 1) * ld];
            else m[u] = u < nu ? Zp[r + (int64_t)(i0 + u) * ld] : T(0);
          }
#pragma unroll
          for (int u = 0; u < kSubCols; ++u) {
            const double mv = u < nu ? (double)m[u] : 0.0;
#pragma unroll
            for (int cc = 0; cc < KS; ++cc) acc[u][cc] += mv * xv[cc];
          }
        }
#pragma unroll
        for (int u = 0; u < kSubCols; ++u)
#pragma unroll
          for (int cc = 0; cc < KS; ++cc) {
            const double v = wave_sum(acc[u][cc]);
            if (lane == 0 && u < nu && cc < kb) {
              const int64_t o = (int64_t)(c0 + cc) * b.n + S.first + i0 + u;
              const double zv = pre ? L.own[S.first + i0 + u - col_lo] : vld<COH>(b.z + o);
              vst<COH>(b.w + o, zv - v);
            }
          }
      }
      __syncthreads();
    }
  }
}

template <typename T, bool COH>
__device__ void subtree_forward(const T* __restrict__ F, const SubtreeArgs& a, int sb,
                                const VecBufs& b, int k, SubtreeLds& L) {
  if (k == 1) subtree_forward_g<T, COH, 1>(F, a, sb, b, k, reinterpret_cast<SubtreeLdsK<1>&>(L));
  else subtree_forward_g<T, COH, kSubtreeRhs>(F, a, sb, b, k, L);
}

template <typename T, bool COH>
__device__ void subtree_backward(const T* __restrict__ F, const SubtreeArgs& a, int sb,
                                 const VecBufs& b, int k, SubtreeLds& L) {
  if (k == 1) subtree_backward_g<T, COH, 1>(F, a, sb, b, k, reinterpret_cast<SubtreeLdsK<1>&>(L));
  else subtree_backward_g<T, COH, kSubtreeRhs>(F, a, sb, b, k, L);
}

template <typename T, bool FWD>
__global__ __launch_bounds__(256) void k_subtrees(const T* __restrict__ F, SubtreeArgs a,
                                                  VecBufs b, int k) {
  __shared__ SubtreeLds L;
  if (b.live && uni(*b.live) == 0) return;
  if (FWD) subtree_forward<T, false>(F, a, blockIdx.x, b, k, L);
  else subtree_backward<T, false>(F, a, blockIdx.x, b, k, L);
}

template <typename T>
void launch_subtrees(const T* factor, const SubtreeArgs& a, int count, bool forward,
                     VecBufs b, int k, hipStream_t s) {
  if (count <= 0) return;
  if (forward) k_subtrees<T, true><<<count, 256, 0, s>>>(factor, a, b, k);
  else k_subtrees<T, false><<<count, 256, 0, s>>>(factor, a, b, k);
}
template void launch_subtrees<double>(const double*, const SubtreeArgs&, int, bool, VecBufs,
                                      int, hipStream_t);
template void launch_subtrees<float>(const float*, const SubtreeArgs&, int, bool, VecBufs, int,
                                     hipStream_t);

// ---- multi-right-hand-side subtree tasks -------------------------------------
// One workgroup runs a subtree of small uncompressed supernodes for sixteen
// right-hand sides per pass.  A supernode's whole front vector (ncols + R
// entries) and y (ncols) sit in LDS; every product runs on the FP64 matrix
// cores (sixteen right-hand sides per MFMA 16x16x4), the matrix loaded once
// per pass.  Per supernode (panel [Lc^{-T}; Zp]):
//   forward:  fv_c = b[perm], fv_R = 0 + children's fv_R (written by this
//             workgroup);  y = Lc^{-1} fv_c;  fv_R -= Zp fv_c;  z = Lc^{-T} y
//   backward: x_R = w[rows];  x_c = z - Zp^T x_R
// The stored Lc^{-T} block is exactly zero below its diagonal, so the
// products run over the full ncols x ncols block.  z, w and the subtree
// root's fv_R are written with device-scope stores (other workgroups read
// them); the other supernodes' fv_R stay with this workgroup.
constexpr int kMsFld = kMsMaxFront + 1;  // front row stride in LDS (odd: banks)
constexpr int kMsYld = kMsMaxCols + 1;
constexpr int kMsLds = kMfmaRhs * kMsFld + kMfmaRhs * kMsYld;  // doubles
// Larger supernodes (up to kMlMaxCols columns, any R): fv_c and y in LDS
// ([16][kMlFld] each), fv_R in device memory; the backward product stages
// x_R through LDS in chunks of kMlChunk rows.
constexpr int kMlFld = kMlMaxCols + 1;
constexpr int kMlChunk = 256;
static_assert(2 * kMfmaRhs * kMlFld <= kMsLds, "fv_c and y of a larger supernode fit");
static_assert(kMfmaRhs * (kMlChunk + 1) <= kMsLds, "an x_R chunk fits");
static_assert(kMlMaxCols <= 128, "at most two 16-row blocks per wave");

// out(i, r, v) for i < rows, r < kb: v = sum_j A(i, j) x(r, j), j < cols,
// with A(i, j) = M[i + j ld] (TRANS false) or M[j + i ld] (TRANS true) and
// x in LDS (right-hand side r at xs + r * xld).  Wave w takes the 16-row
// blocks w, w + 4, ..., two at a time; columns go in chunks of 32 whose
// sixteen loads per lane are issued together, the sum over j in groups of
// four in increasing order for every right-hand side (so a column's bits do
// not depend on kb).
//
// The four old values a lane's epilogue needs (get(row, r), e.g. from
// device memory) are read before the pass's matrix loads, and every result
// (put(row, r, old, v)) is written after them, so a read-modify-write
// epilogue adds no round trip of its own.
template <typename T, bool TRANS, typename Get, typename Put>
__device__ void ms_product(const T* __restrict__ M, int ld, int rows, int cols,
                           const double* xs, int xld, int kb, Get get, Put put) {
  const int tid0 = body_tid(), lane = tid0 & 63, wave = tid0 >> 6;
  const int li = lane & 15, kq = lane >> 4;
  // two 16-row blocks per wave and pass (i0 and i0 + 64), columns in chunks
  // of 32: the sixteen matrix loads of a chunk go out together
  for (int i0 = wave * 16; i0 < rows; i0 += 128) {
    const bool two = i0 + 64 < rows;
    // the old values go out first: they arrive with the matrix loads
    double old[2][4];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = i0 + 64 * h + kq + 4 * v;
        old[h][v] = (row < rows && li < kb) ? get(row, li) : 0.0;
      }
    d4v acc[2];
    acc[0] = d4v{0, 0, 0, 0};
    acc[1] = d4v{0, 0, 0, 0};
    for (int j0 = 0; j0 < cols; j0 += 32) {
      T m[2][8];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int i = i0 + 64 * h + li;
        const bool iok = i < rows;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int j = j0 + 4 * u + kq;
          const bool ok = iok && j < cols;
          const int ii = ok ? i : 0, jj = ok ? j : 0;
          const T v = TRANS ? M[jj + (int64_t)ii * ld] : M[ii + (int64_t)jj * ld];
          m[h][u] = ok ? v : T(0);
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (j0 + 4 * u >= cols) break;
        const int j = j0 + 4 * u + kq;
        const double xv = (j < cols && li < kb) ? xs[li * xld + j] : 0.0;
        acc[0] = mfma64((double)m[0][u], xv, acc[0]);
        if (two) acc[1] = mfma64((double)m[1][u], xv, acc[1]);
      }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = i0 + 64 * h + kq + 4 * v;
        if (row < rows && li < kb) put(row, li, old[h][v], acc[h][v]);
      }
  }
}

// The epilogue of a product whose results are written without reading.
__device__ inline double ms_none(int, int) { return 0.0; }

// Element loop of the subtree bodies over items [0, total) on the
// workgroup's 256 threads, NB items per thread per round: every index load
// of a round (index(i)), then every value load (load(i, ix)), then every
// use (use(i, ix, v)), so a round costs one memory round trip per stage
// instead of one per item.
template <int NB, typename Index, typename Load, typename Use>
__device__ inline void ms_batched(int total, Index index, Load load, Use use) {
  for (int base = body_tid(); base < total; base += 256 * NB) {
    int ix[NB];
    double v[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int i = base + 256 * u;
      ix[u] = i < total ? index(i) : 0;
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int i = base + 256 * u;
      v[u] = i < total ? load(i, ix[u]) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int i = base + 256 * u;
      if (i < total) use(i, ix[u], v[u]);
    }
  }
}

// A processed supernode of a multi-right-hand-side subtree task as its
// parent needs it.  A task runs its supernodes in postorder, so when a
// supernode starts, its children are the top entries of a stack of these,
// in increasing order (checked on the host when the tasks are built).
struct MsChild {
  int64_t fvr;  // its update vector (fv_off + ncols), first right-hand side
  int64_t rm;   // its relative map
  int R, pad;
};
// Descriptors of up to kMsRing consecutive supernodes of a task, loaded
// together (two dependent memory round trips per kMsRing supernodes instead
// of two per supernode), with the original indices of the columns of those
// of at most kMsMaxCols columns.
constexpr int kMsRing = 8;
struct MsNodeDesc {
  int64_t fv_off, rows_off, relmap_off, panel;
  int first, ncols, R, nchild, root, pad;
};
struct MsRing {
  MsNodeDesc d[kMsRing];
  int perm[kMsRing][kMsMaxCols];
};

// Every thread of the workgroup: descriptors of the task's supernodes at
// list positions q .. q + nr into the ring (lane l reads position q + l),
// then the original indices of the small ones' columns (with_perm).
__device__ void ms_ring_load(const SubtreeArgs& a, int q, int nr, bool with_perm, MsRing& rg) {
  const int tid = body_tid();
  __syncthreads();  // earlier readers of the ring are done
  if (tid < nr) {
    const int jr = a.nodes[q + tid];
    const bool root = jr < 0;
    const int j = root ? ~jr : jr;
    const VecNode v = a.vn[j];
    MsNodeDesc& d = rg.d[tid];
    d.fv_off = v.fv_off;
    d.rows_off = v.rows_off;
    d.relmap_off = v.relmap_off;
    d.panel = a.panel[j];
    d.first = v.first;
    d.ncols = v.ncols;
    d.R = v.R;
    d.nchild = v.nchild;
    d.root = root ? 1 : 0;
  }
  __syncthreads();
  if (with_perm) {
    for (int idx = tid; idx < nr * kMsMaxCols; idx += 256) {
      const int l = idx / kMsMaxCols, i = idx - l * kMsMaxCols;
      const int c = rg.d[l].ncols;
      if (i < c && c <= kMsMaxCols) rg.perm[l][i] = a.perm[rg.d[l].first + i];
    }
    __syncthreads();
  }
}

// Forward step of a larger supernode of a multi subtree (fv_R in device
// memory, plain accesses except the subtree root's final values).
template <typename T>
__device__ void ms_forward_large(const T* __restrict__ Lt, const VecNode& S,
                                 const SubtreeArgs& a, const VecBufs& b, int c0, int kb,
                                 bool root, const MsChild* ch, double* lds) {
  const int tid = body_tid();
  const int c = S.ncols, R = S.R, ld = c + R;
  double* Fc = lds;                      // [16][kMlFld] fv_c
  double* Yl = lds + kMfmaRhs * kMlFld;  // [16][kMlFld] y
  double* fvR = b.fv + S.fv_off + c + (int64_t)c0 * b.fvs;
  for (int idx = tid; idx < kMfmaRhs * c; idx += 256) {
    const int r = idx / c, p = idx - r * c;
    if (r >= kb) Fc[r * kMlFld + p] = 0.0;
  }
  ms_batched<4>(
      kb * c, [&](int i) { return a.perm[S.first + i % c]; },
      [&](int i, int o) { return a.B[o + (int64_t)(c0 + i / c) * a.ldb]; },
      [&](int i, int, double v) { Fc[(i / c) * kMlFld + i % c] = v; });
  for (int idx = tid; idx < kb * R; idx += 256) {
    const int r = idx / R, t = idx - r * R;
    fvR[(int64_t)r * b.fvs + t] = 0.0;
  }
  __syncthreads();
  for (int q = 0; q < S.nchild; ++q) {
    const int* rm = a.relmap + uni64(ch[q].rm);
    const double* fc = b.fv + uni64(ch[q].fvr) + (int64_t)c0 * b.fvs;
    const int CR = uni(ch[q].R);
    // a child's entries land on distinct positions: every load of a round
    // (its values, and the old fv_R entries they are added to) goes first;
    // the first child adds to the zeros stored above, so it stores its
    // values without loading them
    const bool fresh = q == 0;
    for (int base = tid; base < kb * CR; base += 256 * 4) {
      int pp[4];
      double v[4], o[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = base + 256 * u;
        pp[u] = i < kb * CR ? rm[i % CR] : 0;
        v[u] = i < kb * CR ? fc[(int64_t)(i / CR) * b.fvs + i % CR] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = base + 256 * u;
        o[u] = (!fresh && i < kb * CR && pp[u] >= c) ? fvR[(int64_t)(i / CR) * b.fvs + pp[u] - c]
                                                     : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = base + 256 * u;
        if (i >= kb * CR) continue;
        const int r = i / CR;
        if (pp[u] < c) Fc[r * kMlFld + pp[u]] += v[u];
        else fvR[(int64_t)r * b.fvs + pp[u] - c] = o[u] + v[u];
      }
    }
    __syncthreads();
  }
  ms_product<T, true>(Lt, ld, c, c, Fc, kMlFld, kb, ms_none,
                      [&](int i, int r, double, double v) { Yl[r * kMlFld + i] = v; });
  ms_product<T, false>(
      Lt + c, ld, R, c, Fc, kMlFld, kb,
      [&](int i, int r) { return fvR[(int64_t)r * b.fvs + i]; },
      [&](int i, int r, double old, double v) {
        double* p = fvR + (int64_t)r * b.fvs + i;
        if (root) vst<true>(p, old - v);
        else *p = old - v;
      });
  __syncthreads();
  double* zb = b.z + (int64_t)c0 * b.n + S.first;
  ms_product<T, false>(Lt, ld, c, c, Yl, kMlFld, kb, ms_none,
                       [&](int i, int r, double, double v) {
                         vst<true>(zb + (int64_t)r * b.n + i, v);
                       });
  __syncthreads();
}

// Backward step of a larger supernode: x_c = z_c - Zp^T x_R with x_R staged
// through LDS in chunks of kMlChunk rows; each wave keeps the sums of its
// (at most two) 16-row blocks of x_c across the chunks, the sum running over
// the rows in increasing order.
template <typename T>
__device__ void ms_backward_large(const T* __restrict__ Zp, const VecNode& S,
                                  const SubtreeArgs& a, const VecBufs& b, int c0, int kb,
                                  double* lds) {
  const int tid = body_tid(), lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, kq = lane >> 4;
  const int c = S.ncols, R = S.R, ld = c + R;
  constexpr int XL = kMlChunk + 1;
  const int* rw = a.rows + S.rows_off;
  const double* wb = b.w + (int64_t)c0 * b.n;
  d4v acc[2];
  acc[0] = d4v{0, 0, 0, 0};
  acc[1] = d4v{0, 0, 0, 0};
  for (int r0 = 0; r0 < R; r0 += kMlChunk) {
    const int nr = min(kMlChunk, R - r0);
    ms_batched<8>(
        kb * nr, [&](int i) { return rw[r0 + i % nr]; },
        [&](int i, int row) { return vld<true>(wb + (int64_t)(i / nr) * b.n + row); },
        [&](int i, int, double v) { lds[(i / nr) * XL + i % nr] = v; });
    __syncthreads();
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) {
      const int i0 = wave * 16 + 64 * bi;
      if (i0 >= c) break;
      const int i = i0 + li;
      const bool iok = i < c;
      for (int j0 = 0; j0 < nr; j0 += 16) {
        T m[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int j = j0 + 4 * u + kq;
          const bool ok = iok && j < nr;
          const T v = Zp[(int64_t)(ok ? r0 + j : 0) + (int64_t)(ok ? i : 0) * ld];
          m[u] = ok ? v : T(0);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (j0 + 4 * u >= nr) break;
          const int j = j0 + 4 * u + kq;
          const double xv = (j < nr && li < kb) ? lds[li * XL + j] : 0.0;
          acc[bi] = mfma64((double)m[u], xv, acc[bi]);
        }
      }
    }
    __syncthreads();
  }
  const double* zb = b.z + (int64_t)c0 * b.n + S.first;
  double* xb = b.w + (int64_t)c0 * b.n + S.first;
#pragma unroll
  for (int bi = 0; bi < 2; ++bi) {
    const int i0 = wave * 16 + 64 * bi;
    if (i0 >= c) break;
    double zo[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = i0 + kq + 4 * v;
      zo[v] = (row < c && li < kb) ? vld<true>(zb + (int64_t)li * b.n + row) : 0.0;
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = i0 + kq + 4 * v;
      if (row < c && li < kb) vst<true>(xb + (int64_t)li * b.n + row, zo[v] - acc[bi][v]);
    }
  }
  __syncthreads();
}

template <typename T>
__device__ void ms_forward(const T* __restrict__ F, const SubtreeArgs& a, int sb,
                           const VecBufs& b, int k, double* lds) {
  __shared__ MsChild s_stk[kMsStack];
  __shared__ MsRing s_ring;
  double* Fr = lds;                       // [16][kMsFld] front vector
  double* Yr = lds + kMfmaRhs * kMsFld;   // [16][kMsYld] y
  const int q0 = uni(a.ptr[sb]), q1 = uni(a.ptr[sb + 1]);
  for (int c0 = 0; c0 < k; c0 += kMfmaRhs) {
    const int kb = min(kMfmaRhs, k - c0);
    int sp = 0;  // stack depth (the same in every thread)
    // the right-hand-side entries of the next supernode, loaded during the
    // step of the one before it when both are small and the next
    // descriptor is in the ring
    bool pre = false;
    double nbv[2] = {0.0, 0.0};
    for (int qi = q0; qi < q1; ++qi) {
      if ((qi - q0) % kMsRing == 0) ms_ring_load(a, qi, min(kMsRing, q1 - qi), true, s_ring);
      const int tid = body_tid();  // per supernode (not hoisted out of the loop)
      const int l = (qi - q0) % kMsRing;
      const MsNodeDesc& D = s_ring.d[l];
      VecNode S{};
      S.first = uni(D.first);
      S.ncols = uni(D.ncols);
      S.R = uni(D.R);
      S.nchild = uni(D.nchild);
      S.fv_off = uni64(D.fv_off);
      S.rows_off = uni64(D.rows_off);
      S.relmap_off = uni64(D.relmap_off);
      const bool root = uni(D.root) != 0;  // its update vector goes to other workgroups
      const int64_t P = uni64(D.panel);
      const int c = S.ncols, R = S.R, ld = c + R;
      const T* __restrict__ Lt = F + P;
      const MsChild* ch = s_stk + sp - S.nchild;  // the children, in order
      if (!(c <= kMsMaxCols && ld <= kMsMaxFront)) {
        ms_forward_large<T>(Lt, S, a, b, c0, kb, root, ch, lds);
        pre = false;
      } else {
        for (int idx = tid; idx < kMfmaRhs * ld; idx += 256) {
          const int r = idx / ld, p = idx - r * ld;
          if (r >= kb || p >= c) Fr[r * kMsFld + p] = 0.0;
        }
        // fv_c = b[perm[cols]] (original indices staged with the
        // descriptors), loaded during the previous supernode's step or now
        {
          const int* pm = s_ring.perm[l];
          double v[2];
          int ii[2];
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            const int i = tid + 256 * u;
            ii[u] = i;
            v[u] = pre ? nbv[u]
                       : (i < kb * c ? a.B[pm[i % c] + (int64_t)(c0 + i / c) * a.ldb] : 0.0);
          }
#pragma unroll
          for (int u = 0; u < 2; ++u)
            if (ii[u] < kb * c) Fr[(ii[u] / c) * kMsFld + ii[u] % c] = v[u];
        }
        __syncthreads();
        pre = false;
        if (l + 1 < kMsRing && qi + 1 < q1) {
          const int cn = uni(s_ring.d[l + 1].ncols), Rn = uni(s_ring.d[l + 1].R);
          if (cn <= kMsMaxCols && cn + Rn <= kMsMaxFront) {
            pre = true;
            const int* pn = s_ring.perm[l + 1];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
              const int i = tid + 256 * u;
              nbv[u] = i < kb * cn ? a.B[pn[i % cn] + (int64_t)(c0 + i / cn) * a.ldb] : 0.0;
            }
          }
        }
        int q = 0;
        if (S.nchild >= 2 && kb * uni(ch[0].R) <= 1024 && kb * uni(ch[1].R) <= 1024) {
          // two children of one round each: both children's loads first,
          // then their adds in order
          const int* rm0 = a.relmap + uni64(ch[0].rm);
          const int* rm1 = a.relmap + uni64(ch[1].rm);
          const double* fc0 = b.fv + uni64(ch[0].fvr);
          const double* fc1 = b.fv + uni64(ch[1].fvr);
          const int CR0 = uni(ch[0].R), CR1 = uni(ch[1].R);
          int p0[4], p1[4];
          double v0[4], v1[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int i = tid + 256 * u;
            const bool k0 = i < kb * CR0, k1 = i < kb * CR1;
            p0[u] = k0 ? rm0[i % CR0] : 0;
            v0[u] = k0 ? fc0[(int64_t)(c0 + i / CR0) * b.fvs + i % CR0] : 0.0;
            p1[u] = k1 ? rm1[i % CR1] : 0;
            v1[u] = k1 ? fc1[(int64_t)(c0 + i / CR1) * b.fvs + i % CR1] : 0.0;
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int i = tid + 256 * u;
            if (i < kb * CR0) Fr[(i / CR0) * kMsFld + p0[u]] += v0[u];
          }
          __syncthreads();
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int i = tid + 256 * u;
            if (i < kb * CR1) Fr[(i / CR1) * kMsFld + p1[u]] += v1[u];
          }
          __syncthreads();
          q = 2;
        }
        for (; q < S.nchild; ++q) {
          const int* rm = a.relmap + uni64(ch[q].rm);
          const double* fc = b.fv + uni64(ch[q].fvr);
          const int CR = uni(ch[q].R);
          ms_batched<4>(
              kb * CR, [&](int i) { return rm[i % CR]; },
              [&](int i, int) { return fc[(int64_t)(c0 + i / CR) * b.fvs + i % CR]; },
              [&](int i, int p, double v) { Fr[(i / CR) * kMsFld + p] += v; });
          __syncthreads();
        }
        // y = Lc^{-1} fv_c (y_i = sum_j Lt[j, i] fv_c[j]) and fv_R -= Zp fv_c
        ms_product<T, true>(Lt, ld, c, c, Fr, kMsFld, kb, ms_none,
                            [&](int i, int r, double, double v) { Yr[r * kMsYld + i] = v; });
        ms_product<T, false>(Lt + c, ld, R, c, Fr, kMsFld, kb, ms_none,
                             [&](int i, int r, double, double v) { Fr[r * kMsFld + c + i] -= v; });
        __syncthreads();
        // z = Lc^{-T} y
        double* zb = b.z + (int64_t)c0 * b.n + S.first;
        ms_product<T, false>(Lt, ld, c, c, Yr, kMsYld, kb, ms_none,
                             [&](int i, int r, double, double v) {
                               vst<true>(zb + (int64_t)r * b.n + i, v);
                             });
        double* fvR = b.fv + S.fv_off + c + (int64_t)c0 * b.fvs;
        for (int idx = tid; idx < kb * R; idx += 256) {
          const int r = idx / R, t = idx - r * R;
          const double v = Fr[r * kMsFld + c + t];
          if (root) vst<true>(fvR + (int64_t)r * b.fvs + t, v);
          else fvR[(int64_t)r * b.fvs + t] = v;
        }
      }
      // the children are replaced by this node on the stack (its entries
      // are read after the barrier below, by its parent)
      sp -= S.nchild;
      if (tid == 0) s_stk[sp] = MsChild{S.fv_off + c, S.relmap_off, R, 0};
      ++sp;
      __syncthreads();
    }
  }
}

template <typename T>
__device__ void ms_backward(const T* __restrict__ F, const SubtreeArgs& a, int sb,
                            const VecBufs& b, int k, double* lds) {
  __shared__ MsRing s_ring;
  __shared__ int s_rw[kMsMaxFront];  // row indices of the current small supernode
  double* Fr = lds;  // [16][kMsFld]: x_R at positions c .. c + R
  const int q0 = uni(a.ptr[sb]), q1 = uni(a.ptr[sb + 1]);
  for (int c0 = 0; c0 < k; c0 += kMfmaRhs) {
    const int kb = min(kMfmaRhs, k - c0);
    bool pre = false;  // s_rw holds supernode qi's rows (loaded during the step before)
    int rbase = q1;  // list position of the ring's first descriptor
    for (int qi = q1 - 1; qi >= q0; --qi) {
      if (qi < rbase) {
        // descriptors of the next (at most) kMsRing supernodes going down
        const int nr = min(kMsRing, qi - q0 + 1);
        rbase = qi - nr + 1;
        ms_ring_load(a, rbase, nr, false, s_ring);
      }
      const int tid = body_tid();  // per supernode (not hoisted out of the loop)
      const MsNodeDesc& D = s_ring.d[qi - rbase];
      VecNode S{};
      S.first = uni(D.first);
      S.ncols = uni(D.ncols);
      S.R = uni(D.R);
      S.nchild = uni(D.nchild);
      S.fv_off = uni64(D.fv_off);
      S.rows_off = uni64(D.rows_off);
      S.relmap_off = uni64(D.relmap_off);
      const int64_t P = uni64(D.panel);
      const int c = S.ncols, R = S.R, ld = c + R;
      if (!(c <= kMsMaxCols && ld <= kMsMaxFront)) {
        ms_backward_large<T>(F + P + c, S, a, b, c0, kb, lds);
        pre = false;
        continue;
      }
      if (!pre) {
        const int* rw = a.rows + S.rows_off;
        for (int t = tid; t < R; t += 256) s_rw[t] = rw[t];
        __syncthreads();
      }
      // the next supernode's rows (going down), when it is small and its
      // descriptor is in the ring: loaded now, staged after the gather
      bool pre_next = false;
      int nrw = 0, Rn = 0;
      if (qi - 1 >= rbase) {
        const MsNodeDesc& Dn = s_ring.d[qi - 1 - rbase];
        const int cn = uni(Dn.ncols);
        Rn = uni(Dn.R);
        if (cn <= kMsMaxCols && cn + Rn <= kMsMaxFront) {
          pre_next = true;
          if (tid < Rn) nrw = a.rows[uni64(Dn.rows_off) + tid];
        }
      }
      const double* wb = b.w + (int64_t)c0 * b.n;
      ms_batched<4>(
          kb * R, [&](int i) { return s_rw[i % R]; },
          [&](int i, int row) { return vld<true>(wb + (int64_t)(i / R) * b.n + row); },
          [&](int i, int, double v) { Fr[(i / R) * kMsFld + c + i % R] = v; });
      __syncthreads();
      // x_c = z_c - Zp^T x_R
      const double* zb = b.z + (int64_t)c0 * b.n + S.first;
      double* xb = b.w + (int64_t)c0 * b.n + S.first;
      ms_product<T, true>(
          F + P + c, ld, c, R, Fr + c, kMsFld, kb,
          [&](int i, int r) { return vld<true>(zb + (int64_t)r * b.n + i); },
          [&](int i, int r, double z, double v) { vst<true>(xb + (int64_t)r * b.n + i, z - v); });
      // s_rw was last read by the gather, before the barrier above
      if (pre_next && tid < Rn) s_rw[tid] = nrw;
      pre = pre_next;
      __syncthreads();
    }
  }
}

constexpr int DF_SCATTER_CHUNK = kScatterChunk;  // x entries per scatter task (8 per thread)
static_assert(DF_SCATTER_CHUNK % (8 * 256) == 0, "scatter chunk: eight entries per thread per pass");

// Thread 0 waits until group g has completed need tasks.  Returns false when
// the wait gave up (timeout here or in another workgroup): the kernel then
// runs to its end without further waiting and the host reports the failure.
__device__ bool df_wait(const DfArgs& a, int g, unsigned need) {
  if (__hip_atomic_load(a.cnt + g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= need)
    return true;
  const unsigned long long t0 = wall_clock64();
  int backoff = 1;
  while (__hip_atomic_load(a.cnt + g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
    if (__hip_atomic_load(a.status + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      return false;
    if (wall_clock64() - t0 > a.timeout_ticks) {
      __hip_atomic_store(a.status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(a.status + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    for (int z = 0; z < backoff; ++z) __builtin_amdgcn_s_sleep(4);
    if (backoff < a.backoff_max) backoff *= 2;
  }
  return true;
}

// LDS of the two dataflow kernels.  The single-right-hand-side kernel runs
// subtree tasks (one right-hand side per pass) and scalar task bodies; the
// multi-right-hand-side kernel the matrix-core bodies and four-column
// assembly, and no subtree tasks (its list has none).  The smaller LDS of the
// first lets more of its workgroups share a compute unit.
struct DfLdsScalar {
  union {
    SubtreeLdsK<1> sl;
    double xl[kXlds];  // also the one-right-hand-side assembly chunk (VCH)
  };
};
static_assert(kXlds >= kSingleChunkGroup * VCH, "the assembly span fits the x staging area");
static_assert(kXlds >= kColRowsMax, "a column task's x fits the staging area");
constexpr int kMultiLds = kMfmaLds > kAsmRhs * VCH ? kMfmaLds : kAsmRhs * VCH;
struct DfLdsMulti {
  double mf[kMultiLds > kMsLds ? kMultiLds : kMsLds];
};
template <bool MULTI>
struct DfLds {
  using type = DfLdsScalar;
};
template <>
struct DfLds<true> {
  using type = DfLdsMulti;
};

// The body of task `ticket` (every thread of the workgroup).
template <typename T, bool MULTI>
__device__ inline void df_run_task(const T* __restrict__ F, const DfArgs& a, const VecBufs& b,
                                   int k, int ticket, double (*red)[64],
                                   typename DfLds<MULTI>::type& lds) {
  const DfTask& tk = a.tasks[ticket];
  const GvTask gt = uniform_task(a.gv[ticket]);
  const int type = uni(tk.type), ta = uni(tk.a), tb = uni(tk.b);
  const int tid = body_tid();
  if (type == 0) {
    // the chunk is summed in LDS (one right-hand side: the first
    // kSingleChunkGroup * VCH entries of the x staging area), then stored
    // once; a task of the single-right-hand-side list spans
    // kSingleChunkGroup chunks
    if constexpr (MULTI) {
      assemble_chunk<true>(uniform_node(a.nodes[ta]), tb, a.nodes, a.children, a.relmap,
                           a.ranges, b, k, a.B, a.ldb, a.perm, lds.mf);
    } else {
      assemble_chunk<true, kSingleChunkGroup>(uniform_node(a.nodes[ta]), tb, a.nodes,
                                              a.children, a.relmap, a.ranges, b, 1, a.B,
                                              a.ldb, a.perm, lds.xl);
    }
  } else if (type == 1) {
    if constexpr (MULTI) {
      gather_chunk<true, MULTI>(uniform_node(a.nodes[ta]), tb, a.rows, b, k);
    } else {
      gather_chunk<true, false, kSingleChunkGroup>(uniform_node(a.nodes[ta]), tb, a.rows, b,
                                                   k);
    }
  } else if (type == 2) {
    if constexpr (MULTI) gemv_lite<T, true, MULTI>(F, gt, b, k, red, lds.mf);
    else gemv_lite<T, true, MULTI>(F, gt, b, k, red, lds.xl);
  } else if (type == 4 || type == 5) {
    if constexpr (MULTI) {
      if (type == 4) ms_forward<T>(F, a.sub, ta, b, k, lds.mf);
      else ms_backward<T>(F, a.sub, ta, b, k, lds.mf);
    } else if (type == 4) {
      subtree_forward_g<T, true, 1>(F, a.sub, ta, b, 1, lds.sl);
    } else {
      subtree_backward_g<T, true, 1>(F, a.sub, ta, b, 1, lds.sl);
    }
  } else {
    // x = P^T w over a chunk of x in its own order, x[j] = w[pinv[j]]: the
    // stores are coalesced (a scattered 8-byte store costs a memory line);
    // eight source indices of a thread, then per right-hand side their
    // eight entries, are loaded together
    const int n = (int)b.n;
    for (int h = 0; h < DF_SCATTER_CHUNK; h += 8 * 256) {
      const int j0 = ta * DF_SCATTER_CHUNK + h + tid;
      int src[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) src[u] = j0 + 256 * u < n ? a.pinv[j0 + 256 * u] : 0;
      for (int col = 0; col < k; ++col) {
        const double* wc = b.w + (int64_t)col * b.n;
        double* xc = a.X + (int64_t)col * a.ldx + j0;
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = vld<true>(wc + src[u]);
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (j0 + 256 * u < n) xc[256 * u] = v[u];
      }
    }
  }
}

// Queue slot values besides task indices: not yet written, and a slot whose
// task the workgroup that made it ready runs itself.
constexpr int kQEmpty = -1;
constexpr int kQSkip = -3;
constexpr int kQAbort = -4;  // returned when the launch is aborted

// Thread 0 spins until queue slot idx is written (or the launch is
// aborted).  Same bound and back-off as a dependency wait.
__device__ int df_pop_wait(const DfArgs& a, int idx) {
  int t = __hip_atomic_load(a.q_queue + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (t != kQEmpty) return t;
  const unsigned long long t0 = wall_clock64();
  int backoff = 1;
  for (;;) {
    t = __hip_atomic_load(a.q_queue + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (t != kQEmpty) return t;
    if (__hip_atomic_load(a.status + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      return kQAbort;
    if (wall_clock64() - t0 > a.timeout_ticks) {
      __hip_atomic_store(a.status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(a.status + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return kQAbort;
    }
    for (int z = 0; z < backoff; ++z) __builtin_amdgcn_s_sleep(4);
    if (backoff < a.backoff_max) backoff *= 2;
  }
}

// Groups made ready by one task's completion whose tasks the whole
// workgroup pushes (more are pushed by thread 0 alone).
constexpr int kQPush = 16;
// Ready-queue loop: pop a slot, wait until it is written, run the task,
// signal its groups.  A group whose last wait completes has its tasks
// pushed (slots reserved at the push counter, written by the workgroup);
// the first such group (successors go most critical first) gives its first
// task to this workgroup as its next task, and its slot holds kQSkip.
// Every slot is written exactly once and a task is popped or continued only
// when all its waits are complete, so no workgroup holds a task that cannot
// run.
template <typename T, bool MULTI>
__device__ void df_queue_loop(const T* __restrict__ F, const DfArgs& a, const VecBufs& b,
                              int k, double (*red)[64], typename DfLds<MULTI>::type& lds) {
  __shared__ int s_task, s_next, s_np;
  __shared__ int s_pb[kQPush], s_pf[kQPush], s_pc[kQPush], s_pn[kQPush];
  __shared__ unsigned long long s_tgrab, s_tready;
  const int tid = body_tid();
  if (tid == 0) s_next = -1;
  for (;;) {
    if (tid == 0) {
      if (a.trace) s_tgrab = wall_clock64();
      int t = s_next;
      s_next = -1;
      while (t < 0) {
        const int idx = (int)__hip_atomic_fetch_add(a.q_ctr, 1u, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
        if (idx >= a.ntasks) {
          t = -2;  // every slot popped: done
          break;
        }
        t = df_pop_wait(a, idx);
        if (t == kQAbort) break;
        if (t == kQSkip) t = -1;  // pop again
      }
      s_task = t;
      if (a.trace) s_tready = wall_clock64();
    }
    __syncthreads();
    const int ticket = uni(s_task);
    if (ticket < 0) break;
    // the completion step's groups and counts, read with the descriptor
    const DfTask& tk = a.tasks[ticket];
    const int g1 = uni(tk.group), g2 = uni(tk.group2);
    const int sp[4] = {uni(tk.succ1[0]), uni(tk.succ1[1]), uni(tk.succ2[0]), uni(tk.succ2[1])};
    const int sf[2] = {uni(tk.sfirst[0]), uni(tk.sfirst[1])};
    const unsigned n1 = (unsigned)uni((int)tk.need1), n2 = (unsigned)uni((int)tk.need2);
    df_run_task<T, MULTI>(F, a, b, k, ticket, red, lds);
    // every thread: its device-scope stores have completed before the count
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
    if (tid == 0) {
      // both completion counts go out together; only the task completing a
      // group does more (its successors' pending counts, then push slots
      // for those that became ready)
      // a task of a small group counts down the pending counts of the groups
      // waiting on it directly; in a larger group only the task completing
      // it does (both completion counts go out together); the task bringing
      // a pending count to zero pushes that group's tasks
      const bool d1 = n1 <= kQDirectMax, d2 = g2 >= 0 && n2 <= kQDirectMax;
      const unsigned o1 = d1 ? 0u
                             : __hip_atomic_fetch_add(a.cnt + g1, 1u, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT);
      const unsigned o2 = (g2 < 0 || d2) ? 0u
                                         : __hip_atomic_fetch_add(a.cnt + g2, 1u,
                                                                  __ATOMIC_RELAXED,
                                                                  __HIP_MEMORY_SCOPE_AGENT);
      int np = 0, next = -1;
      const int gs[2] = {(d1 || o1 + 1u == n1) ? g1 : -1,
                         (g2 >= 0 && (d2 || o2 + 1u == n2)) ? g2 : -1};
      for (int e = 0; e < 2; ++e) {
        const int g = gs[e];
        if (g < 0) continue;
        const int q0 = sp[2 * e], q1 = sp[2 * e + 1];
        for (int qb = q0; qb < q1; qb += 8) {
          // up to eight successors' pending counts at once (the first
          // successor comes with the descriptor), and with them the
          // successors' task records (range, count, first task)
          int hs[8], ps[8];
          int4 gr[8];
#pragma unroll
          for (int u = 0; u < 8; ++u)
            hs[u] = qb + u < q1 ? (qb + u == q0 ? sf[e] : a.q_succ[qb + u]) : -1;
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            ps[u] = hs[u] >= 0 ? __hip_atomic_fetch_add(a.q_pend + hs[u], -1, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT)
                               : 0;
            gr[u] = hs[u] >= 0 ? a.q_ginfo[hs[u]] : make_int4(0, 0, -1, 0);
          }
          for (int u = 0; u < 8; ++u) {
            if (hs[u] < 0 || ps[u] != 1) continue;
            const int4 gi = gr[u];
            const int c0 = gi.x, cn = gi.y;
            if (cn <= 0) continue;  // a group signalled by other groups' tasks
            const int base = (int)__hip_atomic_fetch_add(a.q_ctr + 1, (unsigned)cn,
                                                         __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT);
            int first = 0;
            if (next < 0) {
              next = gi.z;
              first = 1;
            }
            if (np < kQPush) {
              s_pb[np] = base;
              s_pf[np] = first;
              s_pc[np] = c0;
              s_pn[np] = cn;
              ++np;
            } else {
              for (int i = 0; i < cn; ++i)
                __hip_atomic_store(a.q_queue + base + i,
                                   i < first ? kQSkip : a.q_gtasks[c0 + i], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
          }
        }
      }
      s_np = np;
      s_next = next;
      if (a.trace) {
        unsigned long long* tr = a.trace + 4 * (int64_t)ticket;
        tr[0] = s_tgrab;
        tr[1] = s_tready;
        tr[2] = wall_clock64();
        tr[3] = blockIdx.x;
      }
    }
    __syncthreads();
    const int np = uni(s_np);
    for (int r = 0; r < np; ++r) {
      const int base = uni(s_pb[r]), first = uni(s_pf[r]);
      const int c0 = uni(s_pc[r]), cn = uni(s_pn[r]);
      for (int i = tid; i < cn; i += 256)
        __hip_atomic_store(a.q_queue + base + i, i < first ? kQSkip : a.q_gtasks[c0 + i],
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// One tid == 0 region per loop iteration, with a barrier on both sides: the
// completion signal and the next ticket are taken together, so the compiler
// cannot merge a tid == 0 region of one iteration with one of the next (a
// merged region lets the other lanes run the loop without thread 0).  A task's
// descriptors (task entry, its first wait inline, and the matrix-vector entry
// at the same index) depend only on the ticket, so they arrive together.
// MULTI = false is the single-right-hand-side kernel: k is 1 by
// construction (the launcher sends every k > 1 to the other variant), and
// the constant lets the compiler drop the multi-column code paths.
template <typename T, bool MULTI>
__global__ __launch_bounds__(256, 4) void k_apply_df(const T* __restrict__ F, DfArgs a,
                                                  VecBufs b, int k_arg) {
  const int k = MULTI ? k_arg : 1;
  __shared__ double red[4][64];
  __shared__ typename DfLds<MULTI>::type lds;
  __shared__ int s_ticket;
  __shared__ int s_ok;
  __shared__ unsigned long long s_tgrab, s_tready;
  const int tid = threadIdx.x;
  if (b.live && uni(*b.live) == 0) return;
  if (a.q_queue) {
    df_queue_loop<T, MULTI>(F, a, b, k, red, lds);
    return;
  }
  if (tid == 0) {
    s_ok = 1;
    s_ticket = (int)__hip_atomic_fetch_add(a.cnt + a.ngroups, 1u, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
    if (a.trace) s_tgrab = wall_clock64();
  }
  __syncthreads();
  for (;;) {
    const int ticket = uni(s_ticket);
    if (ticket >= a.ntasks) break;
    const DfTask& tk = a.tasks[ticket];
    const int group = uni(tk.group), group2 = uni(tk.group2);
    const int w0 = uni(tk.wbeg), w1 = uni(tk.wend);
    if (w1 > w0) {
      const int wg0 = uni(tk.wg0);
      const unsigned wn0 = (unsigned)uni((int)tk.wn0);
      if (tid == 0) {
        bool ok = s_ok != 0 && df_wait(a, wg0, wn0);
        for (int q = w0 + 1; ok && q < w1; ++q) {
          const DfWait w = a.gwait[q];
          ok = df_wait(a, w.group, w.need);
        }
        s_ok = ok ? 1 : 0;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      }
      __syncthreads();
    }
    if (a.trace && tid == 0) s_tready = wall_clock64();
    df_run_task<T, MULTI>(F, a, b, k, ticket, red, lds);
    // every thread: its device-scope stores have completed before the count
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
    if (tid == 0) {
      __hip_atomic_fetch_add(a.cnt + group, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (group2 >= 0)
        __hip_atomic_fetch_add(a.cnt + group2, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (a.trace) {
        unsigned long long* tr = a.trace + 4 * (int64_t)ticket;
        tr[0] = s_tgrab;
        tr[1] = s_tready;
        tr[2] = wall_clock64();
        tr[3] = blockIdx.x;
      }
      s_ticket = (int)__hip_atomic_fetch_add(a.cnt + a.ngroups, 1u, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
      if (a.trace) s_tgrab = wall_clock64();
    }
    __syncthreads();
  }
}

__global__ void k_df_reset_queue(unsigned* cnt, int* pend, const int* pend0, int ngroups,
                                 int* queue, const int* init, int n0, int ntasks,
                                 unsigned* ctr) {
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ngroups || i < ntasks; i += stride) {
    if (i < ngroups) {
      cnt[i] = 0u;
      pend[i] = pend0[i];
    }
    if (i < ntasks) queue[i] = i < n0 ? init[i] : -1;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctr[0] = 0u;
    ctr[1] = (unsigned)n0;
  }
}

void launch_df_reset_queue(unsigned* cnt, int* pend, const int* pend0, int ngroups,
                           int* queue, const int* init, int n0, int ntasks, unsigned* ctr,
                           hipStream_t s) {
  const int m = std::max(ngroups, ntasks);
  const int blocks = std::max(1, std::min(1024, (m + 255) / 256));
  k_df_reset_queue<<<blocks, 256, 0, s>>>(cnt, pend, pend0, ngroups, queue, init, n0, ntasks,
                                          ctr);
}

__global__ void k_df_reset(unsigned* cnt, int count) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
    cnt[i] = 0u;
}

void launch_df_reset(unsigned* cnt, int count, hipStream_t s) {
  const int blocks = std::min(1024, (count + 255) / 256);
  k_df_reset<<<blocks, 256, 0, s>>>(cnt, count);
}

template <typename T>
void launch_apply_dataflow(const T* factor, const DfArgs& a, VecBufs b, int k, int grid,
                           hipStream_t s) {
  if (a.ntasks <= 0) return;
  // the matrix-core variant is a kernel of its own, so its registers do not
  // weigh on the single-right-hand-side kernel
  if (k > 1 || b.multi_min_k <= 1) k_apply_df<T, true><<<grid, 256, 0, s>>>(factor, a, b, k);
  else k_apply_df<T, false><<<grid, 256, 0, s>>>(factor, a, b, 1);
}

template <typename T>
int dataflow_blocks_per_cu(bool multi) {
  // resident workgroups per compute unit of each variant (registers and LDS
  // differ); every workgroup of a launch is resident, the grid being sized
  // from this count
  int nb = 0;
  const hipError_t e =
      multi ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_apply_df<T, true>, 256, 0)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_apply_df<T, false>, 256, 0);
  if (e != hipSuccess || nb < 1) nb = 1;
  return nb;
}

template void launch_apply_dataflow<double>(const double*, const DfArgs&, VecBufs, int, int,
                                            hipStream_t);
template void launch_apply_dataflow<float>(const float*, const DfArgs&, VecBufs, int, int,
                                           hipStream_t);
template int dataflow_blocks_per_cu<double>(bool);
template int dataflow_blocks_per_cu<float>(bool);

// ---- SpMV ------------------------------------------------------------------
// One thread per row for up to 16 columns per pass: the row's entries and
// column indices are read once for all the pass's columns, and each
// column's sum runs over the row's entries in storage order, so a column's
// result does not depend on how many columns share the launch.  Inactive
// columns (active[col] == 0) are neither computed nor written.  Entries go
// in groups of eight: their values and indices load together, then per
// column the eight x entries (clamped to the row's first entry past the
// end, masked), so a row costs a few memory round trips instead of two per
// entry.
constexpr int kSpmvCols = 16;
constexpr int kSpmvEnt = 8;
__global__ __launch_bounds__(256) void k_spmv(int n, const int* __restrict__ p,
                                              const int* __restrict__ idx,
                                              const double* __restrict__ val,
                                              const double* __restrict__ x, int ldx,
                                              double* __restrict__ y, int ldy, int k,
                                              const int* __restrict__ active) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  const int e0 = p[row], e1 = p[row + 1];
  for (int c0 = 0; c0 < k; c0 += kSpmvCols) {
    const int kc = min(kSpmvCols, k - c0);
    bool on[kSpmvCols];
    double acc[kSpmvCols];
#pragma unroll
    for (int c = 0; c < kSpmvCols; ++c) {
      on[c] = c < kc && (!active || active[c0 + c]);
      acc[c] = 0;
    }
    const double* xc = x + (int64_t)c0 * ldx;
    for (int eb = e0; eb < e1; eb += kSpmvEnt) {
      double v[kSpmvEnt];
      int64_t j[kSpmvEnt];
#pragma unroll
      for (int u = 0; u < kSpmvEnt; ++u) {
        const int e = min(eb + u, e1 - 1);
        v[u] = val[e];
        j[u] = idx[e];
      }
#pragma unroll
      for (int c = 0; c < kSpmvCols; ++c) {
        if (!on[c]) continue;
        const double* xcc = xc + (int64_t)c * ldx;
        double xv[kSpmvEnt];
#pragma unroll
        for (int u = 0; u < kSpmvEnt; ++u) xv[u] = xcc[j[u]];
#pragma unroll
        for (int u = 0; u < kSpmvEnt; ++u)
          if (eb + u < e1) acc[c] += v[u] * xv[u];
      }
    }
#pragma unroll
    for (int c = 0; c < kSpmvCols; ++c)
      if (on[c]) y[row + (int64_t)(c0 + c) * ldy] = acc[c];
  }
}

void launch_spmv(int n, const int* p, const int* idx, const double* val,
                 const double* x, int ldx, double* y, int ldy, int k,
                 const int* active, hipStream_t s) {
  if (n <= 0 || k <= 0) return;
  k_spmv<<<(n + 255) / 256, 256, 0, s>>>(n, p, idx, val, x, ldx, y, ldy, k, active);
}

}  // namespace dev
}  // namespace rsc
