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

namespace rsc {
namespace dev {

constexpr int MAXR = 1024;  // largest rank handled in LDS by the solves

__global__ void k_gather(const double* __restrict__ b, int ldb, double* __restrict__ w,
                         const int* __restrict__ perm, int n, int k) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * k) return;
  int i = (int)(t % n), j = (int)(t / n);
  w[i + (int64_t)j * n] = b[perm[i] + (int64_t)j * ldb];
}

__global__ void k_scatter(const double* __restrict__ w, double* __restrict__ x, int ldx,
                          const int* __restrict__ perm, int n, int k) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * k) return;
  int i = (int)(t % n), j = (int)(t / n);
  x[perm[i] + (int64_t)j * ldx] = w[i + (int64_t)j * n];
}

void launch_gather(const double* b, int ldb, double* w, const int* perm, int n,
                   int k, hipStream_t s) {
  int64_t tot = (int64_t)n * k;
  k_gather<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(b, ldb, w, perm, n, k);
}

void launch_scatter(const double* w, double* x, int ldx, const int* perm, int n,
                    int k, hipStream_t s) {
  int64_t tot = (int64_t)n * k;
  k_scatter<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(w, x, ldx, perm, n, k);
}

// ---- vector gathers of a level -------------------------------------------

// Front vectors are processed in chunks of 1024 entries (blockIdx.y), so a
// large supernode spreads over many workgroups.  In the forward assembly a
// chunk adds, child by child, the child entries whose parent position falls
// in the chunk: deterministic, no atomics.
constexpr int VCH = 1024;

__global__ __launch_bounds__(256) void k_fwd_assemble(const VecNode* __restrict__ nodes,
                                                      const int* __restrict__ lnodes,
                                                      const int* __restrict__ children,
                                                      const int* __restrict__ relmap,
                                                      VecBufs b, int k) {
  const VecNode S = nodes[lnodes[blockIdx.x]];
  const int c = S.ncols, R = S.R, tid = threadIdx.x;
  const int p0 = blockIdx.y * VCH, p1 = min(p0 + VCH, c + R);
  if (p0 >= c + R) return;
  for (int col = 0; col < k; ++col) {
    double* f = b.fv + S.fv_off + (int64_t)col * b.fvs;
    const double* wc = b.w + (int64_t)col * b.n;
    for (int i = p0 + tid; i < p1; i += 256) f[i] = i < c ? wc[S.first + i] : 0.0;
    __syncthreads();
    for (int q = 0; q < S.nchild; ++q) {
      const VecNode C = nodes[children[S.child_off + q]];
      const int* rm = relmap + C.relmap_off;
      const double* fc = b.fv + C.fv_off + (int64_t)col * b.fvs + C.ncols;
      // one predicated pass over the child's rows (no dependent searches)
      for (int a = tid; a < C.R; a += 256) {
        const int p = rm[a];
        if (p >= p0 && p < p1) f[p] += fc[a];
      }
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(256) void k_bwd_gather(const VecNode* __restrict__ nodes,
                                                    const int* __restrict__ lnodes,
                                                    const int* __restrict__ rows,
                                                    VecBufs b, int k) {
  const VecNode S = nodes[lnodes[blockIdx.x]];
  const int c = S.ncols, R = S.R, tid = threadIdx.x;
  const int p0 = blockIdx.y * VCH, p1 = min(p0 + VCH, c + R);
  if (p0 >= c + R) return;
  const int* rw = rows + S.rows_off;
  for (int col = 0; col < k; ++col) {
    double* f = b.fv + S.fv_off + (int64_t)col * b.fvs;
    const double* wc = b.w + (int64_t)col * b.n;
    for (int i = p0 + tid; i < p1; i += 256) f[i] = i < c ? wc[S.first + i] : wc[rw[i - c]];
  }
}

void launch_fwd_assemble(const VecNode* nodes, const int* level_nodes, int count,
                         const int* children, const int* relmap, VecBufs b, int k,
                         int max_front, hipStream_t s) {
  if (count <= 0) return;
  const int ny = (max_front + VCH - 1) / VCH;
  k_fwd_assemble<<<dim3(count, ny), 256, 0, s>>>(nodes, level_nodes, children, relmap, b,
                                                  k);
}

void launch_bwd_gather(const VecNode* nodes, const int* level_nodes, int count,
                       const int* rows, VecBufs b, int k, int max_front, hipStream_t s) {
  if (count <= 0) return;
  const int ny = (max_front + VCH - 1) / VCH;
  k_bwd_gather<<<dim3(count, ny), 256, 0, s>>>(nodes, level_nodes, rows, b, k);
}

// ---- matrix-vector tasks ---------------------------------------------------

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__device__ inline double* vsel(const VecBufs& b, int sel, int col) {
  if (sel == 0) return b.w + (int64_t)col * b.n;
  if (sel == 1) return b.fv + (int64_t)col * b.fvs;
  return b.tv + (int64_t)col * b.tvs;
}

// Small tasks: plain loops, few registers (many workgroups per CU).
template <typename T>
__global__ __launch_bounds__(256) void k_gemv_small(const T* __restrict__ F,
                                                    const GvTask* __restrict__ tasks,
                                                    VecBufs b, int k) {
  __shared__ double red[4][64];
  const GvTask t = tasks[blockIdx.x];
  const T* __restrict__ M = F + t.mat;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
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
__global__ __launch_bounds__(256) void k_gemv_tasks(const T* __restrict__ F,
                                                    const GvTask* __restrict__ tasks,
                                                    VecBufs b, int k) {
  __shared__ double red[4][64];
  const GvTask t = tasks[blockIdx.x];
  const T* __restrict__ M = F + t.mat;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
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

// ---- SpMV ------------------------------------------------------------------
// LPR lanes cooperate on one row; the row's partial sums are combined with
// wavefront shuffles.
template <int LPR>
__global__ __launch_bounds__(256) void k_spmv(int n, const int* __restrict__ p,
                                              const int* __restrict__ idx,
                                              const double* __restrict__ val,
                                              const double* __restrict__ x, int ldx,
                                              double* __restrict__ y, int ldy,
                                              const int* __restrict__ active) {
  const int col = blockIdx.y;
  if (active && !active[col]) return;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int row = (int)(t / LPR), sub = (int)(t % LPR);
  const double* xc = x + (int64_t)col * ldx;
  double acc = 0;
  if (row < n)
    for (int e = p[row] + sub; e < p[row + 1]; e += LPR) acc += val[e] * xc[idx[e]];
  for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, LPR);
  if (row < n && sub == 0) y[row + (int64_t)col * ldy] = acc;
}

void launch_spmv(int n, const int* p, const int* idx, const double* val,
                 const double* x, int ldx, double* y, int ldy, int k,
                 const int* active, hipStream_t s, int lpr) {
  dim3 grid;
  grid.y = k;
  grid.z = 1;
  if (lpr >= 32) {
    grid.x = (unsigned)(((int64_t)n * 32 + 255) / 256);
    k_spmv<32><<<grid, 256, 0, s>>>(n, p, idx, val, x, ldx, y, ldy, active);
  } else {
    grid.x = (unsigned)(((int64_t)n * 8 + 255) / 256);
    k_spmv<8><<<grid, 256, 0, s>>>(n, p, idx, val, x, ldx, y, ldy, active);
  }
}

}  // namespace dev
}  // namespace rsc
