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

__global__ __launch_bounds__(256) void k_fwd_assemble(const VecNode* __restrict__ nodes,
                                                      const int* __restrict__ lnodes,
                                                      const int* __restrict__ children,
                                                      const int* __restrict__ relmap,
                                                      VecBufs b, int k) {
  const VecNode S = nodes[lnodes[blockIdx.x]];
  const int c = S.ncols, R = S.R, tid = threadIdx.x;
  for (int col = 0; col < k; ++col) {
    double* f = b.fv + S.fv_off + (int64_t)col * b.fvs;
    const double* wc = b.w + (int64_t)col * b.n;
    for (int i = tid; i < c; i += 256) f[i] = wc[S.first + i];
    for (int i = tid; i < R; i += 256) f[c + i] = 0.0;
    __syncthreads();
    for (int q = 0; q < S.nchild; ++q) {
      const VecNode C = nodes[children[S.child_off + q]];
      const int* rm = relmap + C.relmap_off;
      const double* fc = b.fv + C.fv_off + (int64_t)col * b.fvs + C.ncols;
      for (int a = tid; a < C.R; a += 256) f[rm[a]] += fc[a];
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
  const int* rw = rows + S.rows_off;
  for (int col = 0; col < k; ++col) {
    double* f = b.fv + S.fv_off + (int64_t)col * b.fvs;
    const double* wc = b.w + (int64_t)col * b.n;
    for (int i = tid; i < c; i += 256) f[i] = wc[S.first + i];
    for (int a = tid; a < R; a += 256) f[c + a] = wc[rw[a]];
  }
}

void launch_fwd_assemble(const VecNode* nodes, const int* level_nodes, int count,
                         const int* children, const int* relmap, VecBufs b, int k,
                         hipStream_t s) {
  if (count <= 0) return;
  k_fwd_assemble<<<count, 256, 0, s>>>(nodes, level_nodes, children, relmap, b, k);
}

void launch_bwd_gather(const VecNode* nodes, const int* level_nodes, int count,
                       const int* rows, VecBufs b, int k, hipStream_t s) {
  if (count <= 0) return;
  k_bwd_gather<<<count, 256, 0, s>>>(nodes, level_nodes, rows, b, k);
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

template <typename T>
__global__ __launch_bounds__(256) void k_gemv_tasks(const T* __restrict__ F,
                                                    const GvTask* __restrict__ tasks,
                                                    VecBufs b, int k) {
  const GvTask t = tasks[blockIdx.x];
  const T* __restrict__ M = F + t.mat;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int col = 0; col < k; ++col) {
    const double* x = vsel(b, t.xsel, col) + t.xoff;
    double* y = vsel(b, t.ysel, col) + t.yoff;
    if (t.kind == 0) {
      for (int i = tid; i < t.rows; i += 256) {
        double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
        const T* row = M + i;
        int j = 0;
        for (; j + 4 <= t.cols; j += 4) {
          a0 += (double)row[(int64_t)j * t.ld] * x[j];
          a1 += (double)row[(int64_t)(j + 1) * t.ld] * x[j + 1];
          a2 += (double)row[(int64_t)(j + 2) * t.ld] * x[j + 2];
          a3 += (double)row[(int64_t)(j + 3) * t.ld] * x[j + 3];
        }
        for (; j < t.cols; ++j) a0 += (double)row[(int64_t)j * t.ld] * x[j];
        const double acc = (a0 + a1) + (a2 + a3);
        if (t.op) y[i] -= acc;
        else y[i] = acc;
      }
    } else {
      const int q0 = wave * 8;
      const int nq = min(8, t.cols - q0);
      if (nq > 0) {
        double acc[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[u] = 0;
        for (int i = lane; i < t.rows; i += 64) {
          const double xi = x[i];
          const T* base = M + i + (int64_t)q0 * t.ld;
#pragma unroll
          for (int u = 0; u < 8; ++u)
            if (u < nq) acc[u] += (double)base[(int64_t)u * t.ld] * xi;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const double v = wave_sum(acc[u]);
          if (lane == 0 && u < nq) {
            if (t.op) y[q0 + u] -= v;
            else y[q0 + u] = v;
          }
        }
      }
    }
  }
}

template <typename T>
void launch_gemv_tasks(const T* factor, const GvTask* tasks, int count, VecBufs b,
                       int k, hipStream_t s) {
  if (count <= 0) return;
  k_gemv_tasks<T><<<count, 256, 0, s>>>(factor, tasks, b, k);
}

template void launch_gemv_tasks<double>(const double*, const GvTask*, int, VecBufs, int,
                                        hipStream_t);
template void launch_gemv_tasks<float>(const float*, const GvTask*, int, VecBufs, int,
                                       hipStream_t);

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
