// Preconditioner apply M^{-1} = P^T L^{-T} L^{-1} P on the device
// (FactorPreconditioner::solve, src/rschol/solver/FactorPreconditioner.cpp:
// 135-162; FactorManager::forwardSolve / backwardSolve, FactorManager.cpp:
// 6245-6341; Supernode::forwardSolve / backwardSolve, Supernode.cpp:
// 4571-5250; DiagonalBlock::BlockNode::diagonalSolve, DiagonalBlock.cpp:
// 1306-1378) and the CSR SpMV of the PCG operator.
//
// Level-scheduled: one launch per elimination-tree level and direction, one
// workgroup per supernode.  The forward sweep is multifrontal: a supernode's
// front vector receives its children's update vectors (no atomics), solves
// with its diagonal block and leaves its own update vector for the parent.
// Vectors are FP64; the factor may be FP32 (widened on load).
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

// ---- workgroup helpers ----------------------------------------------------

// f[0:b] <- L^{-1} f[0:b]; rows b..b+extra of the same columns are updated
// too (L stored b+extra rows tall with leading dimension ld).
template <typename T>
__device__ void wg_lower_solve(const T* __restrict__ L, int ld, int b, int extra,
                               double* f) {
  const int tid = threadIdx.x;
  for (int kk = 0; kk < b; ++kk) {
    if (tid == 0) f[kk] = f[kk] / (double)L[kk + (int64_t)kk * ld];
    __syncthreads();
    const double y = f[kk];
    const T* col = L + (int64_t)kk * ld;
    for (int i = kk + 1 + tid; i < b + extra; i += 256) f[i] -= (double)col[i] * y;
    __syncthreads();
  }
}

// z[0:b] <- L^{-T} z[0:b], L b x b lower (leading dimension ld), row form.
template <typename T>
__device__ void wg_upper_solve(const T* __restrict__ L, int ld, int b, double* z) {
  const int tid = threadIdx.x;
  for (int kk = b - 1; kk >= 0; --kk) {
    if (tid == 0) z[kk] = z[kk] / (double)L[kk + (int64_t)kk * ld];
    __syncthreads();
    const double x = z[kk];
    for (int i = tid; i < kk; i += 256) z[i] -= (double)L[kk + (int64_t)i * ld] * x;
    __syncthreads();
  }
}

// dst[0:m] -= V (m x r) * (U^T src)  with U (c x r): one low-rank pair.
template <typename T>
__device__ void wg_lowrank(const T* __restrict__ U, int c, const T* __restrict__ V,
                           int m, int r, const double* src, double* dst,
                           double* tsh) {
  const int tid = threadIdx.x;
  for (int q = tid; q < r; q += 256) {
    const T* u = U + (int64_t)q * c;
    double acc = 0;
    for (int i = 0; i < c; ++i) acc += (double)u[i] * src[i];
    tsh[q] = acc;
  }
  __syncthreads();
  for (int i = tid; i < m; i += 256) {
    double acc = 0;
    for (int q = 0; q < r; ++q) acc += (double)V[i + (int64_t)q * m] * tsh[q];
    dst[i] -= acc;
  }
  __syncthreads();
}

// dst[0:c] -= U (c x r) * (V^T src), V (m x r): transposed low-rank pair.
template <typename T>
__device__ void wg_lowrank_t(const T* __restrict__ U, int c, const T* __restrict__ V,
                             int m, int r, const double* src, double* dst,
                             double* tsh) {
  const int tid = threadIdx.x;
  for (int q = tid; q < r; q += 256) {
    const T* v = V + (int64_t)q * m;
    double acc = 0;
    for (int i = 0; i < m; ++i) acc += (double)v[i] * src[i];
    tsh[q] = acc;
  }
  __syncthreads();
  for (int i = tid; i < c; i += 256) {
    double acc = 0;
    for (int q = 0; q < r; ++q) acc += (double)U[i + (int64_t)q * c] * tsh[q];
    dst[i] -= acc;
  }
  __syncthreads();
}

// ---- forward sweep ---------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(256) void k_forward(
    const T* __restrict__ F, const SolveNode* __restrict__ nodes,
    const int* __restrict__ lnodes, const SolveHNode* __restrict__ htree,
    const int* __restrict__ hops, const int* __restrict__ children,
    const int* __restrict__ relmap, double* __restrict__ w, int n, int k,
    double* __restrict__ fv, int64_t fvs) {
  __shared__ double tsh[MAXR];
  const int j = lnodes[blockIdx.x];
  const SolveNode S = nodes[j];
  const int c = S.ncols, R = S.R, tid = threadIdx.x;
  for (int col = 0; col < k; ++col) {
    double* f = fv + S.fv_off + (int64_t)col * fvs;
    double* wc = w + (int64_t)col * n;
    for (int i = tid; i < c; i += 256) f[i] = wc[S.first + i];
    for (int i = tid; i < R; i += 256) f[c + i] = 0.0;
    __syncthreads();
    for (int q = 0; q < S.nchild; ++q) {
      const SolveNode C = nodes[children[S.child_off + q]];
      const int* rm = relmap + C.relmap_off;
      const double* fc = fv + C.fv_off + (int64_t)col * fvs + C.ncols;
      for (int a = tid; a < C.R; a += 256) f[rm[a]] += fc[a];
      __syncthreads();
    }
    if (!S.compress) {
      wg_lower_solve<T>(F + S.panel, c + R, c, R, f);
    } else {
      const SolveHNode* H = htree + S.htree;
      for (int o = 0; o < S.nhops; ++o) {
        const SolveHNode h = H[hops[S.hops_off + o]];
        if (h.left < 0) {
          wg_lower_solve<T>(F + h.leaf, h.r1 - h.r0, h.r1 - h.r0, 0, f + h.r0);
        } else {
          const SolveHNode L = H[h.left], Rt = H[h.right];
          wg_lowrank<T>(F + h.U, L.r1 - L.r0, F + h.V, Rt.r1 - Rt.r0, h.rank,
                        f + L.r0, f + Rt.r0, tsh);
        }
      }
      if (R > 0) wg_lowrank<T>(F + S.U, c, F + S.V, R, S.rank, f, f + c, tsh);
    }
    for (int i = tid; i < c; i += 256) wc[S.first + i] = f[i];
    __syncthreads();
  }
}

// ---- backward sweep --------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(256) void k_backward(
    const T* __restrict__ F, const SolveNode* __restrict__ nodes,
    const int* __restrict__ lnodes, const SolveHNode* __restrict__ htree,
    const int* __restrict__ hops, const int* __restrict__ rows,
    double* __restrict__ w, int n, int k, double* __restrict__ fv, int64_t fvs) {
  __shared__ double tsh[MAXR];
  const int j = lnodes[blockIdx.x];
  const SolveNode S = nodes[j];
  const int c = S.ncols, R = S.R, tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int* rw = rows + S.rows_off;
  for (int col = 0; col < k; ++col) {
    double* z = fv + S.fv_off + (int64_t)col * fvs;
    double* xr = z + c;  // gathered ancestor solution
    double* wc = w + (int64_t)col * n;
    for (int i = tid; i < c; i += 256) z[i] = wc[S.first + i];
    for (int a = tid; a < R; a += 256) xr[a] = wc[rw[a]];
    __syncthreads();
    if (!S.compress) {
      const T* P = F + S.panel;
      const int ld = c + R;
      for (int kk = wave; kk < c; kk += 4) {
        const T* colp = P + c + (int64_t)kk * ld;
        double acc = 0;
        for (int a = lane; a < R; a += 64) acc += (double)colp[a] * xr[a];
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
        if (lane == 0) z[kk] -= acc;
      }
      __syncthreads();
      wg_upper_solve<T>(P, ld, c, z);
    } else {
      if (R > 0) wg_lowrank_t<T>(F + S.U, c, F + S.V, R, S.rank, xr, z, tsh);
      const SolveHNode* H = htree + S.htree;
      for (int o = S.nhops - 1; o >= 0; --o) {
        const SolveHNode h = H[hops[S.hops_off + o]];
        if (h.left < 0) {
          wg_upper_solve<T>(F + h.leaf, h.r1 - h.r0, h.r1 - h.r0, z + h.r0);
        } else {
          const SolveHNode L = H[h.left], Rt = H[h.right];
          wg_lowrank_t<T>(F + h.U, L.r1 - L.r0, F + h.V, Rt.r1 - Rt.r0, h.rank,
                          z + Rt.r0, z + L.r0, tsh);
        }
      }
    }
    for (int i = tid; i < c; i += 256) wc[S.first + i] = z[i];
    __syncthreads();
  }
}

template <typename T>
void launch_forward(const T* factor, const SolveNode* nodes, const int* level_nodes,
                    int count, const SolveHNode* htree, const int* hops,
                    const int* children, const int* relmap, double* w, int n, int k,
                    double* fv, int64_t fv_stride, hipStream_t s) {
  if (count <= 0) return;
  k_forward<T><<<count, 256, 0, s>>>(factor, nodes, level_nodes, htree, hops,
                                     children, relmap, w, n, k, fv, fv_stride);
}

template <typename T>
void launch_backward(const T* factor, const SolveNode* nodes, const int* level_nodes,
                     int count, const SolveHNode* htree, const int* hops,
                     const int* rows, double* w, int n, int k, double* fv,
                     int64_t fv_stride, hipStream_t s) {
  if (count <= 0) return;
  k_backward<T><<<count, 256, 0, s>>>(factor, nodes, level_nodes, htree, hops, rows,
                                      w, n, k, fv, fv_stride);
}

template void launch_forward<double>(const double*, const SolveNode*, const int*, int,
                                     const SolveHNode*, const int*, const int*,
                                     const int*, double*, int, int, double*, int64_t,
                                     hipStream_t);
template void launch_forward<float>(const float*, const SolveNode*, const int*, int,
                                    const SolveHNode*, const int*, const int*,
                                    const int*, double*, int, int, double*, int64_t,
                                    hipStream_t);
template void launch_backward<double>(const double*, const SolveNode*, const int*, int,
                                      const SolveHNode*, const int*, const int*,
                                      double*, int, int, double*, int64_t, hipStream_t);
template void launch_backward<float>(const float*, const SolveNode*, const int*, int,
                                     const SolveHNode*, const int*, const int*,
                                     double*, int, int, double*, int64_t, hipStream_t);

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
