// Preconditioner apply M^{-1} = P^T L^{-T} L^{-1} P on the device
// (FactorPreconditioner::solve, src/rschol/solver/FactorPreconditioner.cpp:
// 135-162; FactorManager::forwardSolve / backwardSolve, FactorManager.cpp:
// 6245-6341; Supernode::forwardSolve / backwardSolve, Supernode.cpp:
// 4571-5250; DiagonalBlock::BlockNode::diagonalSolve, DiagonalBlock.cpp:
// 1306-1378) and the CSR SpMV of the PCG operator.
//
// Level-scheduled: one launch per elimination-tree level and direction, one
// workgroup per supernode; diagonal blocks are applied through their stored
// inverses (L^{-T}), so every step is a matrix-vector product.  The forward sweep is multifrontal: a supernode's
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
// Dense diagonal triangles are stored inverted and transposed (Mt = L^{-T},
// upper triangular, leading dimension ld), so both triangular solves are
// matrix-vector products without per-column barriers.

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Column-reduction GEMV: out[q] = sum_{i<m} W[i + q*ld] * src[i] for q < r.
// Each wave owns 8 columns at a time (8 independent load streams per lane),
// lanes stride the contiguous rows, then 8 wavefront reductions.
template <typename T>
__device__ void wg_colreduce(const T* __restrict__ W, int64_t ld, int m, int r,
                             const double* src, double* out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int q0 = wave * 8; q0 < r; q0 += 32) {
    double acc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] = 0;
    const int nq = min(8, r - q0);
    for (int i = lane; i < m; i += 64) {
      const double xi = src[i];
      const T* base = W + i + (int64_t)q0 * ld;
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (u < nq) acc[u] += (double)base[(int64_t)u * ld] * xi;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      double v = wave_sum(acc[u]);
      if (lane == 0 && u < nq) out[q0 + u] = v;
    }
  }
}

// Row-form GEMV: dst[i] (+|-)= sum_{q<r} W[i + q*ld] * t[q] for i < m, one
// thread per row (coalesced), inner loop unrolled for loads in flight.
template <typename T, bool SUB>
__device__ void wg_rowgemv(const T* __restrict__ W, int64_t ld, int m, int r,
                           const double* t, double* dst) {
  for (int i = threadIdx.x; i < m; i += 256) {
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    const T* row = W + i;
    int q = 0;
    for (; q + 4 <= r; q += 4) {
      a0 += (double)row[(int64_t)q * ld] * t[q];
      a1 += (double)row[(int64_t)(q + 1) * ld] * t[q + 1];
      a2 += (double)row[(int64_t)(q + 2) * ld] * t[q + 2];
      a3 += (double)row[(int64_t)(q + 3) * ld] * t[q + 3];
    }
    for (; q < r; ++q) a0 += (double)row[(int64_t)q * ld] * t[q];
    const double acc = (a0 + a1) + (a2 + a3);
    if (SUB) dst[i] -= acc;
    else dst[i] = acc;
  }
}

// out = L^{-1} f with Mt = L^{-T}: out[i] = sum_k Mt[k,i] f[k]
template <typename T>
__device__ void wg_apply_linv(const T* __restrict__ Mt, int ld, int b, const double* f,
                              double* out) {
  wg_colreduce<T>(Mt, ld, b, b, f, out);
}

// out = L^{-T} z with Mt = L^{-T}: out[i] = sum_k Mt[i,k] z[k] (the strictly
// lower part of Mt holds exact zeros)
template <typename T>
__device__ void wg_apply_ltinv(const T* __restrict__ Mt, int ld, int b, const double* z,
                               double* out) {
  wg_rowgemv<T, false>(Mt, ld, b, b, z, out);
}

// t[q] = sum_i W[i,q] src[i] for q < r (W m x r)
template <typename T>
__device__ void wg_tmatvec(const T* __restrict__ W, int m, int r, const double* src,
                           double* t) {
  wg_colreduce<T>(W, m, m, r, src, t);
}

// dst[i] -= sum_q W[i,q] t[q] for i < m (W m x r, ld ldw)
template <typename T>
__device__ void wg_matvec_sub(const T* __restrict__ W, int m, int r, int64_t ldw,
                              const double* t, double* dst) {
  wg_rowgemv<T, true>(W, ldw, m, r, t, dst);
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
    double* tmp = f + c + R;
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
      const T* P = F + S.panel;
      const int ld = c + R;
      wg_apply_linv<T>(P, ld, c, f, tmp);
      __syncthreads();
      // f_R -= L21 y
      wg_rowgemv<T, true>(P + c, ld, R, c, tmp, f + c);
      for (int i = tid; i < c; i += 256) wc[S.first + i] = tmp[i];
      __syncthreads();
    } else {
      const SolveHNode* H = htree + S.htree;
      for (int o = 0; o < S.nhops; ++o) {
        const SolveHNode h = H[hops[S.hops_off + o]];
        if (h.left < 0) {
          const int b = h.r1 - h.r0;
          wg_apply_linv<T>(F + h.leaf, b, b, f + h.r0, tmp);
          __syncthreads();
          for (int i = tid; i < b; i += 256) f[h.r0 + i] = tmp[i];
        } else {
          const SolveHNode L = H[h.left], Rt = H[h.right];
          wg_tmatvec<T>(F + h.U, L.r1 - L.r0, h.rank, f + L.r0, tsh);
          __syncthreads();
          wg_matvec_sub<T>(F + h.V, Rt.r1 - Rt.r0, h.rank, Rt.r1 - Rt.r0, tsh, f + Rt.r0);
        }
        __syncthreads();
      }
      if (R > 0) {
        wg_tmatvec<T>(F + S.U, c, S.rank, f, tsh);
        __syncthreads();
        wg_matvec_sub<T>(F + S.V, R, S.rank, R, tsh, f + c);
      }
      for (int i = tid; i < c; i += 256) wc[S.first + i] = f[i];
      __syncthreads();
    }
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
  const int* rw = rows + S.rows_off;
  for (int col = 0; col < k; ++col) {
    double* z = fv + S.fv_off + (int64_t)col * fvs;
    double* xr = z + c;  // gathered ancestor solution
    double* tmp = z + c + R;
    double* wc = w + (int64_t)col * n;
    for (int i = tid; i < c; i += 256) z[i] = wc[S.first + i];
    for (int a = tid; a < R; a += 256) xr[a] = wc[rw[a]];
    __syncthreads();
    if (!S.compress) {
      const T* P = F + S.panel;
      const int ld = c + R;
      // z -= L21^T x_R
      wg_colreduce<T>(P + c, ld, R, c, xr, tmp);
      __syncthreads();
      for (int i = tid; i < c; i += 256) z[i] -= tmp[i];
      __syncthreads();
      wg_apply_ltinv<T>(P, ld, c, z, tmp);
      __syncthreads();
      for (int i = tid; i < c; i += 256) wc[S.first + i] = tmp[i];
      __syncthreads();
    } else {
      if (R > 0) {
        wg_tmatvec<T>(F + S.V, R, S.rank, xr, tsh);
        __syncthreads();
        wg_matvec_sub<T>(F + S.U, c, S.rank, c, tsh, z);
        __syncthreads();
      }
      const SolveHNode* H = htree + S.htree;
      for (int o = S.nhops - 1; o >= 0; --o) {
        const SolveHNode h = H[hops[S.hops_off + o]];
        if (h.left < 0) {
          const int b = h.r1 - h.r0;
          wg_apply_ltinv<T>(F + h.leaf, b, b, z + h.r0, tmp);
          __syncthreads();
          for (int i = tid; i < b; i += 256) z[h.r0 + i] = tmp[i];
        } else {
          const SolveHNode L = H[h.left], Rt = H[h.right];
          wg_tmatvec<T>(F + h.V, Rt.r1 - Rt.r0, h.rank, z + Rt.r0, tsh);
          __syncthreads();
          wg_matvec_sub<T>(F + h.U, L.r1 - L.r0, h.rank, L.r1 - L.r0, tsh, z + L.r0);
        }
        __syncthreads();
      }
      for (int i = tid; i < c; i += 256) wc[S.first + i] = z[i];
      __syncthreads();
    }
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
