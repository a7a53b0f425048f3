// Batched dense kernels for gfx950 (CDNA4): MFMA GEMM, tile Cholesky,
// tile triangular solve, 2D copies, CholeskyQR shift and the multifrontal
// front assembly.  Compiled with -Xclang -target-feature -Xclang
// -packed-fp32-ops disabled; no packed-FP32 instructions are written here.
#include "kernels.h"

#include <algorithm>
#include <cmath>

namespace rsc {
namespace dev {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Acc;
// Accumulator row of element v in lane l differs between the two shapes:
// f64 16x16x4 holds rows (l/16) + 4v, f32 16x16x4 holds rows 4(l/16) + v
// (column l%16 in both).
template <>
struct Acc<double> {
  typedef d4 type;
  static __device__ inline d4 mfma(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <>
struct Acc<float> {
  typedef f4 type;
  static __device__ inline f4 mfma(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// ---------------------------------------------------------------------------
// GEMM: 64x64 output tile per 256-thread workgroup; each wave owns a 32x32
// quadrant computed as 2x2 MFMA 16x16x4 blocks.  K is staged through LDS in
// slabs of 32, software-pipelined: the global loads of the next slab are
// issued into registers before the matrix-core work on the current slab,
// so a launch pays about one memory round trip instead of one per slab
// (the operands of these small products are usually cold in L2: they were
// written by the previous launch).  MFMA 16x16x4 operand layout: lane l
// supplies A[l%16][l/16] and B[l/16][l%16]; accumulator element v of lane l
// is D[Acc<T>::row(l,v)][l%16].
// ---------------------------------------------------------------------------
typedef short bf4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16 bits, round to nearest even (finite inputs).  BF16 keeps the
// FP32 exponent range, so the products need no block scaling: scaling a
// block by a power of two would change no rounding.
__device__ inline short to_bf16(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

// BF (FP32 storage only): the products run as BF16 MFMA 16x16x16 with FP32
// accumulation (one instruction per 16-deep slab); lane l supplies
// A[l%16][4(l/16)..4(l/16)+3] and B[4(l/16)..][l%16], and the accumulator
// layout is that of the FP32 16x16x4 instruction.
template <typename T, bool TA, bool TB, bool BF = false>
__global__ __launch_bounds__(256) void k_gemm(const GemmJob* __restrict__ jobs,
                                              const GemmTile* __restrict__ tiles,
                                              T* __restrict__ partials,
                                              unsigned* __restrict__ counters) {
  constexpr int KS = 32;
  constexpr int NE = KS * TILE / 256;  // elements of A (and of B) per thread per slab
  __shared__ T As[KS][TILE + 1];
  __shared__ T Bs[KS][TILE + 1];
  __shared__ int last;
  const GemmTile tl = tiles[blockIdx.x];
  const GemmJob J = jobs[tl.job];
  const T* __restrict__ A = (const T*)J.A;
  const T* __restrict__ B = (const T*)J.B;
  T* __restrict__ C = (T*)J.C;
  const int m0 = tl.tm * TILE, n0 = tl.tn * TILE;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  typedef typename Acc<T>::type V;
  V acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int v = 0; v < 4; ++v) acc[a][b][v] = 0;
  const int kend = tl.k1;
  // element q of this thread's share of a slab: row/column index in the
  // tile and k offset in the slab (consecutive threads read consecutive
  // addresses of the stored matrix)
  auto a_idx = [&](int q, int& i, int& kk) {
    const int e = tid + 256 * q;
    if (!TA) { i = e % TILE; kk = e / TILE; }
    else { kk = e % KS; i = e / KS; }
  };
  auto b_idx = [&](int q, int& j, int& kk) {
    const int e = tid + 256 * q;
    if (!TB) { kk = e % KS; j = e / KS; }
    else { j = e % TILE; kk = e / TILE; }
  };
  T ra[NE], rb[NE];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int q = 0; q < NE; ++q) {
      int i, kk;
      a_idx(q, i, kk);
      const int gi = m0 + i, gk = k0 + kk;
      ra[q] = (gi < J.m && gk < kend)
                  ? (TA ? A[gk + (int64_t)gi * J.lda] : A[gi + (int64_t)gk * J.lda])
                  : (T)0;
      int j, kb;
      b_idx(q, j, kb);
      const int gj = n0 + j, gkb = k0 + kb;
      rb[q] = (gj < J.n && gkb < kend)
                  ? (TB ? B[gj + (int64_t)gkb * J.ldb] : B[gkb + (int64_t)gj * J.ldb])
                  : (T)0;
    }
  };
  if (tl.k0 < kend) fetch(tl.k0);
  for (int k0 = tl.k0; k0 < kend; k0 += KS) {
#pragma unroll
    for (int q = 0; q < NE; ++q) {
      int i, kk;
      a_idx(q, i, kk);
      As[kk][i] = ra[q];
      int j, kb;
      b_idx(q, j, kb);
      Bs[kb][j] = rb[q];
    }
    __syncthreads();
    // the next slab's loads are in flight during this slab's products
    if (k0 + KS < kend) fetch(k0 + KS);
    if constexpr (BF) {
#pragma unroll
      for (int ks = 0; ks < KS; ks += 16) {
        const int k4 = ks + 4 * (lane >> 4), li = lane & 15;
        bf4 a0, a1, b0, b1;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          a0[e] = to_bf16((float)As[k4 + e][wm + li]);
          a1[e] = to_bf16((float)As[k4 + e][wm + 16 + li]);
          b0[e] = to_bf16((float)Bs[k4 + e][wn + li]);
          b1[e] = to_bf16((float)Bs[k4 + e][wn + 16 + li]);
        }
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b1, acc[1][1], 0, 0, 0);
      }
    } else {
#pragma unroll
      for (int kk = 0; kk < KS; kk += 4) {
        T a0 = As[kk + (lane >> 4)][wm + (lane & 15)];
        T a1 = As[kk + (lane >> 4)][wm + 16 + (lane & 15)];
        T b0 = Bs[kk + (lane >> 4)][wn + (lane & 15)];
        T b1 = Bs[kk + (lane >> 4)][wn + 16 + (lane & 15)];
        acc[0][0] = Acc<T>::mfma(a0, b0, acc[0][0]);
        acc[0][1] = Acc<T>::mfma(a0, b1, acc[0][1]);
        acc[1][0] = Acc<T>::mfma(a1, b0, acc[1][0]);
        acc[1][1] = Acc<T>::mfma(a1, b1, acc[1][1]);
      }
    }
    __syncthreads();
  }
  if (tl.ns > 1) {
    // split k: publish this share, and let the last share of the tile sum
    // all of them in index order
    T* P = partials + (int64_t)(tl.part + tl.si) * (TILE * TILE);
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int v = 0; v < 4; ++v)
          P[(wm + a * 16 + Acc<T>::row(lane, v)) + (wn + b * 16 + (lane & 15)) * TILE] =
              acc[a][b][v];
    __threadfence();
    __syncthreads();
    if (tid == 0) last = atomicAdd(&counters[tl.cnt], 1u) == (unsigned)(tl.ns - 1);
    __syncthreads();
    if (!last) return;
    __threadfence();
    const T* P0 = partials + (int64_t)tl.part * (TILE * TILE);
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int v = 0; v < 4; ++v) {
          const int o = (wm + a * 16 + Acc<T>::row(lane, v)) + (wn + b * 16 + (lane & 15)) * TILE;
          T sum = 0;
          for (int q = 0; q < tl.ns; ++q) sum += P0[(int64_t)q * (TILE * TILE) + o];
          acc[a][b][v] = sum;
        }
    if (tid == 0) counters[tl.cnt] = 0;  // ready for the next factorization
  }
  const T alpha = (T)J.alpha, beta = (T)J.beta;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int v = 0; v < 4; ++v) {
        int gi = m0 + wm + a * 16 + Acc<T>::row(lane, v);
        int gj = n0 + wn + b * 16 + (lane & 15);
        if (gi < J.m && gj < J.n) {
          T* p = C + gi + (int64_t)gj * J.ldc;
          T r = alpha * acc[a][b][v];
          if (beta != (T)0) r += beta * (*p);
          *p = r;
        }
      }
}

int plan_gemm_tiles(const GemmJob* jobs, int njobs, bool lower,
                    std::vector<GemmTile>& tiles, int* ncounters) {
  const size_t first = tiles.size();
  for (int jb = 0; jb < njobs; ++jb) {
    const GemmJob& J = jobs[jb];
    if (J.m <= 0 || J.n <= 0) continue;
    int tm = (J.m + TILE - 1) / TILE, tn = (J.n + TILE - 1) / TILE;
    for (int a = 0; a < tm; ++a)
      for (int b = 0; b < tn; ++b) {
        if (lower && a < b) continue;
        tiles.push_back({jb, a, b, 0, J.k, 1, 0, 0, 0});
      }
  }
  // about two workgroups per CU of 256, and at least 256 of k per share
  const int64_t ntile = (int64_t)(tiles.size() - first);
  if (ntile == 0 || ntile >= 256) return 0;
  const int want = (int)std::max<int64_t>(1, 512 / ntile);
  std::vector<GemmTile> out;
  int parts = 0;
  for (size_t t = first; t < tiles.size(); ++t) {
    GemmTile g = tiles[t];
    const int k = jobs[g.job].k;
    int ns = std::min(want, std::min(64, (k + 255) / 256));
    if (ns <= 1) {
      out.push_back(g);
      continue;
    }
    int chunk = (k + ns - 1) / ns;
    chunk = (chunk + 31) / 32 * 32;
    ns = (k + chunk - 1) / chunk;
    if (ns <= 1) {
      out.push_back(g);
      continue;
    }
    const int cnt = (*ncounters)++;
    for (int q = 0; q < ns; ++q)
      out.push_back({g.job, g.tm, g.tn, q * chunk, std::min(k, (q + 1) * chunk), ns, q,
                     parts, cnt});
    parts += ns;
  }
  tiles.resize(first);
  tiles.insert(tiles.end(), out.begin(), out.end());
  return parts;
}

template <typename T>
void launch_gemm(bool ta, bool tb, const GemmJob* jobs, const GemmTile* tiles,
                 int ntiles, void* partials, unsigned* counters, hipStream_t s) {
  if (ntiles <= 0) return;
  T* P = (T*)partials;
  if (!ta && !tb) k_gemm<T, false, false><<<ntiles, 256, 0, s>>>(jobs, tiles, P, counters);
  else if (ta && !tb) k_gemm<T, true, false><<<ntiles, 256, 0, s>>>(jobs, tiles, P, counters);
  else if (!ta && tb) k_gemm<T, false, true><<<ntiles, 256, 0, s>>>(jobs, tiles, P, counters);
  else k_gemm<T, true, true><<<ntiles, 256, 0, s>>>(jobs, tiles, P, counters);
}

void launch_gemm_bf16(bool ta, bool tb, const GemmJob* jobs, const GemmTile* tiles,
                      int ntiles, void* partials, unsigned* counters, hipStream_t s) {
  if (ntiles <= 0) return;
  float* P = (float*)partials;
  if (!ta && !tb)
    k_gemm<float, false, false, true><<<ntiles, 256, 0, s>>>(jobs, tiles, P, counters);
  else if (ta && !tb)
    k_gemm<float, true, false, true><<<ntiles, 256, 0, s>>>(jobs, tiles, P, counters);
  else if (!ta && tb)
    k_gemm<float, false, true, true><<<ntiles, 256, 0, s>>>(jobs, tiles, P, counters);
  else k_gemm<float, true, true, true><<<ntiles, 256, 0, s>>>(jobs, tiles, P, counters);
}

// ---------------------------------------------------------------------------
// Tile Cholesky (n <= 64), right-looking.  A non-positive pivot sets
// status[0] = 1 and status[1] = min(node); the pivot is replaced by 1 so the
// remaining arithmetic stays finite (the factorization is then retried).
// ---------------------------------------------------------------------------
// One wave per tile, lane i holding row i in registers; the column loop is
// unrolled so every register index is a constant.  Step k reads the pivot
// from lane k (readlane), every lane scales its entry of column k, the
// column is published in LDS, and each lane updates the entries of its row
// right of column k with the column's values read back from LDS (one
// address for the whole wave).  A step is a short chain (pivot, square
// root, division, one LDS round trip) with no barrier between waves; the
// tile's loads are all issued at once.  Measured on gfx950 this is about
// twice as fast as a four-wave tile with one workgroup barrier per column
// and as the same one-wave tile with rolled, blocked column loops.  The
// arithmetic per entry is that of the textbook right-looking loop in the
// same order.
__device__ inline double lane_value(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}
__device__ inline float lane_value(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

template <typename T>
__global__ __launch_bounds__(64) void k_potrf(const PotrfJob* __restrict__ jobs,
                                              int* __restrict__ status) {
  __shared__ T col[TILE];
  const PotrfJob J = jobs[blockIdx.x];
  T* A = (T*)J.A;
  const int n = J.n, i = threadIdx.x;
  const int ic = min(i, n - 1);
  T a[TILE];
#pragma unroll
  for (int j = 0; j < TILE; ++j) {
    const T v = A[ic + (int64_t)min(j, n - 1) * J.lda];
    a[j] = (i < n && j < n && i >= j) ? v : (T)0;
  }
  // largest positive diagonal entry (0 if none; NaNs do not count)
  const T dg = A[ic + (int64_t)ic * J.lda];
  T mx = (i < n && dg > (T)0) ? dg : (T)0;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const T y = __shfl_xor(mx, o, 64);
    mx = y > mx ? y : mx;
  }
  const T u = sizeof(T) == 8 ? (T)1.1102230246251565e-16 : (T)5.960464477539063e-08;
  const T rel = J.tolerant == 2 ? (T)1e-5 : u;
  const T dfloor = mx > (T)0 ? mx * rel : (T)1;
  int bad = 0;  // the same in every lane
#pragma unroll
  for (int k = 0; k < TILE; ++k) {
    if (k < n) {
      T d = lane_value(a[k], k);
      if (J.tolerant == 2) {
        if (!(d > dfloor)) {
          d = dfloor;
          bad += 1;
        }
      } else if (J.tolerant == 3) {
        // Gram matrix of columns of norm <= 1: a pivot at rounding level
        // is a column dependent on the previous ones; an infinite-like
        // pivot zeroes it in L and in the triangular solve that follows
        if (!(d > (T)1e-13)) d = sizeof(T) == 8 ? (T)1e300 : (T)1e30;
      } else if (!(d > (T)0)) {
        if (J.tolerant) {
          d = dfloor;
        } else {
          bad = 1;
          d = (T)1;
        }
      }
      const T piv = sqrt(d);
      const T l = i > k ? a[k] / piv : piv;
      a[k] = i >= k ? l : a[k];
      const T lc = (i > k && i < n) ? l : (T)0;
      col[i] = lc;
      __syncthreads();
      // entries above the diagonal and rows >= n take harmless values that
      // are never stored (col is zero outside rows k+1..n-1)
#pragma unroll
      for (int j = k + 1; j < TILE; ++j) a[j] -= lc * col[j];
      __syncthreads();
    }
  }
#pragma unroll
  for (int j = 0; j < TILE; ++j)
    if (i < n && j < n && i >= j) A[i + (int64_t)j * J.lda] = a[j];
  if (i == 0 && bad) {
    if (J.tolerant == 2) {
      atomicAdd(&status[2], bad);
    } else {
      atomicOr(&status[0], 1);
      atomicMin(&status[1], J.node);
    }
  }
}

template <typename T>
void launch_potrf(const PotrfJob* jobs, int njobs, int* status, hipStream_t s) {
  if (njobs <= 0) return;
  k_potrf<T><<<njobs, 64, 0, s>>>(jobs, status);
}

// ---------------------------------------------------------------------------
// Row-block triangular solve B <- B L^{-T} (L nb x nb lower, nb <= 64), 64
// rows of B per 256-thread workgroup, blocked by 16 columns on the matrix
// cores.  L is staged in LDS (all loads issued at once; identity beyond
// nb), and wave w inverts the 16 x 16 diagonal block w by substitution
// (lane j computes column j).  Wave w then owns rows 16w..16w+15 of the
// tile, held in registers in the MFMA accumulator layout, and runs the
// block recurrence without waiting for the other waves:
//   X_k = B_k D_k^{-T},   B_j -= X_k L_jk^T  (j > k),
// each a 16 x 16 x 16 product of four MFMA 16x16x4 steps (operands moved to
// the input layout through a per-wave LDS block).  The sequential part is 4
// block steps instead of 64 column steps.  Results agree with column
// substitution to rounding (the 16 x 16 diagonal blocks are applied as
// their computed inverses).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_trsm(const TrsmJob* __restrict__ jobs,
                                              const RowTile* __restrict__ tiles) {
  __shared__ T Ls[TILE][TILE + 1];  // Ls[i][j] = L_ij on and below the diagonal
  __shared__ T Di[4][16][17];       // inverses of the diagonal blocks
  __shared__ T Xs[4][16][17];       // per wave: a block moved to the input layout
  typedef typename Acc<T>::type V;
  const RowTile tl = tiles[blockIdx.x];
  const TrsmJob J = jobs[tl.job];
  const T* L = (const T*)J.L;
  T* B = (T*)J.B;
  const int nb = J.nb, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int rows = min(TILE, J.m - tl.row0);
  T lv[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int e = tid + 256 * q, i = e % TILE, j = e / TILE;
    lv[q] = L[min(i, nb - 1) + (int64_t)min(j, nb - 1) * J.ldl];
  }
  // this wave's 16 rows, accumulator layout: bacc[c][v] = B[16w + row(lane, v)][16c + li]
  V bacc[4];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = 16 * w + Acc<T>::row(lane, v), col = 16 * c + li;
      const T x = B[(tl.row0 + min(r, rows - 1)) + (int64_t)min(col, nb - 1) * J.ldb];
      bacc[c][v] = (r < rows && col < nb) ? x : (T)0;
    }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int e = tid + 256 * q, i = e % TILE, j = e / TILE;
    Ls[i][j] = (i >= j && i < nb) ? lv[q] : (i == j ? (T)1 : (T)0);
  }
  __syncthreads();
  if (lane < 16) {
    // column li of the inverse of diagonal block w (zero above the diagonal)
    const int o = 16 * w;
    T x[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      T sum = i == li ? (T)1 : (T)0;
#pragma unroll
      for (int k = 0; k < i; ++k) sum -= Ls[o + i][o + k] * x[k];
      x[i] = sum / Ls[o + i][o + i];
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) Di[w][i][li] = x[i];
  }
  __syncthreads();
  auto stage = [&](const V& acc) {
#pragma unroll
    for (int v = 0; v < 4; ++v) Xs[w][Acc<T>::row(lane, v)][li] = acc[v];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
    if (16 * kb >= nb) break;
    // X_k = B_k D_k^{-T}: A[i][k] = B_k[i][k], Bop[k][j] = D_k^{-1}[j][k]
    stage(bacc[kb]);
    V x;
#pragma unroll
    for (int v = 0; v < 4; ++v) x[v] = 0;
#pragma unroll
    for (int s = 0; s < 4; ++s)
      x = Acc<T>::mfma(Xs[w][li][4 * s + lk], Di[kb][li][4 * s + lk], x);
    bacc[kb] = x;
    if (16 * (kb + 1) >= nb) break;
    // B_j -= X_k L_jk^T: A[i][k] = -X_k[i][k], Bop[k][j] = L[16j' + j][16kb + k]
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    stage(x);
#pragma unroll
    for (int jb = kb + 1; jb < 4; ++jb)
#pragma unroll
      for (int s = 0; s < 4; ++s)
        bacc[jb] = Acc<T>::mfma(-Xs[w][li][4 * s + lk], Ls[16 * jb + li][16 * kb + 4 * s + lk],
                                bacc[jb]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = 16 * w + Acc<T>::row(lane, v), col = 16 * c + li;
      if (r < rows && col < nb) B[(tl.row0 + r) + (int64_t)col * J.ldb] = bacc[c][v];
    }
}

template <typename T>
void launch_trsm(const TrsmJob* jobs, const RowTile* tiles, int ntiles,
                 hipStream_t s) {
  if (ntiles <= 0) return;
  k_trsm<T><<<ntiles, 256, 0, s>>>(jobs, tiles);
}

// ---------------------------------------------------------------------------
// 2D copy; tiles of 64 columns.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_copy(const CopyJob* __restrict__ jobs,
                                              const RowTile* __restrict__ tiles) {
  const RowTile tl = tiles[blockIdx.x];
  const CopyJob J = jobs[tl.job];
  const T* src = (const T*)J.src;
  T* dst = (T*)J.dst;
  const int c0 = tl.row0, c1 = min(J.n, c0 + TILE);
  for (int j = c0; j < c1; ++j)
    for (int i = threadIdx.x; i < J.m; i += 256)
      dst[i + (int64_t)j * J.ldd] = src[i + (int64_t)j * J.lds];
}

template <typename T>
void launch_copy(const CopyJob* jobs, const RowTile* tiles, int ntiles,
                 hipStream_t s) {
  if (ntiles <= 0) return;
  k_copy<T><<<ntiles, 256, 0, s>>>(jobs, tiles);
}

// ---------------------------------------------------------------------------
// Identity fill; tiles of 64 columns.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_eye(const EyeJob* __restrict__ jobs,
                                             const RowTile* __restrict__ tiles) {
  const RowTile tl = tiles[blockIdx.x];
  const EyeJob J = jobs[tl.job];
  T* A = (T*)J.A;
  const int c1 = min(J.n, tl.row0 + TILE);
  for (int j = tl.row0; j < c1; ++j)
    for (int i = threadIdx.x; i < J.n; i += 256)
      A[i + (int64_t)j * J.lda] = (i == j) ? (T)1 : (T)0;
}

template <typename T>
void launch_eye(const EyeJob* jobs, const RowTile* tiles, int ntiles, hipStream_t s) {
  if (ntiles <= 0) return;
  k_eye<T><<<ntiles, 256, 0, s>>>(jobs, tiles);
}

// ---------------------------------------------------------------------------
// Debug scan for non-finite values (one workgroup per job).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_check(const CheckJob* __restrict__ jobs,
                                               int* __restrict__ status) {
  const CheckJob J = jobs[blockIdx.x];
  const T* p = (const T*)J.p;
  int bad = 0;
  for (int j = 0; j < J.cols; ++j)
    for (int i = threadIdx.x; i < J.rows; i += 256)
      if (!isfinite((double)p[i + (int64_t)j * J.ld])) bad = 1;
  if (bad) atomicMin(&status[3], J.tag);
}

template <typename T>
void launch_check(const CheckJob* jobs, int njobs, int* status, hipStream_t s) {
  if (njobs <= 0) return;
  k_check<T><<<njobs, 256, 0, s>>>(jobs, status);
}

// ---------------------------------------------------------------------------
// Diagonal shift of a Gram matrix (shifted CholeskyQR, Fukaya et al. 2020).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_shift(const ShiftJob* __restrict__ jobs) {
  __shared__ double red[256];
  const ShiftJob J = jobs[blockIdx.x];
  T* W = (T*)J.W;
  double t = 0;
  for (int i = threadIdx.x; i < J.r; i += 256) t += (double)W[i + (int64_t)i * J.ld];
  red[threadIdx.x] = t;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const double u = sizeof(T) == 8 ? 1.1102230246251565e-16 : 5.960464477539063e-08;
  const double sft = 11.0 * ((double)J.m * J.r + (double)J.r * (J.r + 1)) * u * red[0];
  for (int i = threadIdx.x; i < J.r; i += 256) W[i + (int64_t)i * J.ld] += (T)sft;
}

template <typename T>
void launch_shift(const ShiftJob* jobs, int njobs, hipStream_t s) {
  if (njobs <= 0) return;
  k_shift<T><<<njobs, 256, 0, s>>>(jobs);
}

// ---------------------------------------------------------------------------
// Multifrontal front assembly for one level.  A task covers 16 front columns
// of one supernode: zero them, scatter the matrix entries of those columns,
// then add every child's update matrix entries that land in them, child by
// child (deterministic, no atomics).
// ---------------------------------------------------------------------------
constexpr int ASM_W = 16;

template <typename T>
__global__ __launch_bounds__(256) void k_assemble(
    T* __restrict__ arena, const FrontDev* __restrict__ fronts,
    const AsmTask* __restrict__ tasks, const int* __restrict__ children,
    const int* __restrict__ relmap, const int* __restrict__ lowp,
    const int64_t* __restrict__ a_dst, const int64_t* __restrict__ a_src,
    const double* __restrict__ Ax) {
  const AsmTask tk = tasks[blockIdx.x];
  const FrontDev F = fronts[tk.node];
  const int c = F.ncols, R = F.R, ld = c + R;
  const int t0 = tk.col0, t1 = min(t0 + ASM_W, c + R);
  T* panel = arena + F.panel;
  T* upd = arena + F.update;
  // 1. zero
  for (int tb = t0; tb < t1; ++tb) {
    if (tb < c) {
      T* col = panel + (int64_t)tb * ld;
      for (int i = threadIdx.x; i < ld; i += 256) col[i] = 0;
    } else {
      T* col = upd + (int64_t)(tb - c) * R;
      for (int i = threadIdx.x; i < R; i += 256) col[i] = 0;
    }
  }
  __syncthreads();
  // 2. matrix entries of panel columns
  for (int tb = t0; tb < min(t1, c); ++tb) {
    const int k = F.first + tb;
    for (int e = lowp[k] + threadIdx.x; e < lowp[k + 1]; e += 256)
      panel[a_dst[e]] = (T)Ax[a_src[e]];
  }
  __syncthreads();
  // 3. children, one at a time
  for (int q = 0; q < F.nchild; ++q) {
    const FrontDev Cf = fronts[children[F.child_off + q]];
    const int* rm = relmap + Cf.relmap_off;
    const int Rc = Cf.R;
    const T* U = arena + Cf.update;
    // child columns b with rm[b] in [t0, t1): contiguous since rm is sorted
    int lo = 0, hi = Rc;
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (rm[mid] < t0) lo = mid + 1; else hi = mid;
    }
    const int b0 = lo;
    hi = Rc;
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (rm[mid] < t1) lo = mid + 1; else hi = mid;
    }
    const int b1 = lo;
    for (int b = b0; b < b1; ++b) {
      const int tb = rm[b];
      const T* ucol = U + (int64_t)b * Rc;
      if (tb < c) {
        T* col = panel + (int64_t)tb * ld;
        for (int a = b + threadIdx.x; a < Rc; a += 256) col[rm[a]] += ucol[a];
      } else {
        T* col = upd + (int64_t)(tb - c) * R - c;
        for (int a = b + threadIdx.x; a < Rc; a += 256) col[rm[a]] += ucol[a];
      }
    }
    __syncthreads();
  }
}

template <typename T>
void launch_assemble(T* arena, const FrontDev* fronts, const AsmTask* tasks,
                     int ntasks, const int* children, const int* relmap,
                     const int* lowp, const int64_t* a_dst, const int64_t* a_src,
                     const double* Ax, hipStream_t s) {
  if (ntasks <= 0) return;
  k_assemble<T><<<ntasks, 256, 0, s>>>(arena, fronts, tasks, children, relmap,
                                       lowp, a_dst, a_src, Ax);
}

__global__ __launch_bounds__(256) void k_convjobs(const ConvJob* __restrict__ jobs) {
  const ConvJob J = jobs[blockIdx.y];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < J.count;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (J.to64) ((double*)J.dst)[i] = (double)((const float*)J.src)[i];
    else ((float*)J.dst)[i] = (float)((const double*)J.src)[i];
  }
}

void launch_convjobs(const ConvJob* jobs, int njobs, hipStream_t s) {
  if (njobs <= 0) return;
  k_convjobs<<<dim3(64, njobs), 256, 0, s>>>(jobs);
}

#define RSC_INST(T)                                                            \
  template void launch_gemm<T>(bool, bool, const GemmJob*, const GemmTile*,    \
                               int, void*, unsigned*, hipStream_t);           \
  template void launch_potrf<T>(const PotrfJob*, int, int*, hipStream_t);      \
  template void launch_trsm<T>(const TrsmJob*, const RowTile*, int,            \
                               hipStream_t);                                   \
  template void launch_copy<T>(const CopyJob*, const RowTile*, int,            \
                               hipStream_t);                                   \
  template void launch_shift<T>(const ShiftJob*, int, hipStream_t);            \
  template void launch_eye<T>(const EyeJob*, const RowTile*, int, hipStream_t); \
  template void launch_check<T>(const CheckJob*, int, int*, hipStream_t);       \
  template void launch_assemble<T>(T*, const FrontDev*, const AsmTask*, int,   \
                                   const int*, const int*, const int*,         \
                                   const int64_t*, const int64_t*,             \
                                   const double*, hipStream_t);

RSC_INST(double)
RSC_INST(float)

}  // namespace dev
}  // namespace rsc
