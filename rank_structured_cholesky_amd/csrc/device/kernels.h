// Batched dense kernels of the device factor and solve (gfx950).  All
// matrices are column-major.  Each launch processes a device array of jobs
// of varying size; the host precomputes the tile lists.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include <cstdint>

namespace rsc {
namespace dev {

constexpr int TILE = 64;  // GEMM output tile and POTRF/TRSM block width

// C(m x n) = beta*C + alpha*op(A)*op(B); op(A) is m x k, op(B) is k x n.
struct GemmJob {
  const void* A;
  const void* B;
  void* C;
  int m, n, k, lda, ldb, ldc;
  double alpha, beta;
};
// One workgroup's share of a GEMM: output tile (tm, tn) over k in [k0, k1).
// ns > 1 splits the k range of one output tile over ns workgroups: each
// writes its partial tile to partials[part + si], and the last to finish
// (counted in counters[cnt]) sums the ns partials in index order and writes
// C, so the result does not depend on which workgroup finishes last.
struct GemmTile {
  int job, tm, tn;
  int k0, k1;
  int ns, si, part, cnt;
};
// Tiles of jobs[first_job..] for one launch, k split where the launch has
// too few output tiles to fill the device.  Counter indices continue from
// *ncounters; partial slots (TILE*TILE elements each) start at 0 per
// launch, and their count is returned.
int plan_gemm_tiles(const GemmJob* jobs, int njobs, bool lower,
                    std::vector<GemmTile>& tiles, int* ncounters);

// In-place Cholesky of an n x n (n <= TILE) lower block.  tolerant:
//   0  a non-positive pivot is reported (status[0] = 1, status[1] = node)
//   1  (Gram matrices of the thin QR) it is replaced by u * max(tile diagonal)
//   2  (FP32 factor) it is replaced by 1e-5 * max(tile diagonal) and counted
//      in status[2]; L keeps a positive diagonal, so L L^T stays SPD
//   3  (Gram matrices of the unshifted thin-QR passes, whose columns have
//      norm <= 1) a pivot <= 1e-13 marks a dependent column: it is replaced
//      by a huge value, so that column of L and of the orthonormalized
//      matrix come out (numerically) zero: the basis drops the direction
//      instead of amplifying rounding noise into it
struct PotrfJob {
  void* A;
  int n, lda, node, tolerant;
};

// B(m x nb) <- B * L^{-T}, L nb x nb lower (nb <= TILE).
struct TrsmJob {
  const void* L;
  void* B;
  int m, nb, ldl, ldb;
};
struct RowTile {
  int job, row0;
};

// dst(m x n) = src(m x n)
struct CopyJob {
  const void* src;
  void* dst;
  int m, n, lds, ldd;
};

// A(n x n, lda) = I
struct EyeJob {
  void* A;
  int n, lda;
};

// dst = src converted between FP32 and FP64 (to64 != 0: float -> double)
struct ConvJob {
  const void* src;
  void* dst;
  int64_t count;
  int to64, pad;
};
void launch_convjobs(const ConvJob* jobs, int njobs, hipStream_t s);

// Debug scan: status[3] = min(tag) over jobs whose rows x cols block
// (leading dimension ld) holds a non-finite value.
struct CheckJob {
  const void* p;
  int rows, cols, ld, tag;
};

// W(r x r) += s*I with s = 11 (m r + r (r+1)) u trace(W)
struct ShiftJob {
  void* W;
  int r, ld, m, pad;
};

// per-node front geometry for the assembly kernel
struct FrontDev {
  int64_t panel;    // element offset of the (c+R) x c panel in the arena
  int64_t update;   // element offset of the R x R update matrix (-1 if R==0)
  int ncols, R;
  int first;        // first permuted column
  int child_off, nchild;
  int64_t relmap_off;  // offset of this node's relmap (R entries)
};
struct AsmTask {
  int node, col0;
};

template <typename T>
void launch_gemm(bool ta, bool tb, const GemmJob* jobs, const GemmTile* tiles,
                 int ntiles, void* partials, unsigned* counters, hipStream_t s);
// FP32 operands and result, BF16 products with FP32 accumulation
void launch_gemm_bf16(bool ta, bool tb, const GemmJob* jobs, const GemmTile* tiles,
                      int ntiles, void* partials, unsigned* counters, hipStream_t s);
template <typename T>
void launch_potrf(const PotrfJob* jobs, int njobs, int* status, hipStream_t s);
template <typename T>
void launch_trsm(const TrsmJob* jobs, const RowTile* tiles, int ntiles,
                 hipStream_t s);
template <typename T>
void launch_copy(const CopyJob* jobs, const RowTile* tiles, int ntiles,
                 hipStream_t s);
template <typename T>
void launch_eye(const EyeJob* jobs, const RowTile* tiles, int ntiles, hipStream_t s);
template <typename T>
void launch_check(const CheckJob* jobs, int njobs, int* status, hipStream_t s);
template <typename T>
void launch_shift(const ShiftJob* jobs, int njobs, hipStream_t s);
// Assemble fronts: zero, scatter A (values converted to T), extend-add the
// children's update matrices.
template <typename T>
void launch_assemble(T* arena, const FrontDev* fronts, const AsmTask* tasks,
                     int ntasks, const int* children, const int* relmap,
                     const int* lowp, const int64_t* a_dst, const int64_t* a_src,
                     const double* Ax, hipStream_t s);

}  // namespace dev
}  // namespace rsc
