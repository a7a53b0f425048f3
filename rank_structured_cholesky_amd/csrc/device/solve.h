// Device kernels of the preconditioner apply and of PCG.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "factor.h"

namespace rsc {
namespace dev {

// w[i + j*n] = b[perm[i] + j*ldb]
void launch_gather(const double* b, int ldb, double* w, const int* perm, int n,
                   int k, hipStream_t s);
// x[perm[i] + j*ldx] = w[i + j*n]
void launch_scatter(const double* w, double* x, int ldx, const int* perm, int n,
                    int k, hipStream_t s);

// One matrix-vector task of a solve phase.  ROW: y[i] (op)= sum_j M[i,j] x[j]
// for i < rows (one thread per row).  COL: y[j] (op)= sum_i M[i,j] x[i] for
// j < cols (waves reduce over rows).  M is the factor slice at `mat`
// (leading dimension ld); x / y live in w (stride n), the front vectors fv
// (stride fvs) or the rank-sized buffers tv (stride tvs): sel 0 / 1 / 2.
// A COL task over many rows is split into row chunks that write partial sums
// (part >= 0: slot of 32 partials in the partial buffer); a reduce task then
// adds a group's partials in fixed order and applies op to y.
struct GvTask {
  int64_t mat;
  int ld, rows, cols;
  int kind;  // 0 ROW, 1 COL, 2 REDUCE (part = first slot, rows = slot count)
  int op;    // 0 set, 1 subtract
  int xsel, ysel;
  int part;
  int64_t xoff, yoff;
};

// Per-supernode vector gathers of a solve level.
struct VecNode {
  int first, ncols, R, nchild;
  int child_off, pad;
  int64_t fv_off;    // front vector offset (per RHS column)
  int64_t rows_off;  // row list offset
  int64_t relmap_off;
};

struct VecBufs {
  double* w;
  double* fv;
  double* tv;
  double* part;
  int64_t n, fvs, tvs, ps;
};

// forward: fv_c = w[cols], fv_R = 0, plus children's fv_R via relmap
void launch_fwd_assemble(const VecNode* nodes, const int* level_nodes, int count,
                         const int* children, const int* relmap, VecBufs b, int k,
                         int max_front, hipStream_t s);
// backward: fv_c = w[cols], fv_R[a] = w[rows[a]]
void launch_bwd_gather(const VecNode* nodes, const int* level_nodes, int count,
                       const int* rows, VecBufs b, int k, int max_front, hipStream_t s);
template <typename T>
void launch_gemv_tasks(const T* factor, const GvTask* tasks, int count, VecBufs b,
                       int k, bool big, hipStream_t s);

// CSR (= CSC of the symmetric matrix) y = A x for k columns.
void launch_spmv(int n, const int* p, const int* idx, const double* val,
                 const double* x, int ldx, double* y, int ldy, int k,
                 const int* active, hipStream_t s, int lanes_per_row);

}  // namespace dev
}  // namespace rsc
