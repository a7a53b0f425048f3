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

template <typename T>
void launch_forward(const T* factor, const SolveNode* nodes, const int* level_nodes,
                    int count, const SolveHNode* htree, const int* hops,
                    const int* children, const int* relmap, double* w, int n, int k,
                    double* fv, int64_t fv_stride, hipStream_t s);
template <typename T>
void launch_backward(const T* factor, const SolveNode* nodes, const int* level_nodes,
                     int count, const SolveHNode* htree, const int* hops,
                     const int* rows, double* w, int n, int k, double* fv,
                     int64_t fv_stride, hipStream_t s);

// CSR (= CSC of the symmetric matrix) y = A x for k columns.
void launch_spmv(int n, const int* p, const int* idx, const double* val,
                 const double* x, int ldx, double* y, int ldy, int k,
                 const int* active, hipStream_t s, int lanes_per_row);

}  // namespace dev
}  // namespace rsc
