// Device PCG driver and the PCG operator (full symmetric CSC = CSR).
#pragma once

#include "factor.h"

namespace rsc {

struct DevCsr {
  int n = 0;
  int64_t nnz = 0;
  int lanes = 8;
  DBuf p, i, x;
  void upload(const Csc& A, hipStream_t s);
  void update_values(const double* vals, hipStream_t s);
  void spmv(const double* X, int ldx, double* Y, int ldy, int k, const int* active,
            hipStream_t s) const;
};

// k PCG recurrences (device pointers B, X); res has k entries.
void pcg_solve_device(FactorEngine& M, const DevCsr& A, int k, const double* B,
                      int ldb, double* X, int ldx, double eps, int max_iters,
                      rsc_pcg_result* res, hipStream_t s);

}  // namespace rsc
