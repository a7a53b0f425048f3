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

// Device and pinned host buffers of the PCG driver, kept between solves
// (grown when a solve needs more right-hand sides).
struct PcgWork {
  int n = 0, k = 0;
  DBuf r, d, q, sv, part1, part2, stb, flags, anyb;
  int* host_any = nullptr;           // pinned, one slot per polled iteration parity
  hipEvent_t ev[2] = {nullptr, nullptr};
  PcgWork() = default;
  PcgWork(const PcgWork&) = delete;
  PcgWork& operator=(const PcgWork&) = delete;
  ~PcgWork();
  void ensure(int n, int k);
};

// k PCG recurrences (device pointers B, X); res has k entries.
void pcg_solve_device(FactorEngine& M, const DevCsr& A, PcgWork& w, int k, const double* B,
                      int ldb, double* X, int ldx, double eps, int max_iters,
                      rsc_pcg_result* res, hipStream_t s);

}  // namespace rsc
