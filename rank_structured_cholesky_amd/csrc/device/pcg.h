// Device PCG driver and the PCG operator (full symmetric CSC = CSR).
#pragma once

#include "factor.h"

namespace rsc {

struct DevCsr {
  int n = 0;
  int64_t nnz = 0;
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

// Column slots of the request server: a slot is FREE, admitted (INIT),
// iterating (RUN) or converged (DONE) until the host retires it.
enum SlotState { SLOT_FREE = 0, SLOT_INIT = 1, SLOT_RUN = 2, SLOT_DONE = 3 };
constexpr int kSlotList = 64;
struct SlotList {  // passed by value to the kernels
  int n = 0;
  int s[kSlotList];
};
struct SlotResult {  // the host finishes true_relres = sqrt(rr) / bnorm
  int iterations, pad;
  double recurrence_relres, rr, bnorm;
};

// Device state of kcap column slots of one factor: b, x and the PCG vectors
// (n x kcap each, column j = slot j), per-slot scalars, states and results.
struct ServeBatch {
  int n = 0, kcap = 0;
  DBuf b_, x, r, d, q, s, part1, part2, stb, state, run, busy, anyb, res;
  void alloc(int n, int kcap);
  // slots of L (b already copied in): x = 0, r = b, ||b||, state INIT
  void admit(const SlotList& L, hipStream_t st);
  // one iteration over slots [0, kact): RUN slots advance, INIT slots take
  // their first preconditioner apply; states move on in k_serve_commit
  void step(FactorEngine& M, const DevCsr& A, int kact, double eps, int max_iters,
            bool admitted, hipStream_t st);
  // DONE slots of L: ||b - A x||^2 and the PCG results into res; state FREE
  void retire(const DevCsr& A, const SlotList& L, hipStream_t st);
};

// k PCG recurrences (device pointers B, X); res has k entries.
void pcg_solve_device(FactorEngine& M, const DevCsr& A, PcgWork& w, int k, const double* B,
                      int ldb, double* X, int ldx, double eps, int max_iters,
                      rsc_pcg_result* res, hipStream_t s);

}  // namespace rsc
