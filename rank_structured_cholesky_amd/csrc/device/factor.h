// Device numeric factor and solve of one matrix.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "../host/rsc_host.h"
#include "kernels.h"

namespace rsc {

#define RSC_HIP(call)                                                     \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess)                                                 \
      throw ::rsc::DeviceError(std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct IndefiniteError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Owning device buffer.
// A device allocation that failed (an out-of-memory condition: RSC_OOM).
struct DeviceOom : std::bad_alloc {
  std::string msg;
  explicit DeviceOom(size_t bytes)
      : msg("device allocation of " + std::to_string(bytes) + " bytes failed") {}
  const char* what() const noexcept override { return msg.c_str(); }
};

struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t b) {
    reset();
    if (b == 0) b = 8;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      throw DeviceOom(b);
    }
    bytes = b;
  }
  template <typename U>
  void upload(const std::vector<U>& v, hipStream_t s) {
    alloc(v.size() * sizeof(U));
    if (!v.empty())
      RSC_HIP(hipMemcpyAsync(p, v.data(), v.size() * sizeof(U),
                             hipMemcpyHostToDevice, s));
  }
  template <typename U>
  U* as() const { return (U*)p; }
};

// Solve-side description of one supernode (device array).
struct SolveNode {
  int first, ncols, R, compress;
  int64_t panel;        // uncompressed: factor offset of the panel
  int64_t V, U;         // compressed off-diagonal pair (factor offsets)
  int rank;
  int htree;            // compressed: offset of its tree in the tree array
  int hroot;            // index of the root inside its tree
  int child_off, nchild;
  int64_t relmap_off;
  int64_t rows_off;     // offset of the row index list
  int64_t fv_off;       // front-vector scratch offset (per RHS column)
  int hops_off, nhops;  // compressed: in-order HODLR solve ops
};
struct SolveHNode {
  int r0, r1, left, right;  // left/right relative to the tree base (-1 leaf)
  int rank, pad;
  int64_t leaf, V, U;       // factor offsets
};

class FactorEngine {
 public:
  virtual ~FactorEngine() = default;
  virtual void numeric(const double* Ax_host) = 0;          // may throw IndefiniteError
  // X = M^{-1} B (device pointers).  live: optional device flag; when it
  // reads 0 at run time the apply does nothing (converged PCG columns).
  // multi_min_k: matrix-vector tasks use the sixteen-right-hand-side matrix
  // core bodies for k > 1, and for k = 1 when multi_min_k <= 1 (so a
  // column's result does not depend on how many others share the apply)
  virtual void apply(int k, const double* B, int ldb, double* X, int ldx, hipStream_t s,
                     const int* live = nullptr, int multi_min_k = 2) = 0;
  // Doubles of per-call scratch an apply with k right-hand sides uses (the
  // permuted vectors w and z, front vectors, rank vectors, partial sums).
  virtual int64_t apply_scratch_doubles(int k) const = 0;
  // apply() with that scratch supplied by the caller (device memory of at
  // least apply_scratch_doubles(k) doubles, 256-byte aligned; its contents
  // need no initialization): no allocation of any kind.
  virtual void apply_into(int k, const double* B, int ldb, double* X, int ldx, hipStream_t s,
                          double* scratch) = 0;
  // Synchronizes s and throws DeviceError if an apply on it failed.
  virtual void check_apply(hipStream_t s) = 0;
  // Device address of the apply's status words (two ints, both 0 while every
  // apply so far succeeded), for callers that read it asynchronously.
  virtual const int* apply_status_device() const = 0;
  // Throws DeviceError for status words read from apply_status_device().
  virtual void raise_apply_status(const int* words) = 0;
  virtual int64_t factor_bytes() const = 0;
  virtual int64_t workspace_bytes() const = 0;
  virtual int pivot_replacements() const = 0;
  virtual void release_workspace() = 0;
};

std::unique_ptr<FactorEngine> make_engine(const Plan& plan, const Csc& A,
                                          hipStream_t stream);

}  // namespace rsc
