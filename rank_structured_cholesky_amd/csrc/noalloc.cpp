// rsc_apply_noalloc: the preconditioner apply of rsc_apply with its per-call
// scratch in a fixed device array instead of library allocations, for
// callers that cannot allocate (code under graph capture, tight serving
// loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
//
// The scratch is the file-scope array `scratch` below (kNoAllocScratchBytes
// per device, held until the process ends).  Its device address is taken
// once per device with hipGetSymbolAddress and kept; host code uses only
// that address, never the array's name.  Every call on a device shares the
// device's array, so calls that overlap on one device race: the caller
// orders them (one stream, or events between streams).  The array needs no
// clearing: the apply writes every scratch entry it reads within the call.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "capi_internal.h"

namespace {

constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;  // 64 MiB per device
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};  // device address of `scratch` per device

// The device address of this device's array (taken on the first call on the
// device, with that device current).
double* scratch_ptr(int dev) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw rsc::DeviceError(std::string("hipGetSymbolAddress(scratch): ") +
                             hipGetErrorString(e));
    g_ptr[dev] = (double*)p;
  }
  return g_ptr[dev];
}

rsc_status fail(rsc_status st, const std::string& msg) {
  rsc::last_error() = msg;
  return st;
}

}  // namespace

extern "C" {

rsc_status rsc_apply_noalloc(rsc_factor* f, int nrhs, const double* B, int ldb, double* X,
                             int ldx, void* stream) {
  if (!f || !B || !X || nrhs < 0 || ldb < f->plan.n || ldx < f->plan.n)
    return fail(RSC_BAD_INPUT, "bad apply arguments");
  if (nrhs == 0) return RSC_OK;
  try {
    const int64_t need = f->eng->apply_scratch_doubles(nrhs) * (int64_t)sizeof(double);
    if (need > (int64_t)kNoAllocScratchBytes)
      return fail(RSC_BAD_INPUT, "rsc_apply_noalloc: " + std::to_string(need) +
                                     " bytes of scratch needed, the fixed array holds " +
                                     std::to_string(kNoAllocScratchBytes) +
                                     " (use rsc_apply, or fewer right-hand sides)");
    int dev = 0;
    RSC_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(RSC_DEVICE_ERROR, "device index out of range");
    double* sp = scratch_ptr(dev);
    hipStream_t s = stream ? (hipStream_t)stream : f->stream;
    f->eng->apply_into(nrhs, B, ldb, X, ldx, s, sp);
    RSC_HIP(hipGetLastError());
    return RSC_OK;
  } catch (const rsc::DeviceError& e) {
    return fail(RSC_DEVICE_ERROR, e.what());
  } catch (const std::exception& e) {
    return fail(RSC_BAD_INPUT, e.what());
  }
}

}  // extern "C"
