// Preconditioned conjugate gradients on the device, k right-hand sides
// advanced together (PCG_SOLVER::solvePCG, src/rschol/linearalgebra/
// PCG_SOLVER.cpp:226-363): every column keeps its own alpha, beta, delta and
// stop test; one SpMV with k vectors and one preconditioner apply with k
// right-hand sides per iteration.  All vector arithmetic is FP64 and every
// reduction is a fixed-order two-stage sum (run-to-run identical).
#include <chrono>
#include <cmath>
#include <vector>

#include "pcg.h"
#include "solve.h"

namespace rsc {
namespace dev {

constexpr int RB = 256;    // reduction block
constexpr int RMAX = 512;  // partial sums per column

struct PcgCol {
  double delta, alpha, beta, bnorm, maxr, rr;
  int iters, active, next_active, pad;
};

__device__ inline double block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int w = RB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// partial[col*RMAX + b] = sum over this block's slice of x.*y
__global__ __launch_bounds__(RB) void k_dot(int n, const double* __restrict__ x, int ldx,
                                            const double* __restrict__ y, int ldy,
                                            double* __restrict__ partial) {
  __shared__ double sh[RB];
  const int col = blockIdx.y;
  const double* xc = x + (int64_t)col * ldx;
  const double* yc = y + (int64_t)col * ldy;
  double acc = 0;
  for (int64_t i = blockIdx.x * (int64_t)RB + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * RB)
    acc += xc[i] * yc[i];
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[col * RMAX + blockIdx.x] = s;
}

__device__ double sum_partials(const double* p, int nb, double* sh) {
  double acc = 0;
  for (int b = threadIdx.x; b < nb; b += RB) acc += p[b];
  return block_sum(acc, sh);
}

// phase 0: bnorm; phase 1: delta = r.d (initial); phase 2: alpha from d.q;
// phase 3: delta/beta/maxr from r.s and r.r
__global__ __launch_bounds__(RB) void k_scalars(int phase, const double* __restrict__ p1,
                                                const double* __restrict__ p2, int nb,
                                                PcgCol* __restrict__ st, double eps,
                                                int max_iters) {
  __shared__ double sh[RB];
  const int col = blockIdx.x;
  double a = sum_partials(p1 + col * RMAX, nb, sh);
  double b = p2 ? sum_partials(p2 + col * RMAX, nb, sh) : 0.0;
  if (threadIdx.x != 0) return;
  PcgCol& c = st[col];
  if (phase == 0) {
    c.bnorm = sqrt(a);
  } else if (phase == 1) {
    c.delta = a;
    c.maxr = 2.0 * eps;
    c.iters = 0;
    c.active = (max_iters > 0 && c.maxr > eps) ? 1 : 0;
    c.next_active = c.active;
  } else if (phase == 2) {
    if (!c.active) return;
    c.alpha = (fabs(a) > 0.0) ? c.delta / a : a;
  } else {
    if (!c.active) return;
    double old = c.delta;
    c.delta = a;
    c.beta = c.delta / old;
    c.rr = b;
    c.maxr = sqrt(b) / c.bnorm;
    c.iters += 1;
    c.next_active = (c.iters < max_iters && c.maxr > eps) ? 1 : 0;
  }
}

// r = b - r (r holds A x on entry)
__global__ void k_residual(int n, int k, const double* __restrict__ b, int ldb,
                           double* __restrict__ r) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * k) return;
  int i = (int)(t % n), j = (int)(t / n);
  r[t] = b[i + (int64_t)j * ldb] - r[t];
}

// x += alpha d ; r -= alpha q (active columns) with the partial sums of
// r.r: the grid and row order of k_dot (nb x k blocks of RB threads, rows
// strided by nb RB), so the partials are those k_dot(r, r) gives after the
// update.  Columns with mask[col] == 0 (when mask is set) write no partials.
__global__ __launch_bounds__(RB) void k_update_xr_rr(int n, const PcgCol* __restrict__ st,
                                                     double* __restrict__ x, int ldx,
                                                     double* __restrict__ r,
                                                     const double* __restrict__ d,
                                                     const double* __restrict__ q,
                                                     double* __restrict__ partial,
                                                     const int* __restrict__ mask) {
  __shared__ double sh[RB];
  const int col = blockIdx.y;
  const bool act = st[col].active != 0;
  const bool dot = !mask || mask[col];
  if (!act && !dot) return;
  const double al = st[col].alpha;
  double acc = 0;
  for (int64_t i = blockIdx.x * (int64_t)RB + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * RB) {
    const int64_t o = i + (int64_t)col * n;
    double rv = r[o];
    if (act) {
      x[i + (int64_t)col * ldx] += al * d[o];
      rv -= al * q[o];
      r[o] = rv;
    }
    acc += rv * rv;
  }
  if (!dot) return;
  const double sres = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[col * RMAX + blockIdx.x] = sres;
}

// d = beta d + s
__global__ void k_update_d(int n, int k, const PcgCol* __restrict__ st,
                           double* __restrict__ d, const double* __restrict__ s) {
  const int col = blockIdx.y;
  if (!st[col].active) return;
  const double be = st[col].beta;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t o = i + (int64_t)col * n;
    d[o] = be * d[o] + s[o];
  }
}

__global__ void k_commit(int k, PcgCol* __restrict__ st, int* __restrict__ flags,
                         int* __restrict__ any) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  for (int c = threadIdx.x; c < k; c += blockDim.x) {
    st[c].active = st[c].next_active;
    flags[c] = st[c].active;
    if (st[c].active) atomicAdd(&cnt, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) *any = cnt;
}

__global__ void k_flags(int k, const PcgCol* __restrict__ st, int* __restrict__ flags,
                        int* __restrict__ any) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  for (int c = threadIdx.x; c < k; c += blockDim.x) {
    flags[c] = st[c].active;
    if (st[c].active) atomicAdd(&cnt, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) *any = cnt;
}


// ---- continuous batching (request server) ---------------------------------
// A fixed pool of column slots advanced by one PCG iteration per step; a
// slot's column is admitted (INIT), iterates (RUN), converges (DONE) and is
// retired by the host, independently of the other slots.  Every per-column
// operation below repeats the arithmetic of the kernels above in the same
// order, so a column served in any mix gives the bits of a solo solve.

__global__ __launch_bounds__(RB) void k_dot_masked(int n, const double* __restrict__ x,
                                                   int ldx, const double* __restrict__ y,
                                                   int ldy, double* __restrict__ partial,
                                                   const int* __restrict__ flags) {
  __shared__ double sh[RB];
  const int col = blockIdx.y;
  if (!flags[col]) return;
  const double* xc = x + (int64_t)col * ldx;
  const double* yc = y + (int64_t)col * ldy;
  double acc = 0;
  for (int64_t i = blockIdx.x * (int64_t)RB + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * RB)
    acc += xc[i] * yc[i];
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[col * RMAX + blockIdx.x] = s;
}

__global__ __launch_bounds__(RB) void k_dot_list(int n, const double* __restrict__ x,
                                                 int ldx, const double* __restrict__ y,
                                                 int ldy, double* __restrict__ partial,
                                                 SlotList L) {
  __shared__ double sh[RB];
  const int col = L.s[blockIdx.y];
  const double* xc = x + (int64_t)col * ldx;
  const double* yc = y + (int64_t)col * ldy;
  double acc = 0;
  for (int64_t i = blockIdx.x * (int64_t)RB + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * RB)
    acc += xc[i] * yc[i];
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[col * RMAX + blockIdx.x] = s;
}

// admission: x = 0, r = b, ||b|| (k_scalars phase 0), state INIT
__global__ void k_serve_admit(int n, SlotList L, const double* __restrict__ B,
                              double* __restrict__ X, double* __restrict__ R) {
  const int col = L.s[blockIdx.y];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = i + (int64_t)col * n;
    X[o] = 0.0;
    R[o] = B[o] - 0.0;  // r = b - A x with x = 0
  }
}

__global__ __launch_bounds__(RB) void k_serve_bnorm(SlotList L, const double* __restrict__ p1,
                                                    int nb, PcgCol* __restrict__ st,
                                                    int* __restrict__ state,
                                                    int* __restrict__ any) {
  __shared__ double sh[RB];
  const int col = L.s[blockIdx.x];
  double a = sum_partials(p1 + col * RMAX, nb, sh);
  if (threadIdx.x != 0) return;
  PcgCol c{};
  c.bnorm = sqrt(a);
  st[col] = c;
  state[col] = SLOT_INIT;
  *any = 1;
}

// INIT columns: phase 1 on r.s (s = M r); RUN columns: phase 3
__global__ __launch_bounds__(RB) void k_serve_scalars(const double* __restrict__ p1,
                                                      const double* __restrict__ p2, int nb,
                                                      PcgCol* __restrict__ st,
                                                      const int* __restrict__ state,
                                                      double eps, int max_iters) {
  __shared__ double sh[RB];
  const int col = blockIdx.x;
  const int sc = state[col];
  if (sc != SLOT_INIT && sc != SLOT_RUN) return;
  double a = sum_partials(p1 + col * RMAX, nb, sh);
  double b = sum_partials(p2 + col * RMAX, nb, sh);
  if (threadIdx.x != 0) return;
  PcgCol& c = st[col];
  if (sc == SLOT_INIT) {
    c.delta = a;
    c.maxr = 2.0 * eps;
    c.iters = 0;
    c.active = (max_iters > 0 && c.maxr > eps) ? 1 : 0;
    c.next_active = c.active;
  } else {
    if (!c.active) return;
    double old = c.delta;
    c.delta = a;
    c.beta = c.delta / old;
    c.rr = b;
    c.maxr = sqrt(b) / c.bnorm;
    c.iters += 1;
    c.next_active = (c.iters < max_iters && c.maxr > eps) ? 1 : 0;
  }
}

// RUN: d = beta d + s; INIT: d = s
__global__ void k_serve_update_d(int n, const PcgCol* __restrict__ st,
                                 const int* __restrict__ state, double* __restrict__ d,
                                 const double* __restrict__ s) {
  const int col = blockIdx.y;
  const int sc = state[col];
  if (sc == SLOT_INIT) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
      d[i + (int64_t)col * n] = s[i + (int64_t)col * n];
    return;
  }
  if (sc != SLOT_RUN || !st[col].active) return;
  const double be = st[col].beta;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t o = i + (int64_t)col * n;
    d[o] = be * d[o] + s[o];
  }
}

// state transitions after a step; run = RUN mask (SpMV, x/r update), busy =
// INIT or RUN (apply, dots); any = number of busy slots
__global__ void k_serve_commit(int k, PcgCol* __restrict__ st, int* __restrict__ state,
                               int* __restrict__ run, int* __restrict__ busy,
                               int* __restrict__ any) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  for (int c = threadIdx.x; c < k; c += blockDim.x) {
    int sc = state[c];
    if (sc == SLOT_INIT) {
      st[c].active = st[c].next_active;
      sc = st[c].active ? SLOT_RUN : SLOT_DONE;
    } else if (sc == SLOT_RUN) {
      st[c].active = st[c].next_active;
      if (!st[c].active) sc = SLOT_DONE;
    }
    state[c] = sc;
    run[c] = sc == SLOT_RUN;
    busy[c] = sc == SLOT_RUN || sc == SLOT_INIT;
    if (sc == SLOT_RUN) atomicAdd(&cnt, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) *any = cnt;
}

// masks before the first step after admissions (INIT slots are busy)
__global__ void k_serve_masks(int k, PcgCol* __restrict__ st, const int* __restrict__ state,
                              int* __restrict__ run, int* __restrict__ busy) {
  for (int c = threadIdx.x; c < k; c += blockDim.x) {
    const int sc = state[c];
    run[c] = sc == SLOT_RUN;
    busy[c] = sc == SLOT_RUN || sc == SLOT_INIT;
    if (sc == SLOT_INIT) st[c].active = 0;  // no SpMV / x, r update before the first apply
  }
}

// retirement: r = b - A x (A x already in r), then the partials of r.r
// summed sequentially in block order, as the solo driver does on the host
__global__ void k_serve_residual(int n, SlotList L, const double* __restrict__ B,
                                 double* __restrict__ R) {
  const int col = L.s[blockIdx.y];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = i + (int64_t)col * n;
    R[o] = B[o] - R[o];
  }
}

__global__ void k_serve_result(SlotList L, const double* __restrict__ p1, int nb,
                               PcgCol* __restrict__ st, int* __restrict__ state,
                               SlotResult* __restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= L.n) return;
  const int col = L.s[q];
  double rr = 0;
  for (int b = 0; b < nb; ++b) rr += p1[col * RMAX + b];
  out[col].iterations = st[col].iters;
  out[col].recurrence_relres = st[col].maxr;
  out[col].rr = rr;
  out[col].bnorm = st[col].bnorm;
  state[col] = SLOT_FREE;
}

}  // namespace dev

void DevCsr::upload(const Csc& A, hipStream_t s) {
  n = A.n;
  nnz = A.nnz();
  p.upload(A.p, s);
  i.upload(A.i, s);
  x.upload(A.x, s);
}

void DevCsr::update_values(const double* vals, hipStream_t s) {
  RSC_HIP(hipMemcpyAsync(x.p, vals, nnz * sizeof(double), hipMemcpyHostToDevice, s));
}

void DevCsr::spmv(const double* X, int ldx, double* Y, int ldy, int k,
                  const int* active, hipStream_t s) const {
  dev::launch_spmv(n, p.as<int>(), i.as<int>(), x.as<double>(), X, ldx, Y, ldy, k,
                   active, s);
}

PcgWork::~PcgWork() {
  if (host_any) (void)hipHostFree(host_any);
  for (hipEvent_t& e : ev)
    if (e) (void)hipEventDestroy(e);
}

void PcgWork::ensure(int n_, int k_) {
  using namespace dev;
  if (!host_any) {
    RSC_HIP(hipHostMalloc((void**)&host_any, 2 * sizeof(int), hipHostMallocDefault));
    for (hipEvent_t& e : ev) RSC_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  if (n_ == n && k_ <= k) return;
  const size_t vb = (size_t)n_ * k_ * sizeof(double);
  r.alloc(vb);
  d.alloc(vb);
  q.alloc(vb);
  sv.alloc(vb);
  part1.alloc((size_t)k_ * RMAX * sizeof(double));
  part2.alloc((size_t)k_ * RMAX * sizeof(double));
  stb.alloc((size_t)k_ * sizeof(PcgCol));
  flags.alloc((size_t)k_ * sizeof(int));
  anyb.alloc(sizeof(int));
  n = n_;
  k = k_;
}

// The host checks convergence one iteration behind: iteration i+1 is queued
// before the flag of iteration i is read, so the device never idles on the
// check.  A queued iteration whose columns have all converged does no work
// (every kernel of it returns on the flags).
void pcg_solve_device(FactorEngine& M, const DevCsr& A, PcgWork& w, int k, const double* B,
                      int ldb, double* X, int ldx, double eps, int max_iters,
                      rsc_pcg_result* res, hipStream_t s) {
  using namespace dev;
  auto t0 = std::chrono::steady_clock::now();
  const int n = A.n;
  w.ensure(n, k);
  RSC_HIP(hipMemsetAsync(w.stb.p, 0, (size_t)k * sizeof(PcgCol), s));
  PcgCol* st = w.stb.as<PcgCol>();
  double *R = w.r.as<double>(), *D = w.d.as<double>(), *Q = w.q.as<double>(),
         *S = w.sv.as<double>();
  int* any = w.anyb.as<int>();
  const int nb = (int)std::min<int64_t>(RMAX, ((int64_t)n + RB * 4 - 1) / (RB * 4));
  const dim3 dgrid(nb, k);
  const int ublocks = (int)std::min<int64_t>(2048, ((int64_t)n + 255) / 256);
  const dim3 ugrid(ublocks, k);
  const unsigned eb = (unsigned)(((int64_t)n * k + 255) / 256);
  // ||b||
  k_dot<<<dgrid, RB, 0, s>>>(n, B, ldb, B, ldb, w.part1.as<double>());
  k_scalars<<<k, RB, 0, s>>>(0, w.part1.as<double>(), nullptr, nb, st, eps, max_iters);
  // r = b - A x ; d = M^{-1} r ; delta = r.d
  A.spmv(X, ldx, R, n, k, nullptr, s);
  k_residual<<<eb, 256, 0, s>>>(n, k, B, ldb, R);
  M.apply(k, R, n, D, n, s);
  k_dot<<<dgrid, RB, 0, s>>>(n, R, n, D, n, w.part1.as<double>());
  k_scalars<<<k, RB, 0, s>>>(1, w.part1.as<double>(), nullptr, nb, st, eps, max_iters);
  k_flags<<<1, 256, 0, s>>>(k, st, w.flags.as<int>(), any);
  int it = 0;
  while (it < max_iters) {
    A.spmv(D, n, Q, n, k, w.flags.as<int>(), s);
    k_dot<<<dgrid, RB, 0, s>>>(n, D, n, Q, n, w.part1.as<double>());
    k_scalars<<<k, RB, 0, s>>>(2, w.part1.as<double>(), nullptr, nb, st, eps, max_iters);
    // the update with the partials of r.r (r does not change in the apply)
    k_update_xr_rr<<<dgrid, RB, 0, s>>>(n, st, X, ldx, R, D, Q, w.part2.as<double>(), nullptr);
    M.apply(k, R, n, S, n, s, any);
    k_dot<<<dgrid, RB, 0, s>>>(n, R, n, S, n, w.part1.as<double>());
    k_scalars<<<k, RB, 0, s>>>(3, w.part1.as<double>(), w.part2.as<double>(), nb, st, eps,
                               max_iters);
    k_update_d<<<ugrid, 256, 0, s>>>(n, k, st, D, S);
    k_commit<<<1, 256, 0, s>>>(k, st, w.flags.as<int>(), any);
    RSC_HIP(hipMemcpyAsync(w.host_any + (it & 1), any, sizeof(int), hipMemcpyDeviceToHost, s));
    RSC_HIP(hipEventRecord(w.ev[it & 1], s));
    ++it;
    if (it >= 2) {
      RSC_HIP(hipEventSynchronize(w.ev[it & 1]));  // iteration it - 2 (0-based)
      if (w.host_any[it & 1] == 0) break;
    }
  }
  // true residual
  A.spmv(X, ldx, R, n, k, nullptr, s);
  k_residual<<<eb, 256, 0, s>>>(n, k, B, ldb, R);
  k_dot<<<dgrid, RB, 0, s>>>(n, R, n, R, n, w.part1.as<double>());
  std::vector<PcgCol> hst(k);
  std::vector<double> hp((size_t)k * RMAX);
  RSC_HIP(hipMemcpyAsync(hst.data(), w.stb.p, k * sizeof(PcgCol), hipMemcpyDeviceToHost, s));
  RSC_HIP(hipMemcpyAsync(hp.data(), w.part1.p, hp.size() * sizeof(double),
                         hipMemcpyDeviceToHost, s));
  RSC_HIP(hipStreamSynchronize(s));
  RSC_HIP(hipGetLastError());
  M.check_apply(s);
  double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (res)
    for (int c = 0; c < k; ++c) {
      double rr = 0;
      for (int b = 0; b < nb; ++b) rr += hp[(size_t)c * RMAX + b];
      res[c].iterations = hst[c].iters;
      res[c].recurrence_relres = hst[c].maxr;
      res[c].true_relres = std::sqrt(rr) / hst[c].bnorm;
      res[c].seconds = secs;
    }
}


// ---- ServeBatch ------------------------------------------------------------

void ServeBatch::alloc(int n_, int kcap_) {
  using namespace dev;
  n = n_;
  kcap = kcap_;
  const size_t vb = (size_t)n * kcap * sizeof(double);
  for (DBuf* b : {&b_, &x, &r, &d, &q, &s}) b->alloc(vb);
  part1.alloc((size_t)kcap * RMAX * sizeof(double));
  part2.alloc((size_t)kcap * RMAX * sizeof(double));
  stb.alloc((size_t)kcap * sizeof(PcgCol));
  state.alloc((size_t)kcap * sizeof(int));
  run.alloc((size_t)kcap * sizeof(int));
  busy.alloc((size_t)kcap * sizeof(int));
  anyb.alloc(sizeof(int));
  res.alloc((size_t)kcap * sizeof(SlotResult));
  RSC_HIP(hipMemset(state.p, 0, state.bytes));
  RSC_HIP(hipMemset(run.p, 0, run.bytes));
  RSC_HIP(hipMemset(busy.p, 0, busy.bytes));
  RSC_HIP(hipMemset(anyb.p, 0, anyb.bytes));
  RSC_HIP(hipMemset(stb.p, 0, stb.bytes));
}

static int dot_blocks(int n) {
  using namespace dev;
  return (int)std::min<int64_t>(RMAX, ((int64_t)n + RB * 4 - 1) / (RB * 4));
}

void ServeBatch::admit(const SlotList& L, hipStream_t st) {
  using namespace dev;
  if (L.n <= 0) return;
  const int nb = dot_blocks(n);
  const int ublocks = (int)std::min<int64_t>(2048, ((int64_t)n + 255) / 256);
  k_serve_admit<<<dim3(ublocks, L.n), 256, 0, st>>>(n, L, b_.as<double>(), x.as<double>(),
                                                      r.as<double>());
  k_dot_list<<<dim3(nb, L.n), RB, 0, st>>>(n, b_.as<double>(), n, b_.as<double>(), n,
                                            part1.as<double>(), L);
  k_serve_bnorm<<<L.n, RB, 0, st>>>(L, part1.as<double>(), nb, stb.as<PcgCol>(),
                                     state.as<int>(), anyb.as<int>());
}

void ServeBatch::step(FactorEngine& M, const DevCsr& A, int kact, double eps, int max_iters,
                      bool admitted, hipStream_t st) {
  using namespace dev;
  if (kact <= 0) return;
  const int nb = dot_blocks(n);
  const int ublocks = (int)std::min<int64_t>(2048, ((int64_t)n + 255) / 256);
  PcgCol* c = stb.as<PcgCol>();
  double *X = x.as<double>(), *R = r.as<double>(), *D = d.as<double>(),
         *Q = q.as<double>(), *S = s.as<double>();
  int *RUN = run.as<int>(), *BUSY = busy.as<int>();
  if (admitted) k_serve_masks<<<1, 256, 0, st>>>(kact, c, state.as<int>(), RUN, BUSY);
  // RUN columns: q = A d, alpha, x += alpha d, r -= alpha q (as the solo loop)
  A.spmv(D, n, Q, n, kact, RUN, st);
  k_dot_masked<<<dim3(nb, kact), RB, 0, st>>>(n, D, n, Q, n, part1.as<double>(), RUN);
  k_scalars<<<kact, RB, 0, st>>>(2, part1.as<double>(), nullptr, nb, c, eps, max_iters);
  // (with the partials of r.r for INIT and RUN columns, as the solo loop)
  k_update_xr_rr<<<dim3(nb, kact), RB, 0, st>>>(n, c, X, n, R, D, Q, part2.as<double>(), BUSY);
  // INIT and RUN columns: s = M^{-1} r, r.s
  M.apply(kact, R, n, S, n, st, anyb.as<int>(), 1);
  k_dot_masked<<<dim3(nb, kact), RB, 0, st>>>(n, R, n, S, n, part1.as<double>(), BUSY);
  k_serve_scalars<<<kact, RB, 0, st>>>(part1.as<double>(), part2.as<double>(), nb, c,
                                       state.as<int>(), eps, max_iters);
  k_serve_update_d<<<dim3(ublocks, kact), 256, 0, st>>>(n, c, state.as<int>(), D, S);
  k_serve_commit<<<1, 256, 0, st>>>(kact, c, state.as<int>(), RUN, BUSY, anyb.as<int>());
}

void ServeBatch::retire(const DevCsr& A, const SlotList& L, hipStream_t st) {
  using namespace dev;
  if (L.n <= 0) return;
  const int nb = dot_blocks(n);
  const int ublocks = (int)std::min<int64_t>(2048, ((int64_t)n + 255) / 256);
  for (int q0 = 0; q0 < L.n; ++q0) {
    const int col = L.s[q0];
    A.spmv(x.as<double>() + (size_t)col * n, n, r.as<double>() + (size_t)col * n, n, 1,
           nullptr, st);
  }
  k_serve_residual<<<dim3(ublocks, L.n), 256, 0, st>>>(n, L, b_.as<double>(), r.as<double>());
  k_dot_list<<<dim3(nb, L.n), RB, 0, st>>>(n, r.as<double>(), n, r.as<double>(), n,
                                            part1.as<double>(), L);
  k_serve_result<<<1, 64, 0, st>>>(L, part1.as<double>(), nb, stb.as<PcgCol>(),
                                   state.as<int>(), res.as<SlotResult>());
}

}  // namespace rsc
