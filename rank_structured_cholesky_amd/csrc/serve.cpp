// Request server: continuous batching of PCG right-hand sides over factors
// resident in HBM (declared in include/rsc.h).  The reference has no
// counterpart; it is a caller of the preconditioner apply and PCG path
// (FactorPreconditioner::solve, PCG_SOLVER::solvePCG).
//
// One scheduler thread drives one stream per factor.  Per factor a pool of
// column slots (rsc::ServeBatch) is advanced one PCG iteration per step.
// The host reads each step's slot states one step behind (the next step is
// already queued), retires converged columns (x copied out, true residual)
// and admits queued columns into the freed slots between steps, so requests
// that arrive while others run join at the next step.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>

#include "capi_internal.h"

namespace {

using clk = std::chrono::steady_clock;

struct Request {
  int64_t id = 0;
  int lane = 0, m = 0;
  const double* B = nullptr;
  int ldb = 0;
  double* X = nullptr;
  int ldx = 0;
  rsc::SlotResult* res = nullptr;  // pinned, m entries
  int admitted = 0, finished = 0;
  clk::time_point t0, t1;
  bool done = false;
};

struct Retire {
  hipEvent_t ev = nullptr;
  int* apst = nullptr;  // pinned: the apply status words at retirement
  std::vector<std::pair<Request*, int>> cols;
};

struct Lane {
  rsc_factor* f = nullptr;
  hipStream_t st = nullptr;
  rsc::ServeBatch sb;
  int kcap = 0;
  std::vector<Request*> slot_req;
  std::vector<int> slot_col;
  std::vector<int64_t> slot_since;  // index of the first step after admission
  int occupied = 0;
  std::deque<Request*> queue;
  int* snap = nullptr;  // pinned: states after the two most recent steps
  hipEvent_t sev[2] = {nullptr, nullptr};
  int64_t steps = 0, harvested = 0;
  bool admitted = false;
  std::deque<Retire> retq;
  std::vector<hipEvent_t> evpool;
  std::vector<int*> stpool;  // pinned pairs of status words

  int kact() const {
    for (int s = kcap - 1; s >= 0; --s)
      if (slot_req[s]) return s + 1;
    return 0;
  }
  bool idle() const {
    return occupied == 0 && queue.empty() && retq.empty() && steps == harvested;
  }
  int* status_words() {
    if (!stpool.empty()) {
      int* p = stpool.back();
      stpool.pop_back();
      return p;
    }
    void* p = nullptr;
    RSC_HIP(hipHostMalloc(&p, 2 * sizeof(int), hipHostMallocDefault));
    return (int*)p;
  }
  hipEvent_t event() {
    if (!evpool.empty()) {
      hipEvent_t e = evpool.back();
      evpool.pop_back();
      return e;
    }
    hipEvent_t e;
    RSC_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return e;
  }
};

}  // namespace

struct rsc_server {
  rsc_server_options opt{};
  std::vector<std::unique_ptr<Lane>> lanes;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> incoming;
  std::map<int64_t, Request*> requests;
  int64_t next_id = 1;
  bool stop = false;
  bool failed = false;
  std::string error;
  int64_t total_steps = 0;
  int max_busy = 0;
  std::thread worker;

  void run();
  bool service(Lane& L);
};

namespace {

rsc_status fail(rsc_status st, const std::string& msg) {
  rsc::last_error() = msg;
  return st;
}

// Retire every slot that a harvested state snapshot shows converged.
void retire_slots(rsc_server& S, Lane& L, const std::vector<int>& slots) {
  const int n = L.sb.n;
  for (size_t q0 = 0; q0 < slots.size(); q0 += rsc::kSlotList) {
    rsc::SlotList sl;
    Retire rt;
    for (size_t q = q0; q < slots.size() && sl.n < rsc::kSlotList; ++q) sl.s[sl.n++] = slots[q];
    L.sb.retire(L.f->dA, sl, L.st);
    for (int q = 0; q < sl.n; ++q) {
      const int s = sl.s[q];
      Request* r = L.slot_req[s];
      const int c = L.slot_col[s];
      RSC_HIP(hipMemcpyAsync(r->X + (size_t)c * r->ldx, L.sb.x.as<double>() + (size_t)s * n,
                             (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, L.st));
      RSC_HIP(hipMemcpyAsync(r->res + c, L.sb.res.as<rsc::SlotResult>() + s,
                             sizeof(rsc::SlotResult), hipMemcpyDeviceToHost, L.st));
      rt.cols.push_back({r, c});
      L.slot_req[s] = nullptr;
      --L.occupied;
    }
    // the apply's status words travel with the retirement: a column is
    // returned only when every apply that produced it succeeded
    rt.apst = L.status_words();
    RSC_HIP(hipMemcpyAsync(rt.apst, L.f->eng->apply_status_device(), 2 * sizeof(int),
                           hipMemcpyDeviceToHost, L.st));
    rt.ev = L.event();
    RSC_HIP(hipEventRecord(rt.ev, L.st));
    L.retq.push_back(std::move(rt));
  }
  (void)S;
}

}  // namespace

// One pass over a factor's lane; returns whether anything moved.
bool rsc_server::service(Lane& L) {
  bool progress = false;
  const int n = L.sb.n;
  // finished retirements, in order
  while (!L.retq.empty() && hipEventQuery(L.retq.front().ev) == hipSuccess) {
    Retire& rt = L.retq.front();
    L.f->eng->raise_apply_status(rt.apst);  // throws: the server fails
    L.stpool.push_back(rt.apst);
    rt.apst = nullptr;
    {
      std::lock_guard<std::mutex> lk(mu);
      for (auto& rc : rt.cols) {
        Request* r = rc.first;
        if (++r->finished == r->m) {
          r->t1 = clk::now();
          r->done = true;
        }
      }
    }
    cv_done.notify_all();
    L.evpool.push_back(rt.ev);
    L.retq.pop_front();
    progress = true;
  }
  // the oldest queued step's states
  if (L.harvested < L.steps && hipEventQuery(L.sev[L.harvested & 1]) == hipSuccess) {
    const int64_t j = L.harvested;
    const int* sn = L.snap + (j & 1) * L.kcap;
    std::vector<int> done;
    for (int s = 0; s < L.kcap; ++s)
      if (L.slot_req[s] && j >= L.slot_since[s] && sn[s] == rsc::SLOT_DONE) done.push_back(s);
    ++L.harvested;
    if (!done.empty()) retire_slots(*this, L, done);
    progress = true;
  }
  // admissions into free slots, lowest first, in arrival order
  {
    rsc::SlotList sl;
    auto flush = [&]() {
      if (sl.n == 0) return;
      L.sb.admit(sl, L.st);
      L.admitted = true;
      sl.n = 0;
    };
    int s = 0;
    while (!L.queue.empty()) {
      while (s < L.kcap && L.slot_req[s]) ++s;
      if (s >= L.kcap) break;
      Request* r = L.queue.front();
      const int c = r->admitted++;
      RSC_HIP(hipMemcpyAsync(L.sb.b_.as<double>() + (size_t)s * n, r->B + (size_t)c * r->ldb,
                             (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, L.st));
      L.slot_req[s] = r;
      L.slot_col[s] = c;
      L.slot_since[s] = L.steps;
      ++L.occupied;
      sl.s[sl.n++] = s;
      if (sl.n == rsc::kSlotList) flush();
      if (r->admitted == r->m) L.queue.pop_front();
      progress = true;
    }
    flush();
  }
  // one more step while slots are busy (at most two steps in flight)
  if (L.occupied > 0 && L.steps - L.harvested < 2) {
    const int kact = L.kact();
    L.sb.step(*L.f->eng, L.f->dA, kact, opt.eps, opt.max_iters, L.admitted, L.st);
    L.admitted = false;
    const int64_t j = L.steps;
    RSC_HIP(hipMemcpyAsync(L.snap + (j & 1) * L.kcap, L.sb.state.p, (size_t)L.kcap * sizeof(int),
                           hipMemcpyDeviceToHost, L.st));
    RSC_HIP(hipEventRecord(L.sev[j & 1], L.st));
    ++L.steps;
    ++total_steps;
    max_busy = std::max(max_busy, kact);
    progress = true;
  }
  return progress;
}

void rsc_server::run() {
  try {
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu);
        while (!incoming.empty()) {
          Request* r = incoming.front();
          incoming.pop_front();
          lanes[r->lane]->queue.push_back(r);
        }
        bool all_idle = true;
        for (auto& L : lanes) all_idle = all_idle && L->idle();
        if (all_idle) {
          if (stop) break;
          cv_work.wait(lk, [&] { return stop || !incoming.empty(); });
          continue;
        }
      }
      bool progress = false;
      for (auto& L : lanes) {
        const bool was_idle = L->idle();
        progress |= service(*L);
        if (!was_idle && L->idle()) L->f->eng->check_apply(L->st);
      }
      if (!progress) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
  } catch (const std::exception& e) {
    std::lock_guard<std::mutex> lk(mu);
    failed = true;
    error = e.what();
  }
  cv_done.notify_all();
}

namespace {

// Device bytes one column slot of factor f takes: the slot's PCG vectors
// (b, x, r, d, q, s) and the apply buffers of one more column.
int64_t column_bytes(const rsc_factor* f) {
  const int64_t n = f->plan.n;
  return 6 * n * (int64_t)sizeof(double) +
         f->eng->apply_scratch_doubles(1) * (int64_t)sizeof(double) + 4096;
}

// Slots per factor from one memory budget: water-filling over the bytes per
// column (every factor an equal share; a factor that reaches the slot
// limit gives the rest of its share back to the others), at least one slot
// each.  The budget is o.memory_bytes, or half of the free device memory.
std::vector<int> slot_counts(rsc_factor* const* fs, int nf, const rsc_server_options& o) {
  int64_t budget = o.memory_bytes;
  if (budget <= 0) {
    size_t freeb = 0, total = 0;
    RSC_HIP(hipMemGetInfo(&freeb, &total));
    budget = (int64_t)(freeb / 2);
  }
  std::vector<int64_t> cb(nf);
  for (int i = 0; i < nf; ++i) cb[i] = column_bytes(fs[i]);
  std::vector<int> k(nf, 0);
  std::vector<bool> open(nf, true);
  int64_t left = budget;
  for (int round = 0; round < nf; ++round) {
    int nopen = 0;
    for (int i = 0; i < nf; ++i) nopen += open[i] ? 1 : 0;
    if (nopen == 0 || left <= 0) break;
    const int64_t share = left / nopen;
    bool capped = false;
    for (int i = 0; i < nf; ++i) {
      if (!open[i]) continue;
      const int64_t want = (int64_t)o.slots - k[i];
      const int64_t can = std::min<int64_t>(want, share / cb[i]);
      if (can >= want) {
        k[i] = o.slots;
        left -= want * cb[i];
        open[i] = false;
        capped = true;
      }
    }
    if (!capped) {
      // no factor reaches the limit: every open one takes its share
      for (int i = 0; i < nf; ++i)
        if (open[i]) {
          const int64_t can = std::min<int64_t>((int64_t)o.slots - k[i], share / cb[i]);
          k[i] += (int)can;
          left -= can * cb[i];
          open[i] = false;
        }
    }
  }
  for (int i = 0; i < nf; ++i) k[i] = std::max(1, k[i]);
  return k;
}

}  // namespace

extern "C" {

rsc_status rsc_server_options_default(rsc_server_options* o) {
  if (!o) return fail(RSC_BAD_INPUT, "null options");
  o->eps = 1e-8;
  o->max_iters = 10000;
  o->slots = rsc::kSlotList;
  o->memory_bytes = 0;
  o->release_workspace = 1;
  o->reserved = 0;
  return RSC_OK;
}

rsc_status rsc_server_create(rsc_factor* const* factors, int nfactors,
                             const rsc_server_options* opt, rsc_server** out) {
  if (!out) return fail(RSC_BAD_INPUT, "out is null");
  *out = nullptr;
  if (!factors || nfactors <= 0) return fail(RSC_BAD_INPUT, "no factors");
  for (int i = 0; i < nfactors; ++i) {
    if (!factors[i]) return fail(RSC_BAD_INPUT, "null factor");
    // two lanes on one factor would share its apply buffers and counters
    for (int j = 0; j < i; ++j)
      if (factors[j] == factors[i]) return fail(RSC_BAD_INPUT, "factor passed twice");
  }
  rsc_server_options o;
  if (opt) o = *opt;
  else rsc_server_options_default(&o);
  if (o.slots < 1 || o.slots > rsc::kSlotList || o.max_iters < 0 || !(o.eps >= 0) ||
      o.memory_bytes < 0)
    return fail(RSC_BAD_INPUT, "bad server options (slots 1..64, memory_bytes >= 0)");
  try {
    std::unique_ptr<rsc_server> S(new rsc_server());
    S->opt = o;
    if (o.release_workspace)
      for (int i = 0; i < nfactors; ++i) {
        factors[i]->eng->release_workspace();
        factors[i]->info.workspace_bytes = factors[i]->eng->workspace_bytes();
      }
    const std::vector<int> caps = slot_counts(factors, nfactors, o);
    for (int i = 0; i < nfactors; ++i) {
      std::unique_ptr<Lane> L(new Lane());
      L->f = factors[i];
      const int kc = caps[i];
      L->kcap = kc;
      const int n = L->f->plan.n;
      RSC_HIP(hipStreamCreate(&L->st));
      L->sb.alloc(n, kc);
      L->slot_req.assign(kc, nullptr);
      L->slot_col.assign(kc, 0);
      L->slot_since.assign(kc, 0);
      RSC_HIP(hipHostMalloc((void**)&L->snap, 2 * (size_t)kc * sizeof(int),
                            hipHostMallocDefault));
      for (hipEvent_t& e : L->sev) RSC_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      // the apply's per-column buffers, grown once here (never mid-stream)
      RSC_HIP(hipMemsetAsync(L->sb.r.p, 0, L->sb.r.bytes, L->st));
      L->f->eng->apply(kc, L->sb.r.as<double>(), n, L->sb.s.as<double>(), n, L->st);
      RSC_HIP(hipStreamSynchronize(L->st));
      L->f->eng->check_apply(L->st);
      S->lanes.push_back(std::move(L));
    }
    rsc_server* raw = S.release();
    raw->worker = std::thread([raw] { raw->run(); });
    *out = raw;
    return RSC_OK;
  } catch (const std::bad_alloc&) {
    return fail(RSC_OOM, "server slot memory");
  } catch (const std::exception& e) {
    return fail(RSC_DEVICE_ERROR, e.what());
  }
}

rsc_status rsc_server_submit(rsc_server* s, int factor, int nrhs, const double* B, int ldb,
                             double* X, int ldx, int64_t* ticket) {
  if (!s || !B || !X || !ticket || nrhs <= 0 || factor < 0 ||
      factor >= (int)s->lanes.size())
    return fail(RSC_BAD_INPUT, "bad request");
  const int n = s->lanes[factor]->sb.n;
  if (ldb < n || ldx < n) return fail(RSC_BAD_INPUT, "leading dimension below n");
  Request* r = new Request();
  r->lane = factor;
  r->m = nrhs;
  r->B = B;
  r->ldb = ldb;
  r->X = X;
  r->ldx = ldx;
  if (hipHostMalloc((void**)&r->res, (size_t)nrhs * sizeof(rsc::SlotResult),
                    hipHostMallocDefault) != hipSuccess) {
    delete r;
    return fail(RSC_OOM, "pinned result buffer");
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->failed) {
      (void)hipHostFree(r->res);
      delete r;
      return fail(RSC_DEVICE_ERROR, s->error);
    }
    r->id = s->next_id++;
    r->t0 = clk::now();
    s->requests[r->id] = r;
    s->incoming.push_back(r);
    *ticket = r->id;
  }
  s->cv_work.notify_one();
  return RSC_OK;
}

rsc_status rsc_server_wait(rsc_server* s, int64_t ticket, rsc_pcg_result* res) {
  if (!s) return fail(RSC_BAD_INPUT, "null server");
  Request* r = nullptr;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->requests.find(ticket);
    if (it == s->requests.end()) return fail(RSC_BAD_INPUT, "unknown ticket");
    r = it->second;
    s->cv_done.wait(lk, [&] { return r->done || s->failed; });
    if (!r->done) return fail(RSC_DEVICE_ERROR, s->error);
    s->requests.erase(it);
  }
  const double secs = std::chrono::duration<double>(r->t1 - r->t0).count();
  rsc_status st = RSC_OK;
  for (int c = 0; c < r->m; ++c) {
    const rsc::SlotResult& q = r->res[c];
    if (res) {
      res[c].iterations = q.iterations;
      res[c].recurrence_relres = q.recurrence_relres;
      res[c].true_relres = std::sqrt(q.rr) / q.bnorm;
      res[c].seconds = secs;
    }
    if (!(q.recurrence_relres <= s->opt.eps)) st = RSC_NOT_CONVERGED;
  }
  (void)hipHostFree(r->res);
  delete r;
  return st;
}

rsc_status rsc_server_stats(const rsc_server* s, int64_t* steps, int* max_busy) {
  if (!s) return fail(RSC_BAD_INPUT, "null server");
  std::lock_guard<std::mutex> lk(const_cast<rsc_server*>(s)->mu);
  if (steps) *steps = s->total_steps;
  if (max_busy) *max_busy = s->max_busy;
  return RSC_OK;
}

rsc_status rsc_server_slots(const rsc_server* s, int factor, int* slots) {
  if (!s || !slots || factor < 0 || factor >= (int)s->lanes.size())
    return fail(RSC_BAD_INPUT, "bad argument");
  *slots = s->lanes[factor]->kcap;
  return RSC_OK;
}

void rsc_server_destroy(rsc_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  for (auto& L : s->lanes) {
    (void)hipStreamSynchronize(L->st);
    for (hipEvent_t e : L->evpool) (void)hipEventDestroy(e);
    for (int* p : L->stpool) (void)hipHostFree(p);
    for (auto& rt : L->retq) {
      (void)hipEventDestroy(rt.ev);
      if (rt.apst) (void)hipHostFree(rt.apst);
    }
    for (hipEvent_t e : L->sev)
      if (e) (void)hipEventDestroy(e);
    if (L->snap) (void)hipHostFree(L->snap);
    (void)hipStreamDestroy(L->st);
  }
  for (auto& kv : s->requests) {
    (void)hipHostFree(kv.second->res);
    delete kv.second;
  }
  delete s;
}

}  // extern "C"
