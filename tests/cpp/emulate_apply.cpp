// CPU emulation of the preconditioner apply's task lists (host code of the
// library only; no device needed).  For a synthetic system it builds the
// symbolic plan, the vector layout, the level-scheduled task list and the
// dataflow task list, fills the factor storage with seeded random values and
// then runs
//   (1) the level-scheduled list in launch order,
//   (2) the dataflow list sequentially in ticket order, checking that every
//       group a task waits on is complete when the task starts,
//   (3) the dataflow list on several threads that take tickets and wait on
//       completion counters, the way the persistent kernel does,
//   (4) the ready-queue form on several threads (pop a slot, run, push the
//       tasks of every group whose last wait completed, run the first of
//       them directly), as the kernel's queue dispatch does,
// with every buffer access bounds-checked.  All four must give bitwise the
// same result: each task's result depends only on its inputs, so any missing
// dependency or out-of-range index shows up.
//
//   emulate_apply <problem> <N> <max_dense_block> <max_diagonal_block> <nrhs> <threads>
//                 [subtree_elems [workers [wide_elems [list [max_tasks]]]]]
//   workers: ticket order by a list schedule on that many workgroups
//   (default 1024; 0 keeps the level order); wide_elems: products split
//   finer from this many elements (default 2^20); list: "multi" runs the
//   multi-right-hand-side list (its own subtree tasks, supernode limits of
//   the sixteen-column LDS bodies) instead of the single one; max_tasks (after
//   list): at most that many multi subtree tasks (the smallest merged)
//   problem: p3d (7-point log-normal Poisson), p2d, el (elasticity)
// Prints "ok ..." and exits 0 on success.
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../rank_structured_cholesky_amd/csrc/host/rsc_host.h"
#include "../../rank_structured_cholesky_amd/csrc/host/solve_tasks.h"

using namespace rsc;

namespace {

[[noreturn]] void fail(const std::string& m) {
  std::fprintf(stderr, "FAIL: %s\n", m.c_str());
  std::fflush(stderr);
  std::exit(1);
}

struct Buf {
  std::vector<double> v;
  const char* name;
  double& at(int64_t i) {
    if (i < 0 || i >= (int64_t)v.size())
      fail(std::string("out of range in ") + name + ": " + std::to_string(i) + " of " +
           std::to_string(v.size()));
    return v[i];
  }
};

template <typename V>
auto& chk(V& vec, int64_t i, const char* what) {
  if (i < 0 || i >= (int64_t)vec.size())
    fail(std::string("out of range in ") + what + ": " + std::to_string(i) + " of " +
         std::to_string(vec.size()));
  return vec[i];
}

struct State {
  const Plan* P;
  const SolveLayout* L;
  Buf F{{}, "factor"}, B{{}, "B"}, X{{}, "X"}, w{{}, "w"}, fv{{}, "fv"}, tv{{}, "tv"},
      part{{}, "part"}, z{{}, "z"};
  int n = 0, k = 1;
  int64_t ps = 1;
  const Subtrees* dsub = nullptr;  // subtree tasks of the dataflow list
  bool multi = false;
  std::vector<int> perm;
  void reset_work() {
    std::fill(w.v.begin(), w.v.end(), std::nan(""));
    std::fill(fv.v.begin(), fv.v.end(), std::nan(""));
    std::fill(tv.v.begin(), tv.v.end(), std::nan(""));
    std::fill(part.v.begin(), part.v.end(), std::nan(""));
    std::fill(X.v.begin(), X.v.end(), std::nan(""));
    std::fill(z.v.begin(), z.v.end(), std::nan(""));
  }
  Buf& vec(int sel) { return sel == 0 ? w : sel == 1 ? fv : sel == 2 ? tv : z; }
  int64_t stride(int sel) const {
    return sel == 0 || sel == 3 ? n : sel == 1 ? L->fv_stride : L->tv_stride;
  }
};

// forward assembly chunk; from_w: source is the permuted vector w (level
// path) instead of B through the permutation (dataflow path)
void assemble(State& S, int j, int chunk, bool from_w) {
  const dev::VecNode& N = chk(S.L->vn, j, "vn");
  const int c = N.ncols, R = N.R;
  const int p0 = chunk * kSolveChunk, p1 = std::min(p0 + kSolveChunk, c + R);
  if (p0 >= c + R) return;
  for (int col = 0; col < S.k; ++col) {
    const int64_t fb = N.fv_off + (int64_t)col * S.L->fv_stride;
    for (int i = p0; i < p1; ++i) {
      double v = 0.0;
      if (i < c) {
        const int g = N.first + i;
        v = from_w ? S.w.at(g + (int64_t)col * S.n)
                   : S.B.at(chk(S.perm, g, "perm") + (int64_t)col * S.n);
      }
      S.fv.at(fb + i) = v;
    }
    for (int q = 0; q < N.nchild; ++q) {
      const int cj = chk(S.L->children, N.child_off + q, "children");
      const dev::VecNode& C = chk(S.L->vn, cj, "vn");
      // the device visits only the precomputed range of child rows: it must
      // be exactly the rows whose parent position falls in the chunk
      const int64_t rq = 2 * ((int64_t)N.range_off + (int64_t)chunk * N.nchild + q);
      const int a0 = chk(S.L->asm_ranges, rq, "asm_ranges");
      const int a1 = chk(S.L->asm_ranges, rq + 1, "asm_ranges");
      for (int a = 0; a < C.R; ++a) {
        const int p = chk(S.L->relmap, C.relmap_off + a, "relmap");
        if (p < 0 || p >= c + R) fail("relmap entry outside the parent front");
        const bool in = p >= p0 && p < p1;
        if (in != (a >= a0 && a < a1)) fail("assembly range does not match the chunk");
        if (in)
          S.fv.at(fb + p) += S.fv.at(C.fv_off + (int64_t)col * S.L->fv_stride + C.ncols + a);
      }
    }
  }
}

void gather(State& S, int j, int chunk) {
  const dev::VecNode& N = chk(S.L->vn, j, "vn");
  const int c = N.ncols, R = N.R;
  const int p0 = chunk * kSolveChunk, p1 = std::min(p0 + kSolveChunk, c + R);
  if (p0 >= c + R) return;
  for (int col = 0; col < S.k; ++col)
    for (int i = p0; i < p1; ++i) {
      if (i < c) {
        S.w.at(N.first + i + (int64_t)col * S.n) = S.z.at(N.first + i + (int64_t)col * S.n);
      } else {
        const int src = chk(S.L->rows, N.rows_off + i - c, "rows");
        S.fv.at(N.fv_off + (int64_t)col * S.L->fv_stride + i) = S.w.at(src + (int64_t)col * S.n);
      }
    }
}

void gemv(State& S, const dev::GvTask& t) {
  for (int col = 0; col < S.k; ++col) {
    Buf& Y = S.vec(t.ysel);
    const int64_t yb = (int64_t)col * S.stride(t.ysel) + t.yoff;
    if (t.kind == 2) {
      const int64_t rs = t.ld > 0 ? t.ld : 1;
      for (int j = 0; j < t.cols; ++j) {
        double acc = 0;
        for (int s2 = 0; s2 < t.rows; ++s2)
          acc += S.part.at((int64_t)col * S.ps + ((int64_t)t.part + s2 * rs) * 32 + j);
        if (t.op) Y.at(yb + j) -= acc;
        else Y.at(yb + j) = acc;
      }
      continue;
    }
    Buf& Xv = S.vec(t.xsel);
    const int64_t xb = (int64_t)col * S.stride(t.xsel) + t.xoff;
    if (t.kind == 0) {
      if (t.rows > 64) fail("ROW task with more than 64 rows");
      // x window: the vector, or the partial slots of a fused reduction
      // summed in slot order
      std::vector<double> xw(t.cols);
      for (int j = 0; j < t.cols; ++j) {
        if (t.xpart < 0) {
          xw[j] = Xv.at(xb + j);
          continue;
        }
        double acc = 0;
        for (int s2 = 0; s2 < t.xnsub; ++s2)
          acc += S.part.at((int64_t)col * S.ps +
                           ((int64_t)t.xpart + (int64_t)(j >> 5) * t.xnsub + s2) * 32 + (j & 31));
        xw[j] = acc;
      }
      for (int i = 0; i < t.rows; ++i) {
        double acc = 0;
        for (int j = 0; j < t.cols; ++j)
          acc += S.F.at(t.mat + i + (int64_t)j * t.ld) * xw[j];
        if (t.part >= 0) S.part.at((int64_t)col * S.ps + ((int64_t)t.part + (i >> 5)) * 32 + (i & 31)) = acc;
        else if (t.op) Y.at(yb + i) -= acc;
        else Y.at(yb + i) = acc;
      }
    } else {
      if (t.cols > 32) fail("COL task with more than 32 columns");
      for (int j = 0; j < t.cols; ++j) {
        double acc = 0;
        for (int i = 0; i < t.rows; ++i)
          acc += S.F.at(t.mat + i + (int64_t)j * t.ld) * Xv.at(xb + i);
        if (t.part >= 0) S.part.at((int64_t)col * S.ps + (int64_t)t.part * 32 + j) = acc;
        else if (t.op) Y.at(yb + j) -= acc;
        else Y.at(yb + j) = acc;
      }
    }
  }
}

void scatter(State& S, int chunk) {
  const int i0 = chunk * kScatterChunk, i1 = std::min(i0 + kScatterChunk, S.n);
  for (int col = 0; col < S.k; ++col)
    for (int i = i0; i < i1; ++i)
      S.X.at(chk(S.perm, i, "perm") + (int64_t)col * S.n) = S.w.at(i + (int64_t)col * S.n);
}

// subtree tasks: every node of subtree sb, children first (forward) or
// parents first (backward); panel = [Lc^{-T}; Zp]
void subtree(State& S, int sb, bool fwd, const Subtrees& T, bool multi) {
  const int q0 = chk(T.ptr, sb, "st ptr"), q1 = chk(T.ptr, sb + 1, "st ptr");
  for (int step = 0; step < q1 - q0; ++step) {
    const int j = chk(T.nodes, fwd ? q0 + step : q1 - 1 - step, "st nodes");
    const dev::VecNode& N = chk(S.L->vn, j, "vn");
    const int64_t P = chk(S.L->panel, j, "panel");
    if (P < 0) fail("subtree node without a panel");
    const int c = N.ncols, R = N.R, ld = c + R;
    if (c > kSubtreeMaxCols || R > kSubtreeMaxRows) fail("subtree node too large");
    if (multi && !multi_subtree_node(c, R)) fail("multi subtree node exceeds its LDS limits");
    for (int col = 0; col < S.k; ++col) {
      const int64_t fb = N.fv_off + (int64_t)col * S.L->fv_stride;
      const int64_t vb = (int64_t)col * S.n;
      if (fwd) {
        std::vector<double> fc(c), y(c);
        for (int i = 0; i < c; ++i) fc[i] = S.B.at(chk(S.perm, N.first + i, "perm") + vb);
        for (int r = 0; r < R; ++r) S.fv.at(fb + c + r) = 0.0;
        for (int q = 0; q < N.nchild; ++q) {
          const int cj = chk(S.L->children, N.child_off + q, "children");
          if (T.of[cj] != sb) fail("subtree child outside the subtree");
          const dev::VecNode& C = chk(S.L->vn, cj, "vn");
          for (int a = 0; a < C.R; ++a) {
            const int p = chk(S.L->relmap, C.relmap_off + a, "relmap");
            const double v = S.fv.at(C.fv_off + (int64_t)col * S.L->fv_stride + C.ncols + a);
            if (p < c) fc[p] += v;
            else S.fv.at(fb + p) += v;
          }
        }
        for (int i = 0; i < c; ++i) {
          double acc = 0;
          for (int jj = 0; jj <= i; ++jj) acc += S.F.at(P + jj + (int64_t)i * ld) * fc[jj];
          y[i] = acc;
        }
        for (int r = 0; r < R; ++r) {
          double acc = 0;
          for (int jj = 0; jj < c; ++jj) acc += S.F.at(P + c + r + (int64_t)jj * ld) * fc[jj];
          S.fv.at(fb + c + r) -= acc;
        }
        for (int i = 0; i < c; ++i) {
          double acc = 0;
          for (int jj = i; jj < c; ++jj) acc += S.F.at(P + i + (int64_t)jj * ld) * y[jj];
          S.z.at(N.first + i + vb) = acc;
        }
      } else {
        std::vector<double> xr(R);
        for (int r = 0; r < R; ++r) xr[r] = S.w.at(chk(S.L->rows, N.rows_off + r, "rows") + vb);
        for (int i = 0; i < c; ++i) {
          double acc = 0;
          for (int r = 0; r < R; ++r) acc += S.F.at(P + c + r + (int64_t)i * ld) * xr[r];
          S.w.at(N.first + i + vb) = S.z.at(N.first + i + vb) - acc;
        }
      }
    }
  }
}

void run_df_task(State& S, const DataflowTasks& D, const dev::DfTask& t) {
  // an assembly / gather task (a = supernode) covers chunk_group
  // consecutive chunks
  const int G = S.multi ? 1 : kSingleChunkGroup;
  auto nchunks = [&]() {
    const Supernode& sn = S.P->nodes.at(t.a);
    return (sn.ncols + sn.R() + kSolveChunk - 1) / kSolveChunk;
  };
  switch (t.type) {
    case 0:
      for (int ch = t.b * G; ch < std::min(nchunks(), (t.b + 1) * G); ++ch)
        assemble(S, t.a, ch, false);
      break;
    case 1:
      for (int ch = t.b * G; ch < std::min(nchunks(), (t.b + 1) * G); ++ch) gather(S, t.a, ch);
      break;
    case 2: gemv(S, chk(D.gv, t.a, "df gv")); break;
    case 3: scatter(S, t.a); break;
    case 4: subtree(S, t.a, true, *S.dsub, S.multi); break;
    case 5: subtree(S, t.a, false, *S.dsub, S.multi); break;
    default: fail("bad task type");
  }
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 7) {
    std::fprintf(stderr, "usage: %s p3d|p2d|el N mdb mdiag nrhs threads [subtree_elems [workers]]\n",
                 argv[0]);
    return 2;
  }
  const std::string prob = argv[1];
  const int N = std::atoi(argv[2]), mdb = std::atoi(argv[3]), mdiag = std::atoi(argv[4]);
  const int k = std::atoi(argv[5]), nthreads = std::atoi(argv[6]);
  Csc A;
  std::vector<double> xyz;
  if (prob == "p2d") {
    A = poisson2d(N);
    xyz = grid_positions(N, N, 1);
  } else if (prob == "el") {
    A = elasticity3d(N, 1.0, 0.3, 7);
  } else {
    A = poisson3d_lognormal(N, 1.0, 1234);
    xyz = grid_positions(N, N, N);
  }
  rsc_options opt;
  rsc_options_default(&opt);
  opt.max_dense_block = mdb;
  opt.max_diagonal_block = mdiag;
  opt.explicit_block_threshold = 2 * mdiag;
  Plan P;
  build_plan(A, xyz.empty() ? nullptr : xyz.data(), opt, nullptr, P);
  SolveLayout L;
  const int64_t sub_elems = argc > 7 ? std::atoll(argv[7]) : std::max<int64_t>(256, P.factor_elems / 64);
  const bool multi = argc > 10 && std::string(argv[10]) == "multi";
  const int max_tasks = argc > 11 ? std::atoi(argv[11]) : 0;
  build_solve_layout(P, L, sub_elems, multi ? (int64_t)1 << 20 : 0, max_tasks);
  if (multi && max_tasks > 0 && (int)L.sub_multi.root.size() > max_tasks)
    fail("more subtree tasks than the limit");
  LevelTasks LT;
  build_level_tasks(P, L, LT);
  DataflowTasks D;
  TicketModel tm;
  tm.workers = argc > 8 ? std::atoi(argv[8]) : 1024;
  tm.wide_elems = argc > 9 ? std::atoll(argv[9]) : 0;
  tm.nrhs = k;
  tm.chunk_group = multi ? 1 : kSingleChunkGroup;
  const auto tb0 = std::chrono::steady_clock::now();
  build_dataflow_tasks(P, L, D, multi ? &L.sub_multi : &L.sub, tm);
  const double build_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tb0).count();

  State S;
  S.P = &P;
  S.L = &L;
  S.n = P.n;
  S.k = k;
  S.perm = P.perm;
  S.dsub = multi ? &L.sub_multi : &L.sub;
  S.multi = multi;
  S.ps = std::max<int64_t>(std::max(LT.slots, D.slots) * 32, 1);
  std::mt19937_64 g(99);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  S.F.v.resize(std::max<int64_t>(P.factor_elems, 1));
  for (double& x : S.F.v) x = U(g) * 0.1;
  // the stored triangular inverses (L^{-T}) are exactly zero below their
  // diagonals, as the factor leaves them
  auto zero_lower = [&](int64_t off, int b, int ld) {
    for (int j = 0; j < b; ++j)
      for (int i = j + 1; i < b; ++i) S.F.at(off + i + (int64_t)j * ld) = 0.0;
  };
  for (const Supernode& s : P.nodes) {
    if (!s.compress) zero_lower(s.off_panel, s.ncols, s.ncols + s.R());
    else
      for (const HNode& h : s.hnodes)
        if (h.leaf()) zero_lower(h.off_leaf, h.r1 - h.r0, h.r1 - h.r0);
  }
  S.B.v.resize((size_t)S.n * k);
  for (double& x : S.B.v) x = U(g);
  S.X.v.resize((size_t)S.n * k);
  S.w.v.resize((size_t)S.n * k);
  S.fv.v.resize((size_t)L.fv_stride * k);
  S.tv.v.resize((size_t)L.tv_stride * k);
  S.part.v.resize((size_t)S.ps * k);
  S.z.v.resize((size_t)S.n * k);

  // (1) level-scheduled order
  S.reset_work();
  for (int col = 0; col < k; ++col)
    for (int i = 0; i < S.n; ++i)
      S.w.at(i + (int64_t)col * S.n) = S.B.at(S.perm[i] + (int64_t)col * S.n);
  auto run_level = [&](const std::vector<SolvePhase>& ph, int lv) {
    for (const SolvePhase& p : ph) {
      if (p.kind == 4 || p.kind == 5) {
        for (int sb = 0; sb < p.count; ++sb) subtree(S, sb, p.kind == 4, L.sub, false);
      } else if (p.kind == 0 || p.kind == 1) {
        for (int j : LT.level_nodes[lv]) {
          const int nch = (P.nodes[j].ncols + P.nodes[j].R() + kSolveChunk - 1) / kSolveChunk;
          for (int c = 0; c < nch; ++c) {
            if (p.kind == 0) assemble(S, j, c, true);
            else gather(S, j, c);
          }
        }
      } else {
        for (int q = p.first; q < p.first + p.count; ++q) gemv(S, chk(LT.gv, q, "level gv"));
      }
    }
  };
  const int nlev = (int)P.levels.size();
  for (int lv = 0; lv < nlev; ++lv) run_level(LT.fwd[lv], lv);
  for (int lv = nlev - 1; lv >= 0; --lv) run_level(LT.bwd[lv], lv);
  for (int c = 0; c * kScatterChunk < S.n; ++c) scatter(S, c);
  const std::vector<double> x_level = S.X.v;
  for (double v : x_level)
    if (!std::isfinite(v)) fail("level path left a non-finite output entry");

  // (2) dataflow list, sequential, with the wait check
  const int ng = (int)D.need.size();
  if ((int)D.gwoff.size() != ng + 1) fail("group wait offsets size");
  S.reset_work();
  {
    std::vector<unsigned> cnt(ng, 0u);
    for (size_t i = 0; i < D.tasks.size(); ++i) {
      const dev::DfTask& t = D.tasks[i];
      if (t.group < 0 || t.group >= ng) fail("task group out of range");
      for (int q = D.gwoff[t.group]; q < D.gwoff[t.group + 1]; ++q) {
        const int wg = chk(D.gw, q, "gw");
        if (wg < 0 || wg >= ng) fail("waited group out of range");
        if (cnt[wg] != D.need[wg]) {
          int first = -1, last = -1, nsig = 0;
          for (size_t u = 0; u < D.tasks.size(); ++u)
            if (D.tasks[u].group == wg || D.tasks[u].group2 == wg) {
              if (first < 0) first = (int)u;
              last = (int)u;
              ++nsig;
            }
          fail("task " + std::to_string(i) + " (type " + std::to_string(t.type) + ", group " +
               std::to_string(t.group) + ") starts before group " + std::to_string(wg) +
               " completed: count " + std::to_string(cnt[wg]) + " need " +
               std::to_string(D.need[wg]) + ", signalled by " + std::to_string(nsig) +
               " tasks at " + std::to_string(first) + ".." + std::to_string(last));
        }
      }
      run_df_task(S, D, t);
      cnt[t.group] += 1u;
      if (t.group2 >= 0) chk(cnt, t.group2, "group2") += 1u;
    }
    for (int q = 0; q < ng; ++q)
      if (cnt[q] != D.need[q]) fail("group " + std::to_string(q) + " count mismatch");
  }
  // split ROW tasks (long rows, partial sums) exist only in the dataflow
  // list: then the two lists agree to rounding, otherwise bit for bit
  bool exact = true;
  for (const dev::GvTask& t : D.gv)
    if (t.kind == 0 && t.part >= 0) exact = false;
  auto agrees = [&](const std::vector<double>& x) {
    if (exact) return x == x_level;
    double dn = 0, rn = 0;
    for (size_t i = 0; i < x.size(); ++i) {
      dn += (x[i] - x_level[i]) * (x[i] - x_level[i]);
      rn += x_level[i] * x_level[i];
    }
    return std::sqrt(dn) <= 1e-12 * std::sqrt(rn);
  };
  if (!agrees(S.X.v)) fail("sequential dataflow result differs from the level path");

  // (3) dataflow list on threads: tickets + completion counters + waits
  S.reset_work();
  {
    std::vector<std::atomic<unsigned>> cnt(ng);
    for (auto& c : cnt) c.store(0u);
    std::atomic<int> ticket{0};
    std::atomic<bool> stuck{false};
    auto wait_for = [&](int wg, unsigned need) {
      const auto t0 = std::chrono::steady_clock::now();
      while (cnt[wg].load(std::memory_order_acquire) < need) {
        if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(20)) return false;
        std::this_thread::yield();
      }
      return true;
    };
    auto worker = [&]() {
      for (;;) {
        const int i = ticket.fetch_add(1);
        if (i >= (int)D.tasks.size()) return;
        const dev::DfTask& t = D.tasks[i];
        if (t.wbeg != D.gwoff[t.group] || t.wend != D.gwoff[t.group + 1])
          fail("task wait range differs from its group's");
        if (t.wend > t.wbeg && (t.wg0 != D.gw[t.wbeg] || t.wn0 != D.need[t.wg0]))
          fail("inline first wait differs from the wait list");
        bool ok = t.wend <= t.wbeg || wait_for(t.wg0, t.wn0);
        for (int q = t.wbeg + 1; ok && q < t.wend; ++q) ok = wait_for(D.gw[q], D.need[D.gw[q]]);
        if (!ok) {
          stuck = true;
          return;
        }
        run_df_task(S, D, t);
        cnt[t.group].fetch_add(1u, std::memory_order_release);
        if (t.group2 >= 0) cnt[t.group2].fetch_add(1u, std::memory_order_release);
      }
    };
    std::vector<std::thread> th;
    for (int q = 0; q < nthreads; ++q) th.emplace_back(worker);
    for (auto& t : th) t.join();
    if (stuck) fail("threaded dataflow run stalled");
  }
  if (!agrees(S.X.v)) fail("threaded dataflow result differs from the level path");

  // (4) the ready-queue form on threads, as the device runs it: pop a slot,
  // wait until it is written, run the task (every wait must be complete),
  // count down the pending counts of the groups waiting on its groups (a
  // small group's tasks each, a larger group's last task once); the task
  // bringing a group's count to zero pushes the tasks of that group
  S.reset_work();
  {
    const int nt = (int)D.tasks.size();
    std::vector<std::atomic<unsigned>> cnt(ng);
    std::vector<std::atomic<int>> pend(ng), slot(nt);
    std::vector<std::atomic<int>> runs(nt);
    for (int g = 0; g < ng; ++g) {
      cnt[g].store(0u);
      pend[g].store(D.q_pend0[g]);
    }
    for (int i = 0; i < nt; ++i) {
      slot[i].store(i < (int)D.q_init.size() ? D.q_init[i] : -1);
      runs[i].store(0);
    }
    std::atomic<int> head{0}, tail{(int)D.q_init.size()};
    std::atomic<bool> stuck{false};
    if ((int)D.q_gptr.size() != ng + 1 || (int)D.q_sptr.size() != ng + 1 ||
        (int)D.q_gtasks.size() != nt || (int)D.q_ginfo.size() != 4 * ng)
      fail("ready-queue arrays have the wrong sizes");
    for (int g = 0; g < ng; ++g) {
      const int c0 = D.q_gptr[g], cn = D.q_gptr[g + 1] - c0;
      if (D.q_ginfo[4 * g] != c0 || D.q_ginfo[4 * g + 1] != cn ||
          D.q_ginfo[4 * g + 2] != (cn > 0 ? D.q_gtasks[c0] : -1))
        fail("a group's ready-queue record disagrees with its task list");
    }
    for (const auto& t : D.tasks)
      if (t.succ1[0] != D.q_sptr[t.group] || t.succ1[1] != D.q_sptr[t.group + 1] ||
          (t.group2 >= 0 &&
           (t.succ2[0] != D.q_sptr[t.group2] || t.succ2[1] != D.q_sptr[t.group2 + 1])))
        fail("a task descriptor's successor ranges disagree with the queue lists");
    for (const auto& t : D.tasks)
      if (t.sfirst[0] != (t.succ1[1] > t.succ1[0] ? D.q_succ[t.succ1[0]] : -1) ||
          (t.group2 >= 0 && t.sfirst[1] != (t.succ2[1] > t.succ2[0] ? D.q_succ[t.succ2[0]] : -1)))
        fail("a task descriptor's first successors disagree with the queue lists");
    constexpr int kSkip = -3;
    auto worker = [&]() {
      int next = -1;  // the task this worker made ready and runs itself
      for (;;) {
        int i = next;
        next = -1;
        while (i < 0) {
          const int idx = head.fetch_add(1);
          if (idx >= nt) return;
          const auto t0 = std::chrono::steady_clock::now();
          while ((i = slot[idx].load(std::memory_order_acquire)) == -1) {
            if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(20)) {
              stuck = true;
              return;
            }
            std::this_thread::yield();
          }
          if (i == kSkip) i = -1;
        }
        if (i >= nt) fail("queue slot holds an invalid task");
        if (runs[i].fetch_add(1) != 0) fail("task run twice");
        const dev::DfTask& t = D.tasks[i];
        for (int q = D.gwoff[t.group]; q < D.gwoff[t.group + 1]; ++q)
          if (cnt[D.gw[q]].load(std::memory_order_acquire) != D.need[D.gw[q]])
            fail("queued task run before its waits completed");
        run_df_task(S, D, t);
        if (t.need1 != D.need[t.group] || t.need2 != (t.group2 >= 0 ? D.need[t.group2] : 0u))
          fail("task descriptor holds wrong completion counts");
        const int gs[2] = {t.group, t.group2};
        for (int g : gs) {
          if (g < 0) continue;
          // a task of a small group counts down the groups waiting on g
          // directly, in a larger group only the task completing it
          const unsigned o = cnt[g].fetch_add(1u, std::memory_order_acq_rel);
          if (D.need[g] > dev::kQDirectMax && o + 1u != D.need[g]) continue;
          for (int q = D.q_sptr[g]; q < D.q_sptr[g + 1]; ++q) {
            const int h = D.q_succ[q];
            if (pend[h].fetch_sub(1, std::memory_order_acq_rel) != 1) continue;
            const int c0 = D.q_gptr[h], cn = D.q_gptr[h + 1] - c0;
            if (cn <= 0) continue;
            const int base = tail.fetch_add(cn);
            if (base + cn > nt) fail("ready queue overflows");
            int first = 0;
            if (next < 0) {
              next = D.q_gtasks[c0];
              first = 1;
            }
            for (int u = 0; u < cn; ++u)
              slot[base + u].store(u < first ? kSkip : D.q_gtasks[c0 + u],
                                   std::memory_order_release);
          }
        }
      }
    };
    std::vector<std::thread> th;
    for (int q = 0; q < nthreads; ++q) th.emplace_back(worker);
    for (auto& t : th) t.join();
    if (stuck) fail("ready-queue run stalled");
    for (int i = 0; i < nt; ++i)
      if (runs[i].load() != 1) fail("a task did not run exactly once in the ready-queue run");
    if (tail.load() != nt) fail("ready queue did not receive every task");
  }
  if (!agrees(S.X.v)) fail("ready-queue result differs from the level path");
  int fused = 0;
  for (size_t i = 0; i < D.tasks.size(); ++i)
    if (D.tasks[i].type == 2 && D.gv[i].kind == 0 && D.gv[i].xpart >= 0) ++fused;
  std::printf(
      "ok%s n=%d nodes=%d levels=%d subtrees=%zu level_tasks=%zu df_tasks=%zu groups=%d "
      "fused_row_tasks=%d df_build_ms=%.1f\n",
      exact ? "" : " (split rows)", S.n, (int)P.nodes.size(), nlev, S.dsub->root.size(),
      LT.gv.size(), D.tasks.size(), ng, fused, build_ms);
  return 0;
}
