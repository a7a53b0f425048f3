// Symbolic plan of the rank-structured supernodal factor.
//
// Follows SparseSolver::factorSymbolic (src/rschol/solver/SparseSolver.cpp:
// 95-154) and FactorManager::factorSymbolic (FactorManager.cpp:280-349):
// ordering -> supernodes -> separator reordering -> per-supernode row sets
// (Supernode::findSystemRows / propagateFillIn, Supernode.cpp:11771-11967)
// -> diagonal block trees (DiagonalBlock::init / buildTraversal,
// DiagonalBlock.cpp:92-173, 735-790) -> ranks (FactorManager::blockRank /
// diagonalBlockRank, FactorManager.cpp:7283-7338) -> storage.
//
// Storage differs from the reference by design: every supernode's front is
// formed explicitly on the device (multifrontal), uncompressed supernodes
// keep their dense (ncols+R) x ncols panel, compressed ones keep only their
// diagonal-block leaves and low-rank (V, U) pairs.
#include <algorithm>
#include <atomic>
#include <memory>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <functional>
#include <map>
#include <numeric>
#include <random>
#include <stdexcept>
#include <thread>

#include "rsc_host.h"

namespace rsc {

int block_rank(int nrows, int ncols, double constant, int oversampling,
               int min_rank, int max_rank) {
  double nz = std::sqrt((double)std::min(ncols, nrows));
  int b = (int)(nz * std::log2(nz) * constant);
  b += oversampling;
  b = std::max(b, min_rank);
  b = std::min(b, nrows);
  b = std::min(b, ncols);
  if (max_rank > 0) b = std::min(b, max_rank);
  return b;
}

namespace {

// MT19937 (Matsumoto & Nishimura 1998) in block form: init_genrand seeding
// and the 624-word refill of MERSENNETWISTER.h:74-78,285-297, written
// branch-free so the refill loops vectorize.
struct Mt19937 {
  static constexpr int N = 624, M = 397;
  uint32_t mt[N];
  int idx = N;
  explicit Mt19937(uint32_t seed) {
    mt[0] = seed;
    for (int i = 1; i < N; ++i) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
  }
  static uint32_t twist(uint32_t a, uint32_t b) {
    uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
    return (y >> 1) ^ ((0u - (b & 1u)) & 0x9908b0dfu);
  }
  void refill() {
    for (int i = 0; i < N - M; ++i) mt[i] = mt[i + M] ^ twist(mt[i], mt[i + 1]);
    for (int i = N - M; i < N - 1; ++i) mt[i] = mt[i + M - N] ^ twist(mt[i], mt[i + 1]);
    mt[N - 1] = mt[M - 1] ^ twist(mt[N - 1], mt[0]);
    idx = 0;
  }
  static uint32_t temper(uint32_t y) {
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    return y ^ (y >> 18);
  }
  // the next n outputs
  void fill(uint32_t* out, int64_t n) {
    while (n > 0) {
      if (idx == N) refill();
      int64_t m = std::min<int64_t>(n, N - idx);
      for (int64_t k = 0; k < m; ++k) out[k] = temper(mt[idx + k]);
      idx += (int)m;
      out += m;
      n -= m;
    }
  }
};

// randNorm on one pair of draws: r from randDblExc (the first draw), phi
// from randExc (the second), value r cos(phi) (MERSENNETWISTER.h:153-177).
inline double box_muller(uint32_t a, uint32_t b) {
  double u1 = ((double)a + 0.5) * (1.0 / 4294967296.0);
  double r = std::sqrt(-2.0 * std::log(1.0 - u1));
  double phi = 2.0 * 3.14159265358979323846264338328 * ((double)b * (1.0 / 4294967296.0));
  return r * std::cos(phi);
}

// The raw MT19937 stream is sequential; its transform is not.  The calling
// thread generates the raw pairs chunk by chunk into a small ring of
// buffers and publishes its progress; the other threads transform published
// chunks in order and hand their ring slots back.  Every entry is computed
// by the same scalar code as a single-threaded loop, so the result does not
// depend on the thread count.
template <typename T>
void gaussian_fill(uint32_t seed, int64_t count, T* out, int nthreads) {
  if (count <= 0) return;
  constexpr int64_t kChunk = 1 << 16;  // values per chunk
  const int64_t nchunks = (count + kChunk - 1) / kChunk;
  nthreads = (int)std::max<int64_t>(1, std::min<int64_t>(nthreads, nchunks));
  if (nthreads == 1) {
    Mt19937 g(seed);
    uint32_t pair[2];
    for (int64_t k = 0; k < count; ++k) {
      g.fill(pair, 2);
      out[k] = (T)box_muller(pair[0], pair[1]);
    }
    return;
  }
  const int slots = 4 * nthreads;
  std::unique_ptr<uint32_t[]> ring(new uint32_t[(size_t)slots * 2 * kChunk]);
  // chunk c uses slot c % slots once chunk c - slots has been transformed
  std::unique_ptr<std::atomic<int64_t>[]> free_for(new std::atomic<int64_t>[slots]);
  for (int q = 0; q < slots; ++q) free_for[q].store(q);
  std::atomic<int64_t> produced{0};  // chunks whose raw draws are written
  std::atomic<int64_t> next{0};      // next chunk to transform
  auto worker = [&]() {
    for (;;) {
      const int64_t c = next.fetch_add(1);
      if (c >= nchunks) return;
      while (produced.load(std::memory_order_acquire) <= c) std::this_thread::yield();
      const uint32_t* raw = ring.get() + (size_t)(c % slots) * 2 * kChunk;
      const int64_t b = c * kChunk, e = std::min(count, b + kChunk);
      for (int64_t k = b; k < e; ++k)
        out[k] = (T)box_muller(raw[2 * (k - b)], raw[2 * (k - b) + 1]);
      free_for[c % slots].store(c + slots, std::memory_order_release);
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nthreads; ++t) pool.emplace_back(worker);
  {
    Mt19937 g(seed);
    for (int64_t c = 0; c < nchunks; ++c) {
      const int q = (int)(c % slots);
      while (free_for[q].load(std::memory_order_acquire) != c) std::this_thread::yield();
      const int64_t b = c * kChunk, e = std::min(count, b + kChunk);
      g.fill(ring.get() + (size_t)q * 2 * kChunk, 2 * (e - b));
      produced.store(c + 1, std::memory_order_release);
    }
  }
  worker();
  for (auto& t : pool) t.join();
}

}  // namespace

int stream_threads() {
  unsigned hw = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(hw, 16u));
}

void gaussian_stream(uint32_t seed, int64_t count, double* out) {
  gaussian_fill<double>(seed, count, out, stream_threads());
}

void gaussian_stream(uint32_t seed, int64_t count, float* out) {
  gaussian_fill<float>(seed, count, out, stream_threads());
}

void gaussian_stream_threads(uint32_t seed, int64_t count, double* out, int nthreads) {
  gaussian_fill<double>(seed, count, out, nthreads);
}

LowerCsc permuted_lower(const Csc& A, const std::vector<int>& pinv) {
  const int n = A.n;
  std::vector<int> perm(n);
  for (int o = 0; o < n; ++o) perm[pinv[o]] = o;
  LowerCsc L;
  L.p.assign(n + 1, 0);
  // columns are independent: count, prefix sum, then fill and sort each
  parallel_for(n, 4096, [&](int64_t b, int64_t e, int) {
    for (int64_t k = b; k < e; ++k) {
      const int o = perm[k];
      int cnt = 0;
      for (int p = A.p[o]; p < A.p[o + 1]; ++p)
        if (pinv[A.i[p]] >= k) ++cnt;
      L.p[k + 1] = cnt;
    }
  });
  for (int k = 0; k < n; ++k) L.p[k + 1] += L.p[k];
  L.i.resize(L.p[n]);
  L.src.resize(L.p[n]);
  parallel_for(n, 4096, [&](int64_t b, int64_t e, int) {
    std::vector<std::pair<int, int64_t>> tmp;
    for (int64_t k = b; k < e; ++k) {
      const int o = perm[k];
      tmp.clear();
      for (int p = A.p[o]; p < A.p[o + 1]; ++p) {
        const int r = pinv[A.i[p]];
        if (r >= k) tmp.push_back({r, (int64_t)p});
      }
      std::sort(tmp.begin(), tmp.end());
      int q = L.p[k];
      for (auto& x : tmp) {
        L.i[q] = x.first;
        L.src[q] = x.second;
        ++q;
      }
    }
  });
  return L;
}

namespace {

// Build the HODLR tree over a compressed separator's diagonal leaves
// (bottom-up pairing of adjacent blocks, DiagonalBlock.cpp:92-120) and the
// in-order traversal with explicit blocks (:735-756).
void build_hodlr(Supernode& s, const std::vector<int>& leaf_off, int explicit_thr) {
  s.hnodes.clear();
  s.hsteps.clear();
  std::vector<int> cur;
  int nleaf = (int)leaf_off.size() - 1;
  if (nleaf <= 0) {
    HNode h;
    h.r0 = 0;
    h.r1 = s.ncols;
    s.hnodes.push_back(h);
    s.hsteps.push_back({0, HStep::LEAF});
    return;
  }
  if ((nleaf & (nleaf - 1)) != 0)
    throw std::runtime_error("diagonal block count must be a power of two");
  // store leaves first, then internal levels; root moved to the end
  for (int b = 0; b < nleaf; ++b) {
    HNode h;
    h.r0 = leaf_off[b];
    h.r1 = leaf_off[b + 1];
    s.hnodes.push_back(h);
    cur.push_back((int)s.hnodes.size() - 1);
  }
  while (cur.size() > 1) {
    std::vector<int> next;
    for (size_t k = 0; k < cur.size(); k += 2) {
      HNode h;
      h.left = cur[k];
      h.right = cur[k + 1];
      h.r0 = s.hnodes[h.left].r0;
      h.r1 = s.hnodes[h.right].r1;
      s.hnodes.push_back(h);
      int id = (int)s.hnodes.size() - 1;
      s.hnodes[cur[k]].parent = id;
      s.hnodes[cur[k + 1]].parent = id;
      next.push_back(id);
    }
    cur.swap(next);
  }
  int root = cur[0];
  for (auto& h : s.hnodes)
    if (!h.leaf()) h.height = std::max(s.hnodes[h.left].height, s.hnodes[h.right].height) + 1;
  bool use_explicit = nleaf > 1 && explicit_thr > 0;
  std::function<void(int)> trav = [&](int id) {
    HNode& h = s.hnodes[id];
    int rows = h.r1 - h.r0;
    bool big = !use_explicit || rows > explicit_thr;
    if (!h.leaf() && big) trav(h.left);
    if (use_explicit && rows <= explicit_thr) {
      h.explicit_build = true;
      s.hsteps.push_back({id, HStep::EXPLICIT});
    } else if (h.leaf()) {
      s.hsteps.push_back({id, HStep::LEAF});
    } else {
      s.hsteps.push_back({id, HStep::OFFDIAG});
    }
    if (!h.leaf() && big) trav(h.right);
  };
  trav(root);
  // mark descendants of explicit nodes as explicit too (built inside)
  std::function<void(int)> mark = [&](int id) {
    s.hnodes[id].explicit_build = true;
    if (!s.hnodes[id].leaf()) {
      mark(s.hnodes[id].left);
      mark(s.hnodes[id].right);
    }
  };
  for (auto& st : s.hsteps)
    if (st.kind == HStep::EXPLICIT) mark(st.node);
  // move root to index 0 is not needed; record it via hsteps
}

}  // namespace

int hodlr_root(const Supernode& s) {
  for (int k = 0; k < (int)s.hnodes.size(); ++k)
    if (s.hnodes[k].parent < 0) return k;
  return 0;
}

namespace {

// Best-fit offsets in one region (holes indexed by length and by offset),
// with coalescing frees.
struct FirstFit {
  std::map<int64_t, int64_t> holes;              // offset -> length
  std::multimap<int64_t, int64_t> by_len;        // length -> offset
  int64_t top = 0;
  void erase_len(int64_t len, int64_t off) {
    auto r = by_len.equal_range(len);
    for (auto it = r.first; it != r.second; ++it)
      if (it->second == off) {
        by_len.erase(it);
        return;
      }
  }
  int64_t take(int64_t len) {
    auto it = by_len.lower_bound(len);
    if (it == by_len.end()) {
      const int64_t off = top;
      top += len;
      return off;
    }
    const int64_t hl = it->first, off = it->second;
    by_len.erase(it);
    holes.erase(off);
    if (hl > len) {
      holes[off + len] = hl - len;
      by_len.emplace(hl - len, off + len);
    }
    return off;
  }
  void give(int64_t off, int64_t len) {
    auto it = holes.emplace(off, len).first;
    auto nx = std::next(it);
    if (nx != holes.end() && it->first + it->second == nx->first) {
      erase_len(nx->second, nx->first);
      it->second += nx->second;
      holes.erase(nx);
    }
    if (it != holes.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        erase_len(pv->second, pv->first);
        pv->second += it->second;
        holes.erase(it);
        it = pv;
      }
    }
    by_len.emplace(it->second, it->first);
  }
};

}  // namespace

// Update matrices (R x R) and compressed supernodes' work panels live only
// while the level program needs them.  Level L's assembly launch zeroes and
// writes the fronts and update matrices of its supernodes while it reads
// their children's update matrices, so level L's regions are taken while
// the children's are still held; the children's updates and level L's work
// panels are given back after level L, for the levels above to reuse (the
// program runs in level order on one stream).
void place_transient(Plan& plan) {
  FirstFit U, W;
  for (auto& lv : plan.levels) {
    for (int j : lv) {
      Supernode& s = plan.nodes[j];
      const int64_t c = s.ncols, R = s.R();
      s.off_update = R > 0 ? U.take(R * R) : -1;
      s.off_work_panel = s.compress ? W.take((c + R) * c) : -1;
    }
    for (int j : lv) {
      const Supernode& s = plan.nodes[j];
      for (int ch : s.children) {
        const Supernode& cs = plan.nodes[ch];
        if (cs.R() > 0) U.give(cs.off_update, (int64_t)cs.R() * cs.R());
      }
      if (s.compress) W.give(s.off_work_panel, (int64_t)(s.ncols + s.R()) * s.ncols);
    }
  }
  plan.update_elems = U.top;
  plan.work_elems = W.top;
}

void replan_ranks(Plan& plan, double diag_const) {
  const rsc_options& o = plan.opt;
  int64_t foff = 0, rng = 0;
  double flops = 0, apply = 0;
  const int64_t es = o.factor_precision == RSC_FP32 ? 4 : 8;
  int64_t dense_root_bytes = (int64_t)512 << 20;
  if (const char* e = std::getenv("RSC_DENSE_ROOT_BYTES")) dense_root_bytes = std::atoll(e);
  auto potrf_f = [](double c) { return c * c * c / 3.0 + c * c / 2.0 + c / 6.0; };
  for (auto& s : plan.nodes) {
    const int64_t c = s.ncols, R = s.R();
    if (!s.compress) {
      s.off_panel = foff;
      foff += (c + R) * c;
      flops += potrf_f((double)c) + (double)R * c * c + (double)c * R * (R + 1);
      apply += 2.0 * ((double)c * (c + 1) / 2 + (double)R * c);
      continue;
    }
    s.off_panel = -1;
    double tree_apply = 0;
    // diagonal tree: ranks, leaf/V/U storage, RNG offsets in draw order
    for (auto& h : s.hnodes) {
      if (h.leaf()) {
        int64_t b = h.r1 - h.r0;
        h.off_leaf = foff;
        foff += b * b;
        tree_apply += 2.0 * (double)b * (b + 1) / 2;
      } else {
        const HNode& L = s.hnodes[h.left];
        const HNode& Rt = s.hnodes[h.right];
        int rows = Rt.r1 - Rt.r0, cols = L.r1 - L.r0;
        h.rank = block_rank(rows, cols, diag_const, o.oversampling, o.min_rank,
                            o.max_rank);
        h.off_V = foff;
        foff += (int64_t)rows * h.rank;
        h.off_U = foff;
        foff += (int64_t)cols * h.rank;
        h.off_W = foff;
        foff += (int64_t)rows * h.rank;
        h.off_X = foff;
        foff += (int64_t)cols * h.rank;
        // forward reads U and W, backward V and X
        tree_apply += 2.0 * (double)h.rank * (rows + cols);
      }
    }
    // A root separator (no off-diagonal rows) small enough keeps its
    // diagonal inverse dense: the apply then needs two matrix-vector stages
    // there instead of two per tree height.
    s.off_dinv = -1;
    if (R == 0 && s.hnodes.size() > 1 && c * c * es <= dense_root_bytes) {
      s.off_dinv = foff;
      foff += c * c;
      apply += (double)c * (c + 1);
    } else {
      apply += tree_apply;
    }
    // draw order: traversal; explicit blocks draw in pre-order
    std::function<void(int)> pre = [&](int id) {
      HNode& h = s.hnodes[id];
      if (h.leaf()) return;
      int rows = s.hnodes[h.right].r1 - s.hnodes[h.right].r0;
      h.rng_off = rng;
      rng += (int64_t)rows * h.rank;
      pre(h.left);
      pre(h.right);
    };
    for (auto& st : s.hsteps) {
      HNode& h = s.hnodes[st.node];
      if (st.kind == HStep::EXPLICIT) {
        pre(st.node);
        int64_t b = h.r1 - h.r0;
        flops += potrf_f((double)b);
      } else if (st.kind == HStep::OFFDIAG) {
        int rows = s.hnodes[h.right].r1 - s.hnodes[h.right].r0;
        h.rng_off = rng;
        rng += (int64_t)rows * h.rank;
      } else {
        flops += potrf_f((double)(h.r1 - h.r0));
      }
    }
    for (auto& h : s.hnodes)
      if (!h.leaf()) {
        double rows = s.hnodes[h.right].r1 - s.hnodes[h.right].r0;
        double cols = s.hnodes[h.left].r1 - s.hnodes[h.left].r0;
        double r = h.rank;
        flops += 4 * 2.0 * rows * cols * r + 2.0 * cols * r * r - 2.0 * r * r * r / 3.0;
        flops += 2.0 * r * r * cols + 2.0 * rows * r * r + (double)r * rows * (rows + 1);
      }
    if (R > 0) {
      s.rank = block_rank((int)R, (int)c, o.rank_constant, o.oversampling,
                          o.min_rank, o.max_rank);
      s.off_V = foff;
      foff += R * s.rank;
      s.off_U = foff;
      foff += c * s.rank;
      s.off_Xoff = foff;
      foff += c * s.rank;
      s.rng_off = rng;
      rng += R * s.rank;
      double r = s.rank;
      flops += (double)R * c * c;  // L21 = F21 L11^-T (leaf-dominated bound)
      flops += 4 * 2.0 * R * c * r + 2.0 * c * r * r - 2.0 * r * r * r / 3.0;
      flops += 2.0 * r * r * c + 2.0 * R * r * r + r * (double)R * (R + 1);
      apply += 2.0 * r * (R + c);
    } else {
      s.rank = 0;
      s.off_V = s.off_U = s.off_Xoff = -1;
      s.rng_off = -1;
    }
  }
  plan.factor_elems = foff;
  place_transient(plan);
  plan.rng_count = rng;
  plan.factor_flops = flops;
  plan.apply_bytes_per_elem = apply;
}

void build_plan(const Csc& A, const double* xyz, const rsc_options& opt,
                const NDResult* nd_in, Plan& plan) {
  plan = Plan();
  plan.opt = opt;
  const int n = A.n;
  plan.n = n;
  auto tp = std::chrono::steady_clock::now();
  auto phase = [&](const char* what) {
    if (!opt.verbose) return;
    auto now = std::chrono::steady_clock::now();
    std::printf("symbolic %s: %.3f s\n", what,
                std::chrono::duration<double>(now - tp).count());
    tp = now;
  };
  NDResult nd = nd_in ? *nd_in : nested_dissection(A, xyz, opt.nd_small);
  phase("nested dissection");
  std::vector<int> pinv(n);
  for (int k = 0; k < n; ++k) pinv[nd.perm[k]] = k;
  std::vector<int> parent = elimination_tree(A, pinv);
  std::vector<int> post(n);
  std::iota(post.begin(), post.end(), 0);
  std::vector<int> cc = column_counts(A, pinv, parent, post);
  phase("etree and column counts");
  std::vector<Separator> seps = separators_from_nd(nd, n);
  plan.specs = build_relaxed_supernodes(parent, cc, seps, nd.cparent,
                                        opt.max_dense_block);
  phase("relaxed supernodes");
  plan.perm = nd.perm;
  plan.diag_blocks = reorder_nested_separators(
      plan.specs, xyz, opt.max_diagonal_block, plan.perm,
      opt.reorder_separators != 0 && xyz != nullptr);
  plan.pinv.assign(n, 0);
  for (int k = 0; k < n; ++k) plan.pinv[plan.perm[k]] = k;
  phase("separator trees");

  // supernodes and row sets
  const int ns = (int)plan.specs.size();
  plan.nodes.resize(ns);
  plan.col2node.assign(n, 0);
  for (int s = 0; s < ns; ++s) {
    Supernode& sn = plan.nodes[s];
    sn.first = plan.specs[s].first;
    sn.ncols = plan.specs[s].ncols();
    sn.compress = plan.specs[s].compress;
    for (int k = sn.first; k <= plan.specs[s].last; ++k) plan.col2node[k] = s;
  }
  // The supernode tree from the elimination tree: the parent of s is the
  // supernode holding the etree parent of s's last column (the smallest row
  // of s's structure below it).  Row sets then go level by level (children
  // first), the supernodes of one level in parallel; each is the union of
  // its columns' entries of A and its children's rows below its last column.
  std::vector<int> sparent(ns, -1), height(ns, 0);
  for (int s = 0; s < ns; ++s) {
    const int lastcol = plan.nodes[s].first + plan.nodes[s].ncols - 1;
    const int pc = parent[lastcol];
    if (pc >= 0) {
      const int ps = plan.col2node[pc];
      if (ps <= s) throw std::logic_error("supernode tree is not in postorder");
      sparent[s] = ps;
      plan.nodes[ps].children.push_back(s);
    }
  }
  int hmax = 0;
  for (int s = 0; s < ns; ++s)
    if (sparent[s] >= 0) {
      height[sparent[s]] = std::max(height[sparent[s]], height[s] + 1);
      hmax = std::max(hmax, height[sparent[s]]);
    }
  std::vector<std::vector<int>> by_height(hmax + 1);
  for (int s = 0; s < ns; ++s) by_height[height[s]].push_back(s);
  std::vector<std::vector<int>> marks(stream_threads());
  for (const std::vector<int>& lv : by_height) {
    parallel_for((int64_t)lv.size(), 64, [&](int64_t b, int64_t e, int t) {
      std::vector<int>& mark = marks[t];
      if (mark.empty()) mark.assign(n, -1);
      for (int64_t x = b; x < e; ++x) {
        const int s = lv[x];
        Supernode& sn = plan.nodes[s];
        const int last = sn.first + sn.ncols - 1;
        std::vector<int>& rows = sn.rows;
        for (int k = sn.first; k <= last; ++k) {
          const int o = plan.perm[k];
          for (int p = A.p[o]; p < A.p[o + 1]; ++p) {
            const int r = plan.pinv[A.i[p]];
            if (r > last && mark[r] != s) {
              mark[r] = s;
              rows.push_back(r);
            }
          }
        }
        for (int ch : sn.children)
          for (int r : plan.nodes[ch].rows)
            if (r > last && mark[r] != s) {
              mark[r] = s;
              rows.push_back(r);
            }
        std::sort(rows.begin(), rows.end());
        sn.parent = sparent[s];
        if ((rows.empty() ? -1 : plan.col2node[rows[0]]) != sparent[s])
          throw std::logic_error("row sets disagree with the elimination tree");
      }
    });
  }
  phase("row sets");
  // levels (height) and relative maps into the parent's front
  int maxlev = 0;
  for (int s = 0; s < ns; ++s) {
    Supernode& sn = plan.nodes[s];
    int lv = 0;
    for (int ch : sn.children) lv = std::max(lv, plan.nodes[ch].level + 1);
    sn.level = lv;
    maxlev = std::max(maxlev, lv);
    plan.max_front = std::max(plan.max_front, sn.ncols + sn.R());
  }
  plan.levels.assign(maxlev + 1, {});
  for (int s = 0; s < ns; ++s) plan.levels[plan.nodes[s].level].push_back(s);
  {
    // position of every row in the parent's front, one parent at a time
    std::vector<int> pos(n, -1);
    for (int p = 0; p < ns; ++p) {
      const Supernode& P = plan.nodes[p];
      if (P.children.empty()) continue;
      for (size_t q = 0; q < P.rows.size(); ++q) pos[P.rows[q]] = P.ncols + (int)q;
      const int plast = P.first + P.ncols - 1;
      for (int ch : P.children) {
        Supernode& sn = plan.nodes[ch];
        sn.relmap.resize(sn.rows.size());
        for (size_t a = 0; a < sn.rows.size(); ++a) {
          const int r = sn.rows[a];
          if (r <= plast) {
            sn.relmap[a] = r - P.first;
          } else {
            if (pos[r] < 0)
              throw std::runtime_error("row structure is not closed under the parent");
            sn.relmap[a] = pos[r];
          }
        }
      }
      for (int r : P.rows) pos[r] = -1;
    }
  }
  phase("levels and relative maps");
  // diagonal block trees of compressed supernodes
  for (int s = 0; s < ns; ++s) {
    Supernode& sn = plan.nodes[s];
    if (!sn.compress) continue;
    build_hodlr(sn, plan.diag_blocks[s], opt.explicit_block_threshold);
  }
  phase("diagonal trees");
  replan_ranks(plan, opt.diagonal_rank_constant);
  phase("ranks and storage");
}

}  // namespace rsc
