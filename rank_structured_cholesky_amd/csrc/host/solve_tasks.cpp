// Host construction of the apply's task lists (declarations and the
// meaning of each task in solve_tasks.h).
#include "solve_tasks.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <map>
#include <queue>
#include <stdexcept>
#include <utility>

namespace rsc {

void build_subtrees(const Plan& P, int64_t max_elems, Subtrees& S, bool multi,
                    int max_tasks) {
  const int ns = (int)P.nodes.size();
  S.of.assign(ns, -1);
  S.ptr.assign(1, 0);
  S.nodes.clear();
  S.root.clear();
  S.rptr.assign(1, 0);
  S.roots.clear();
  S.cols.clear();
  if (max_elems <= 0) return;
  // children are numbered before their parents (postordered supernodes)
  std::vector<char> ok(ns, 0);
  std::vector<int64_t> cost(ns, 0);
  for (int j = 0; j < ns; ++j) {
    const Supernode& s = P.nodes[j];
    bool e = !s.compress && (multi ? multi_subtree_node(s.ncols, s.R())
                                   : s.ncols <= kSubtreeMaxCols && s.R() <= kSubtreeMaxRows);
    int64_t c = (int64_t)s.ncols * (s.ncols + s.R());
    for (int ch : s.children) {
      if (ch >= j) throw std::logic_error("supernodes not in postorder");
      e = e && ok[ch];
      c += cost[ch];
    }
    ok[j] = e && c <= max_elems;
    cost[j] = c;
  }
  // one tree per maximal subtree: its nodes in increasing (post) order
  struct Tree {
    std::vector<int> nodes, roots;
    double work = 0;  // supernodes plus front entries / 150 (relative cost)
  };
  std::vector<Tree> trees;
  for (int j = 0; j < ns; ++j) {
    const int par = P.nodes[j].parent;
    if (!ok[j] || (par >= 0 && ok[par])) continue;
    Tree t;
    t.roots.push_back(j);
    // depth-first postorder, children in increasing order: every supernode
    // follows its whole subtree, so the processed supernodes of a task form
    // a stack whose top entries are the next supernode's children
    std::vector<std::pair<int, int>> dfs{{j, 0}};  // (node, next child)
    while (!dfs.empty()) {
      auto& top = dfs.back();
      const std::vector<int>& ch = P.nodes[top.first].children;
      if (top.second < (int)ch.size()) {
        const int c = ch[top.second++];
        dfs.push_back({c, 0});
        continue;
      }
      const int q = top.first;
      dfs.pop_back();
      t.nodes.push_back(q);
      t.work += 1.0 + (P.nodes[q].ncols + P.nodes[q].R()) / 150.0;
    }
    trees.push_back(std::move(t));
  }
  if (max_tasks > 0 && (int)trees.size() > max_tasks) {
    // merge the two smallest tasks until max_tasks remain (a task runs its
    // trees one after the other)
    using Item = std::pair<double, int>;
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
    for (int i = 0; i < (int)trees.size(); ++i) heap.push({trees[i].work, i});
    std::vector<char> live(trees.size(), 1);
    int count = (int)trees.size();
    while (count > max_tasks && heap.size() >= 2) {
      const int x = heap.top().second;
      heap.pop();
      const int y = heap.top().second;
      heap.pop();
      Tree& tx = trees[x];
      Tree& ty = trees[y];
      tx.nodes.insert(tx.nodes.end(), ty.nodes.begin(), ty.nodes.end());
      tx.roots.insert(tx.roots.end(), ty.roots.begin(), ty.roots.end());
      tx.work += ty.work;
      live[y] = 0;
      heap.push({tx.work, x});
      --count;
    }
    std::vector<Tree> kept;
    for (size_t i = 0; i < trees.size(); ++i)
      if (live[i]) kept.push_back(std::move(trees[i]));
    trees.swap(kept);
    // tasks in the order of their first root
    std::sort(trees.begin(), trees.end(),
              [](const Tree& u, const Tree& v) { return u.roots[0] < v.roots[0]; });
  }
  if (multi) {
    // the device runs a task's supernodes in list order with the processed
    // ones on a stack: each supernode's children must be the top entries,
    // in increasing order, and the stack must fit; otherwise no subtree
    // tasks (every supernode runs as ordinary tasks)
    size_t depth = 0;
    std::vector<int> stk;
    for (const Tree& t : trees) {
      stk.clear();
      for (int q : t.nodes) {
        const std::vector<int>& ch = P.nodes[q].children;
        if (stk.size() < ch.size() || !std::equal(ch.begin(), ch.end(), stk.end() - ch.size()))
          throw std::logic_error("subtree task: children are not the stack's top entries");
        stk.resize(stk.size() - ch.size());
        stk.push_back(q);
        depth = std::max(depth, stk.size());
      }
    }
    if (depth > (size_t)kMsStack) {
      S.of.assign(ns, -1);
      return;
    }
  }
  for (const Tree& t : trees) {
    const int id = (int)S.root.size();
    S.root.push_back(t.roots[0]);
    S.roots.insert(S.roots.end(), t.roots.begin(), t.roots.end());
    S.rptr.push_back((int)S.roots.size());
    for (int q : t.nodes) S.of[q] = id;
    S.nodes.insert(S.nodes.end(), t.nodes.begin(), t.nodes.end());
    S.ptr.push_back((int)S.nodes.size());
    // column range of the task when its supernodes' columns are contiguous
    int lo = INT_MAX, hi = 0;
    int64_t tot = 0;
    for (int q : t.nodes) {
      lo = std::min(lo, P.nodes[q].first);
      hi = std::max(hi, P.nodes[q].first + P.nodes[q].ncols);
      tot += P.nodes[q].ncols;
    }
    S.cols.push_back(lo);
    S.cols.push_back(tot == (int64_t)hi - lo ? hi - lo : -1);
  }
}

void build_solve_layout(const Plan& P, SolveLayout& L, int64_t subtree_elems,
                        int64_t multi_subtree_elems, int multi_max_tasks) {
  const int ns = (int)P.nodes.size();
  L.children.clear();
  L.relmap.clear();
  L.rows.clear();
  std::vector<int> child_off(ns);
  std::vector<int64_t> relmap_off(ns), rows_off(ns);
  for (int j = 0; j < ns; ++j) {
    const Supernode& s = P.nodes[j];
    child_off[j] = (int)L.children.size();
    L.children.insert(L.children.end(), s.children.begin(), s.children.end());
    relmap_off[j] = (int64_t)L.relmap.size();
    L.relmap.insert(L.relmap.end(), s.relmap.begin(), s.relmap.end());
    if (s.relmap.empty()) L.relmap.resize(L.relmap.size() + s.rows.size(), 0);
    rows_off[j] = (int64_t)L.rows.size();
    L.rows.insert(L.rows.end(), s.rows.begin(), s.rows.end());
  }
  L.vn.assign(ns, dev::VecNode{});
  L.tv_node.assign(ns, -1);
  L.fv_off.assign(ns, 0);
  L.tv_h.assign(ns, {});
  int64_t fvo = 0, tvo = 0;
  for (int j = 0; j < ns; ++j) {
    const Supernode& s = P.nodes[j];
    dev::VecNode& d = L.vn[j];
    d.first = s.first;
    d.ncols = s.ncols;
    d.R = s.R();
    d.nchild = (int)s.children.size();
    d.child_off = child_off[j];
    d.range_off = 0;
    d.fv_off = fvo;
    L.fv_off[j] = fvo;
    d.rows_off = rows_off[j];
    d.relmap_off = relmap_off[j];
    fvo += (int64_t)s.ncols + s.R();
    if (s.compress) {
      L.tv_node[j] = tvo;
      tvo += s.rank;
      L.tv_h[j].assign(s.hnodes.size(), -1);
      for (size_t q = 0; q < s.hnodes.size(); ++q)
        if (!s.hnodes[q].leaf()) {
          L.tv_h[j][q] = tvo;
          tvo += s.hnodes[q].rank;
        }
    }
  }
  L.asm_ranges.clear();
  for (int j = 0; j < ns; ++j) {
    const Supernode& s = P.nodes[j];
    if (s.children.empty()) continue;
    L.vn[j].range_off = (int)(L.asm_ranges.size() / 2);
    const int len = s.ncols + s.R();
    for (int p0 = 0; p0 < len; p0 += kSolveChunk) {
      const int p1 = std::min(p0 + kSolveChunk, len);
      for (int ch : s.children) {
        const std::vector<int>& rm = P.nodes[ch].relmap;
        L.asm_ranges.push_back((int)(std::lower_bound(rm.begin(), rm.end(), p0) - rm.begin()));
        L.asm_ranges.push_back((int)(std::lower_bound(rm.begin(), rm.end(), p1) - rm.begin()));
      }
    }
  }
  if (L.asm_ranges.empty()) L.asm_ranges.assign(2, 0);
  L.fv_stride = std::max<int64_t>(fvo, 1);
  L.tv_stride = std::max<int64_t>(tvo, 1);
  L.panel.assign(ns, -1);
  for (int j = 0; j < ns; ++j)
    if (!P.nodes[j].compress) L.panel[j] = P.nodes[j].off_panel;
  build_subtrees(P, subtree_elems, L.sub);
  build_subtrees(P, multi_subtree_elems, L.sub_multi, true, multi_max_tasks);
}

// ---- per-node stages ------------------------------------------------------
//
// The factor is held in partitioned-inverse form: an uncompressed panel
// stores [Lc^{-T} (upper); Zp = L21 Lc^{-1}], a compressed node its HODLR
// diagonal (leaves L^{-T}, W/X helpers), V and X_off = L_D^{-T} U.  Per
// supernode (b = front vector fv = [fv_c; fv_R], w = permuted vector,
// z = L^{-T} y):
//   A  (forward, towards the parent: only what the parent's assembly needs)
//      A0  fv_R -= Zp fv_c (ROW) or tv = X_off^T fv_c (COL)
//      A1  compressed: fv_R -= V tv (ROW)
//   B  (forward, off the path to the parent; starts after the assembly)
//      y_c = Lc^{-1} fv_c (uncompressed: COL on Lc^{-T}; compressed: HODLR
//      leaves, then heights); then z_c = Lc^{-T} y_c (ROW on the stored upper
//      inverse; HODLR leaves and heights with V_M / X_M)
//   C  (backward, after the gather fv_R = x[rows], x_c = z_c)
//      uncompressed: x_c -= Zp^T fv_R (COL); compressed: tv = V^T fv_R (COL),
//      then x_c -= X_off tv (ROW)
namespace {

struct Mv {
  bool col;
  int64_t mat;
  int ld, rows, cols, op, xs;
  int64_t xo;
  int ys;
  int64_t yo;
  bool upper;
};
using Stage = std::vector<Mv>;

struct NodeStages {
  std::vector<Stage> A, B, C;
};

constexpr int W = 0, FV = 1, TV = 2, ZB = 3;

NodeStages node_stages(const Plan& P, const SolveLayout& L, int j) {
  const Supernode& s = P.nodes[j];
  const int64_t fo = L.fv_off[j];
  const int c = s.ncols, R = s.R();
  NodeStages ns;
  int maxh = 0;
  for (auto& h : s.hnodes) maxh = std::max(maxh, h.height);
  if (!s.compress) {
    if (R > 0) ns.A.push_back({{false, s.off_panel + c, c + R, R, c, 1, FV, fo, FV, fo + c, false}});
    ns.B.push_back({{true, s.off_panel, c + R, c, c, 0, FV, fo, W, s.first, true}});
    ns.B.push_back({{false, s.off_panel, c + R, c, c, 0, W, s.first, ZB, s.first, true}});
    if (R > 0)
      ns.C.push_back({{true, s.off_panel + c, c + R, R, c, 1, FV, fo + c, W, s.first, false}});
    return ns;
  }
  if (s.off_dinv >= 0) {
    // dense L_D^{-T} (root separator, no off-diagonal rows)
    ns.B.push_back({{true, s.off_dinv, c, c, c, 0, FV, fo, W, s.first, true}});
    ns.B.push_back({{false, s.off_dinv, c, c, c, 0, W, s.first, ZB, s.first, true}});
    return ns;
  }
  const int64_t tvn = L.tv_node[j];
  const std::vector<int64_t>& tvh = L.tv_h[j];
  if (R > 0 && s.rank > 0) {
    ns.A.push_back({{true, s.off_Xoff, c, c, s.rank, 0, FV, fo, TV, tvn, false}});
    ns.A.push_back({{false, s.off_V, R, R, s.rank, 1, TV, tvn, FV, fo + c, false}});
  }
  // y_c: HODLR leaves, then heights (forward)
  Stage yl;
  for (auto& h : s.hnodes)
    if (h.leaf())
      yl.push_back({true, h.off_leaf, h.r1 - h.r0, h.r1 - h.r0, h.r1 - h.r0, 0, FV, fo + h.r0,
                    W, s.first + h.r0, true});
  ns.B.push_back(yl);
  for (int hh = 1; hh <= maxh; ++hh) {
    Stage c1, c2;
    for (size_t q = 0; q < s.hnodes.size(); ++q) {
      const HNode& h = s.hnodes[q];
      if (h.leaf() || h.height != hh) continue;
      const HNode& A = s.hnodes[h.left];
      const HNode& B = s.hnodes[h.right];
      c1.push_back({true, h.off_U, A.r1 - A.r0, A.r1 - A.r0, h.rank, 0, W, s.first + A.r0, TV,
                    tvh[q], false});
      c2.push_back({false, h.off_W, B.r1 - B.r0, B.r1 - B.r0, h.rank, 1, TV, tvh[q], W,
                    s.first + B.r0, false});
    }
    ns.B.push_back(c1);
    ns.B.push_back(c2);
  }
  // z_c = L_D^{-T} y_c: leaves, then heights (transposed)
  Stage zl;
  for (auto& h : s.hnodes)
    if (h.leaf())
      zl.push_back({false, h.off_leaf, h.r1 - h.r0, h.r1 - h.r0, h.r1 - h.r0, 0, W,
                    s.first + h.r0, ZB, s.first + h.r0, true});
  ns.B.push_back(zl);
  for (int hh = 1; hh <= maxh; ++hh) {
    Stage c1, c2;
    for (size_t q = 0; q < s.hnodes.size(); ++q) {
      const HNode& h = s.hnodes[q];
      if (h.leaf() || h.height != hh) continue;
      const HNode& A = s.hnodes[h.left];
      const HNode& B = s.hnodes[h.right];
      c1.push_back({true, h.off_V, B.r1 - B.r0, B.r1 - B.r0, h.rank, 0, ZB, s.first + B.r0, TV,
                    tvh[q], false});
      c2.push_back({false, h.off_X, A.r1 - A.r0, A.r1 - A.r0, h.rank, 1, TV, tvh[q], ZB,
                    s.first + A.r0, false});
    }
    ns.B.push_back(c1);
    ns.B.push_back(c2);
  }
  if (R > 0 && s.rank > 0) {
    ns.C.push_back({{true, s.off_V, R, R, s.rank, 0, FV, fo + c, TV, tvn, false}});
    ns.C.push_back({{false, s.off_Xoff, c, c, s.rank, 1, TV, tvn, W, s.first, false}});
  }
  return ns;
}

// Matrix-vector specs -> GvTasks.  COL: 32 columns per task, more than CHUNK
// rows split into partial sums (slot per chunk) plus one reduction task;
// ROW: 64 rows per task.
struct TaskEmitter {
  int64_t slot = 0;
  int CHUNK = 512;     // rows per COL task (longer columns: partial sums)
  int row_split = 0;  // > 0: ROW tasks longer than this many columns are split
  // fuse_tv: a split COL task that sets a rank-sized vector (tv) gets no
  // reduce tasks; the ROW tasks that read the vector next sum its partial
  // slots (first slot and slots per 32-entry block, by tv offset)
  bool fuse_tv = false;
  std::map<int64_t, std::pair<int, int>> tv_parts;
  // A product of at least wide_elems matrix elements is split finer (down to
  // 128 rows per COL task, 256 columns per ROW task) while it still yields at
  // most half as many tasks as there are workgroups: the few huge supernodes
  // at the top of the tree run nearly alone, and there a stage's time is its
  // longest task's.
  int workers = 0;
  int64_t wide_elems = int64_t(1) << 20;
  void emit(const Mv& m, std::vector<dev::GvTask>& main, std::vector<dev::GvTask>& red) {
    const int chunk0 = CHUNK, split0 = row_split;
    if (workers > 0 && (int64_t)m.rows * m.cols >= wide_elems) {
      if (m.col) {
        auto count = [&](int ch) {
          int64_t n = 0;
          for (int j0 = 0; j0 < m.cols; j0 += 32) {
            const int rr = m.upper ? std::min(m.rows, j0 + 32) : m.rows;
            n += (rr + ch - 1) / ch;
          }
          return n;
        };
        while (CHUNK > 128 && 2 * count(CHUNK / 2) <= workers) CHUNK /= 2;
      } else if (row_split > 0) {
        auto count = [&](int rs) {
          int64_t n = 0;
          for (int i0 = 0; i0 < m.rows; i0 += 64) {
            const int cc = m.upper ? m.cols - i0 : m.cols;
            n += (cc + rs - 1) / rs;
          }
          return n;
        };
        while (row_split > 256 && 2 * count(row_split / 2) <= workers) row_split /= 2;
      }
    }
    emit_split(m, main, red);
    CHUNK = chunk0;
    row_split = split0;
  }
  // The matrix-core ROW body addresses its block with 32-bit byte offsets
  // from the block's start (8-byte elements at most).
  static void check_row_span(const dev::GvTask& t) {
    const int64_t span = ((int64_t)(t.cols - 1) * t.ld + t.rows) * 8;
    if (span >= (int64_t(1) << 32))
      throw std::runtime_error("solve task: a row block spans 4 GiB or more");
  }
  void emit_split(const Mv& m, std::vector<dev::GvTask>& main, std::vector<dev::GvTask>& red) {
    if (m.col) {
      const bool fuse = fuse_tv && m.ys == TV && m.op == 0 && !m.upper;
      if (fuse) tv_parts.erase(m.yo);
      const int64_t first_slot = slot;
      for (int j0 = 0; j0 < m.cols; j0 += 32) {
        const int rr = m.upper ? std::min(m.rows, j0 + 32) : m.rows;
        if (rr <= 0) continue;
        dev::GvTask t{};
        t.xpart = -1;
        t.ld = m.ld;
        t.cols = std::min(32, m.cols - j0);
        t.kind = 1;
        t.op = m.op;
        t.xsel = m.xs;
        t.ysel = m.ys;
        t.yoff = m.yo + j0;
        if (rr <= CHUNK) {
          t.mat = m.mat + (int64_t)j0 * m.ld;
          t.rows = rr;
          t.xoff = m.xo;
          t.part = -1;
          main.push_back(t);
          continue;
        }
        const int nsub = (rr + CHUNK - 1) / CHUNK;
        for (int s2 = 0; s2 < nsub; ++s2) {
          dev::GvTask u = t;
          u.mat = m.mat + (int64_t)j0 * m.ld + (int64_t)s2 * CHUNK;
          u.rows = std::min(CHUNK, rr - s2 * CHUNK);
          u.xoff = m.xo + (int64_t)s2 * CHUNK;
          u.part = (int)(slot + s2);
          main.push_back(u);
        }
        if (fuse) {
          // every 32-column block of this vector is split alike (same rows)
          tv_parts[m.yo] = {(int)first_slot, nsub};
          slot += nsub;
          continue;
        }
        dev::GvTask rd{};
        rd.xpart = -1;
        rd.kind = 2;
        rd.part = (int)slot;
        rd.rows = nsub;
        rd.cols = t.cols;
        rd.op = m.op;
        rd.ysel = m.ys;
        rd.yoff = m.yo + j0;
        red.push_back(rd);
        slot += nsub;
      }
    } else {
      int xpart = -1, xnsub = 0;
      if (fuse_tv && m.xs == TV && !m.upper && row_split % 32 == 0) {
        const auto it = tv_parts.find(m.xo);
        if (it != tv_parts.end()) {
          xpart = it->second.first;
          xnsub = it->second.second;
        }
      }
      for (int i0 = 0; i0 < m.rows; i0 += 64) {
        dev::GvTask t{};
        t.xpart = xpart;
        t.xnsub = xnsub;
        t.mat = m.mat + i0 + (m.upper ? (int64_t)i0 * m.ld : 0);
        t.ld = m.ld;
        t.rows = std::min(64, m.rows - i0);
        t.cols = m.upper ? m.cols - i0 : m.cols;
        t.kind = 0;
        t.op = m.op;
        t.xsel = m.xs;
        t.ysel = m.ys;
        t.part = -1;
        t.xoff = m.xo + (m.upper ? i0 : 0);
        t.yoff = m.yo + i0;
        if (row_split <= 0 || t.cols <= row_split) {
          check_row_span(t);
          main.push_back(t);
          continue;
        }
        // long rows: column chunks write partial sums, rows 0..31 of chunk
        // s2 into slot base + 2 s2 and rows 32..63 into base + 2 s2 + 1;
        // one reduction per half sums its slots (stride 2)
        const int nsub = (t.cols + row_split - 1) / row_split;
        for (int s2 = 0; s2 < nsub; ++s2) {
          dev::GvTask u = t;
          const int c0 = s2 * row_split;
          u.mat = t.mat + (int64_t)c0 * t.ld;
          u.cols = std::min(row_split, t.cols - c0);
          u.xoff = t.xoff + c0;
          if (xpart >= 0) u.xpart = xpart + (c0 / 32) * xnsub;
          u.part = (int)(slot + 2 * s2);
          check_row_span(u);
          main.push_back(u);
        }
        for (int h = 0; h < 2; ++h) {
          const int nr = std::min(32, t.rows - 32 * h);
          if (nr <= 0) continue;
          dev::GvTask rd{};
          rd.xpart = -1;
          rd.kind = 2;
          rd.part = (int)(slot + h);
          rd.rows = nsub;
          rd.cols = nr;
          rd.ld = 2;
          rd.op = m.op;
          rd.ysel = m.ys;
          rd.yoff = t.yoff + 32 * h;
          red.push_back(rd);
        }
        slot += 2 * nsub;
      }
    }
  }
};

}  // namespace

// Level-scheduled form: per level and direction a sequence of phases, each
// one launch of independent tasks; stage k of every node of the level runs
// in phase k (a phase's split column tasks are followed by a launch of their
// reductions).
void build_level_tasks(const Plan& P, const SolveLayout& L, LevelTasks& out) {
  std::vector<dev::GvTask> gv;
  out.fwd.assign(P.levels.size(), {});
  out.bwd.assign(P.levels.size(), {});
  TaskEmitter em;
  auto phases = [&](std::vector<SolvePhase>& ph, const std::vector<std::vector<Stage>>& per) {
    size_t nst = 0;
    for (auto& v : per) nst = std::max(nst, v.size());
    for (size_t k = 0; k < nst; ++k) {
      std::vector<dev::GvTask> main, red;
      for (auto& v : per)
        if (k < v.size())
          for (const Mv& m : v[k]) em.emit(m, main, red);
      if (!main.empty()) {
        ph.push_back({2, (int)gv.size(), (int)main.size()});
        gv.insert(gv.end(), main.begin(), main.end());
      }
      if (!red.empty()) {
        ph.push_back({2, (int)gv.size(), (int)red.size()});
        gv.insert(gv.end(), red.begin(), red.end());
      }
    }
  };
  const int nsub = (int)L.sub.root.size();
  out.level_nodes.assign(P.levels.size(), {});
  for (size_t lv = 0; lv < P.levels.size(); ++lv) {
    std::vector<std::vector<Stage>> A, B, C;
    for (int j : P.levels[lv]) {
      if (L.sub.of[j] >= 0) continue;
      out.level_nodes[lv].push_back(j);
      NodeStages ns = node_stages(P, L, j);
      A.push_back(ns.A);
      B.push_back(ns.B);
      C.push_back(ns.C);
    }
    if (lv == 0 && nsub > 0) out.fwd[lv].push_back({4, 0, nsub});
    if (!out.level_nodes[lv].empty()) {
      out.fwd[lv].push_back({0, 0, 0});
      phases(out.fwd[lv], A);
      phases(out.fwd[lv], B);
      out.bwd[lv].push_back({1, 0, 0});
      phases(out.bwd[lv], C);
    }
    if (lv == 0 && nsub > 0) out.bwd[lv].push_back({5, 0, nsub});
  }
  // large tasks use the high-MLP kernel
  for (auto* dir : {&out.fwd, &out.bwd})
    for (auto& lv : *dir) {
      std::vector<SolvePhase> sp;
      for (const SolvePhase& p : lv) {
        if (p.kind < 2 || p.kind > 3) {
          sp.push_back(p);
          continue;
        }
        auto big = [&](const dev::GvTask& t) {
          return t.kind == 0 ? t.cols > 64 : (t.kind == 1 && t.rows > 128);
        };
        std::stable_partition(gv.begin() + p.first, gv.begin() + p.first + p.count,
                              [&](const dev::GvTask& t) { return !big(t); });
        int mid = p.first;
        while (mid < p.first + p.count && !big(gv[mid])) ++mid;
        if (mid > p.first) sp.push_back({2, p.first, mid - p.first});
        if (p.first + p.count > mid) sp.push_back({3, mid, p.first + p.count - mid});
      }
      lv.swap(sp);
    }
  out.gv.swap(gv);
  out.slots = em.slot;
}

// Dataflow form: per supernode three chains of groups (A towards the
// parent, B for y and z, C backward).  A node's assembly waits on its
// children's last A group; chain B starts after the assembly; the backward gather
// waits on the parent's last C group (nothing for a root) and on the node's
// own chain B.  The scatter waits on every node's last C group.  Tasks are
// emitted level by level, so the list is in topological order.
namespace {

// Estimated duration of one task (ns) on one of many busy workgroups, fitted
// to per-task timings of the dataflow kernel on MI355X (64^3 Poisson, one
// right-hand side, scalar bodies; 128^3, sixteen, matrix-core bodies): a fixed
// part (ticket, waits, barriers, descriptor and first-data round trips, larger
// for the matrix-core bodies) plus terms in the matrix bytes, the 64-row /
// 128-column blocks of a matrix-vector task, the vector entries moved and the
// children or supernodes visited.
struct CostModel {
  const Plan& P;
  const SolveLayout& L;
  const Subtrees& SB;
  const std::vector<dev::GvTask>& gv;
  const TicketModel& m;
  double us(const dev::DfTask& t) const {
    const bool multi = m.nrhs > 1;
    const double k = std::max(1, m.nrhs), e = m.elem_bytes;
    const double passes = multi ? std::ceil(k / 16.0) : 1.0;
    switch (t.type) {
      case 0: {
        const dev::VecNode& v = L.vn[t.a];
        const int G = std::max(1, m.chunk_group);
        const int nch = (v.ncols + v.R + kSolveChunk - 1) / kSolveChunk;
        double us = 0;
        for (int ch = t.b * G; ch < std::min(nch, (t.b + 1) * G); ++ch) {
          double ent = std::min<int64_t>(kSolveChunk, v.ncols + v.R - (int64_t)ch * kSolveChunk);
          for (int q = 0; q < v.nchild; ++q) {
            const int64_t r = 2 * ((int64_t)v.range_off + (int64_t)ch * v.nchild + q);
            ent += L.asm_ranges[r + 1] - L.asm_ranges[r];
          }
          us += multi ? 2.4 + 1.2e-3 * ent * k + 3.6 * v.nchild
                      : 0.4 + 2.1e-3 * ent * k + 0.8 * v.nchild;
        }
        return us;
      }
      case 1: {
        const dev::VecNode& v = L.vn[t.a];
        const int G = std::max(1, m.chunk_group);
        const int nch = (v.ncols + v.R + kSolveChunk - 1) / kSolveChunk;
        double us = 0;
        for (int ch = t.b * G; ch < std::min(nch, (t.b + 1) * G); ++ch) {
          const double ent =
              std::min<int64_t>(kSolveChunk, v.ncols + v.R - (int64_t)ch * kSolveChunk);
          us += multi ? 1.9 + 1.6e-3 * ent * k : 1.3 + 1.2e-3 * ent * k;
        }
        return us;
      }
      case 2: {
        const dev::GvTask& g = gv[t.a];
        if (g.kind == 2) return multi ? 8.7 : 2.6;
        const double kb = (double)g.rows * g.cols * e / 1000.0;
        if (g.kind == 0)
          return multi ? 5.3 + 0.079 * kb * passes + 3.8 * std::ceil(g.cols / 128.0) * passes
                       : 5.5 + 0.044 * kb;
        return multi ? 6.5 + 0.036 * kb * passes + 1.9 * std::ceil(g.rows / 64.0) * passes
                     : 6.1 + 0.10 * kb;
      }
      case 3:
        return multi ? 10.0 * k : 6.0;
      default: {
        double kb = 0, nodes = 0, front = 0;
        for (int q = SB.ptr[t.a]; q < SB.ptr[t.a + 1]; ++q) {
          const Supernode& s = P.nodes[SB.nodes[q]];
          kb += (double)s.ncols * (s.ncols + s.R()) * e / 1000.0;
          front += s.ncols + s.R();
          nodes += 1;
        }
        if (multi)  // sixteen right-hand sides per pass, fronts in LDS
          return t.type == 4 ? (11.8 * nodes + 0.080 * front) * passes
                             : (2.07 * nodes + 0.112 * front) * passes;
        const double sp = std::ceil(k / kSubtreeRhs);
        return t.type == 4 ? 10.2 + (4.68 * nodes + 0.107 * kb) * sp
                           : 15.9 + (3.75 * nodes + 0.22 * kb) * sp;
      }
    }
  }
  int64_t ns(const dev::DfTask& t) const { return std::max<int64_t>(1, (int64_t)(1000.0 * us(t))); }
};

// Ticket order by list scheduling.  Workgroups take tickets in list order,
// so a ticket whose task is not ready holds a workgroup idle; in level
// order, the long chains of HODLR products that only the backward pass needs
// fill the ticket window ahead of the next level's critical assembly.  The
// list is replayed on `workers` simulated workgroups with the estimated
// costs: among the ready tasks the one with the longest estimated path to
// the end goes first, and the tickets follow the simulated start order.  A
// task starts only after every group it waits on has completed, so the
// order stays topological (every wait ends).
double order_by_list_schedule(std::vector<dev::DfTask>& tasks, const std::vector<unsigned>& need,
                              const std::vector<int>& gwoff, const std::vector<int>& gw,
                              const CostModel& cm, int workers) {
  const int ng = (int)need.size();
  const int nt = (int)tasks.size();
  std::vector<int64_t> cost(nt);
  for (int i = 0; i < nt; ++i) cost[i] = cm.ns(tasks[i]);
  // tasks per primary group, in list order
  std::vector<int> tptr(ng + 1, 0), tlist(nt);
  for (const auto& t : tasks) ++tptr[t.group + 1];
  for (int g = 0; g < ng; ++g) tptr[g + 1] += tptr[g];
  {
    std::vector<int> fill(tptr.begin(), tptr.end() - 1);
    for (int i = 0; i < nt; ++i) tlist[fill[tasks[i].group]++] = i;
  }
  // groups waiting on each group (one entry per wait, counting sort)
  std::vector<int> wptr(ng + 1, 0), wlist(gw.size());
  for (int q : gw) ++wptr[q + 1];
  for (int g = 0; g < ng; ++g) wptr[g + 1] += wptr[g];
  {
    std::vector<int> fill(wptr.begin(), wptr.end() - 1);
    for (int h = 0; h < ng; ++h)
      for (int q = gwoff[h]; q < gwoff[h + 1]; ++q) wlist[fill[gw[q]]++] = h;
  }
  // second group of a group's tasks (the same for all of them), -1 if none
  std::vector<int> g2(ng, -1);
  for (const auto& t : tasks)
    if (t.group2 >= 0) g2[t.group] = t.group2;
  // bottom levels: every group waits only on groups created before it, and
  // second groups are created after the groups that feed them
  std::vector<int64_t> bl(ng, 0);
  for (int g = ng - 1; g >= 0; --g) {
    int64_t c = 0, below = g2[g] >= 0 ? bl[g2[g]] : 0;
    for (int q = tptr[g]; q < tptr[g + 1]; ++q) c = std::max(c, cost[tlist[q]]);
    for (int q = wptr[g]; q < wptr[g + 1]; ++q) below = std::max(below, bl[wlist[q]]);
    bl[g] = c + below;
  }
  // waits outstanding per group; a group without tasks of its own (the
  // second group) completes when its count of signalling tasks is reached
  std::vector<int> pending(ng);
  for (int h = 0; h < ng; ++h) pending[h] = gwoff[h + 1] - gwoff[h];
  std::vector<unsigned> done(ng, 0u);
  using Ready = std::pair<int64_t, int>;  // (-bottom level, group), smallest first
  std::priority_queue<Ready, std::vector<Ready>, std::greater<Ready>> ready;
  std::vector<int> cursor(tptr.begin(), tptr.end() - 1);
  for (int g = 0; g < ng; ++g)
    if (pending[g] == 0 && tptr[g + 1] > tptr[g]) ready.push({-bl[g], g});
  using Run = std::pair<int64_t, int>;  // (finish time, task)
  std::priority_queue<Run, std::vector<Run>, std::greater<Run>> running;
  std::vector<int> order;
  order.reserve(nt);
  int64_t now = 0;
  int free_w = std::max(1, workers);
  auto finish = [&](int g) {
    if (++done[g] != need[g]) return;
    for (int q = wptr[g]; q < wptr[g + 1]; ++q) {
      const int h = wlist[q];
      if (--pending[h] == 0 && tptr[h + 1] > tptr[h]) ready.push({-bl[h], h});
    }
  };
  while ((int)order.size() < nt) {
    while (free_w > 0 && !ready.empty()) {
      const int g = ready.top().second;
      const int i = tlist[cursor[g]++];
      if (cursor[g] == tptr[g + 1]) ready.pop();
      order.push_back(i);
      running.push({now + cost[i], i});
      --free_w;
    }
    if (running.empty()) throw std::logic_error("dataflow apply: task graph has a cycle");
    const Run r = running.top();
    running.pop();
    now = r.first;
    ++free_w;
    finish(tasks[r.second].group);
    if (tasks[r.second].group2 >= 0) finish(tasks[r.second].group2);
  }
  while (!running.empty()) {
    now = std::max(now, running.top().first);
    running.pop();
  }
  std::vector<dev::DfTask> out(nt);
  for (int q = 0; q < nt; ++q) out[q] = tasks[order[q]];
  tasks.swap(out);
  return now / 1000.0;
}

}  // namespace

void build_dataflow_tasks(const Plan& P, const SolveLayout& L, DataflowTasks& out,
                          const Subtrees* subtrees, const TicketModel& tm) {
  const int ns = (int)P.nodes.size();
  const int VCH = kSolveChunk, SCH = kScatterChunk;
  std::vector<dev::GvTask> gv;
  std::vector<dev::DfTask> tasks;
  std::vector<unsigned> need;
  std::vector<int> gwoff{0}, gw;
  std::vector<int> lastA(ns, -1), lastB(ns, -1), lastC(ns, -1), asmg(ns, -1);
  TaskEmitter em;
  em.row_split = tm.row_cols > 0 ? tm.row_cols : kRowSplitCols;
  // fused reductions: every consumer task re-sums the partial slots, cheap
  // for one right-hand side; for sixteen the extra loads cost more than the
  // saved reduce tasks (measured)
  em.fuse_tv = tm.nrhs <= 1;
  if (tm.col_rows > 0) em.CHUNK = std::min(tm.col_rows, kColRowsMax);
  // finer splits of huge products: measured faster with sixteen right-hand
  // sides (matrix-core bodies), not with one
  em.workers = tm.nrhs > 1 || tm.wide_elems > 0 ? tm.workers : 0;
  if (tm.wide_elems > 0) em.wide_elems = tm.wide_elems;
  auto new_group = [&](const std::vector<int>& waits) {
    need.push_back(0u);
    gw.insert(gw.end(), waits.begin(), waits.end());
    gwoff.push_back((int)gw.size());
    return (int)need.size() - 1;
  };
  int cur_node = -1, cur_stage = 0;
  auto add = [&](int type, int a, int b, int g) {
    dev::DfTask t{};
    t.type = type;
    t.a = a;
    t.b = b;
    t.group = g;
    t.group2 = -1;
    t.tag = ((cur_node + 1) << 8) | std::min(cur_stage, 255);
    tasks.push_back(t);
    need[g] += 1u;
  };
  // one stage: its tasks wait on `after`; returns the last group (reductions
  // after the main tasks), or `after` when the stage is empty
  auto stage = [&](const Stage& st, int after) {
    std::vector<dev::GvTask> main, red;
    for (const Mv& m : st) em.emit(m, main, red);
    int last = after;
    if (!main.empty()) {
      const int g = new_group(after >= 0 ? std::vector<int>{after} : std::vector<int>{});
      for (const dev::GvTask& t : main) {
        gv.push_back(t);
        add(2, (int)gv.size() - 1, 0, g);
      }
      last = g;
    }
    if (!red.empty()) {
      const int g = new_group({last});
      for (const dev::GvTask& t : red) {
        gv.push_back(t);
        add(2, (int)gv.size() - 1, 0, g);
      }
      last = g;
    }
    return last;
  };
  // assembly / gather tasks of a supernode: groups of chunk_group chunks
  const int CG = std::max(1, tm.chunk_group);
  auto chunks = [&](int j) {
    const int nch = (P.nodes[j].ncols + P.nodes[j].R() + VCH - 1) / VCH;
    return (nch + CG - 1) / CG;
  };
  std::vector<NodeStages> st(ns);
  for (int j = 0; j < ns; ++j) st[j] = node_stages(P, L, j);
  Subtrees none;
  none.of.assign(ns, -1);
  none.ptr.assign(1, 0);
  none.rptr.assign(1, 0);
  const Subtrees& SB = subtrees ? *subtrees : none;
  const int nsub = (int)SB.root.size();
  std::vector<std::vector<int>> lnodes(P.levels.size());
  for (size_t lv = 0; lv < P.levels.size(); ++lv)
    for (int j : P.levels[lv])
      if (SB.of[j] < 0) lnodes[lv].push_back(j);
  for (int sb = 0; sb < nsub; ++sb) {
    cur_node = SB.root[sb];
    cur_stage = 3;
    const int g = new_group({});
    add(4, sb, 0, g);
    for (int q = SB.rptr[sb]; q < SB.rptr[sb + 1]; ++q) lastA[SB.roots[q]] = lastB[SB.roots[q]] = g;
  }
  for (size_t lv = 0; lv < P.levels.size(); ++lv) {
    const std::vector<int>& nodes = lnodes[lv];
    for (int j : nodes) {
      std::vector<int> waits;
      for (int c : P.nodes[j].children) waits.push_back(lastA[c]);
      cur_node = j;
      cur_stage = 0;
      const int g = new_group(waits);
      for (int c = 0; c < chunks(j); ++c) add(0, j, c, g);
      lastA[j] = asmg[j] = g;
    }
    size_t nA = 0, nB = 0;
    for (int j : nodes) {
      nA = std::max(nA, st[j].A.size());
      nB = std::max(nB, st[j].B.size());
    }
    for (size_t k = 0; k < nA; ++k)
      for (int j : nodes)
        if (k < st[j].A.size()) {
          cur_node = j;
          cur_stage = 10 + (int)k;
          lastA[j] = stage(st[j].A[k], lastA[j]);
        }
    for (int j : nodes) lastB[j] = asmg[j];
    for (size_t k = 0; k < nB; ++k)
      for (int j : nodes)
        if (k < st[j].B.size()) {
          cur_node = j;
          cur_stage = 20 + (int)k;
          lastB[j] = stage(st[j].B[k], lastB[j]);
        }
  }
  for (int lv = (int)P.levels.size() - 1; lv >= 0; --lv) {
    const std::vector<int>& nodes = lnodes[lv];
    for (int j : nodes) {
      const int par = P.nodes[j].parent;
      std::vector<int> waits = {lastB[j]};
      if (par >= 0) waits.push_back(lastC[par]);
      cur_node = j;
      cur_stage = 1;
      const int g = new_group(waits);
      for (int c = 0; c < chunks(j); ++c) add(1, j, c, g);
      lastC[j] = g;
    }
    size_t nC = 0;
    for (int j : nodes) nC = std::max(nC, st[j].C.size());
    for (size_t k = 0; k < nC; ++k)
      for (int j : nodes)
        if (k < st[j].C.size()) {
          cur_node = j;
          cur_stage = 30 + (int)k;
          lastC[j] = stage(st[j].C[k], lastC[j]);
        }
  }
  for (int sb = 0; sb < nsub; ++sb) {
    const int r = SB.root[sb];
    std::vector<int> waits;
    for (int q = SB.rptr[sb]; q < SB.rptr[sb + 1]; ++q) {
      const int rq = SB.roots[q], par = P.nodes[rq].parent;
      waits.push_back(lastB[rq]);
      if (par >= 0) waits.push_back(lastC[par]);
    }
    cur_node = r;
    cur_stage = 4;
    const int g = new_group(waits);
    add(5, sb, 0, g);
    for (int q = SB.ptr[sb]; q < SB.ptr[sb + 1]; ++q) lastC[SB.nodes[q]] = g;
  }
  // scatter after every node's last backward group
  const int ready = new_group({});
  std::vector<char> final_group(need.size(), 0);
  for (int j = 0; j < ns; ++j) final_group[lastC[j]] = 1;
  for (auto& t : tasks)
    if (final_group[t.group]) {
      t.group2 = ready;
      need[ready] += 1u;
    }
  const int gsc = new_group({ready});
  cur_node = -1;
  cur_stage = 2;
  for (int c = 0; c * SCH < P.n; ++c) add(3, c, 0, gsc);
  out.est_span_us = 0;
  if (tm.workers > 0)
    out.est_span_us =
        order_by_list_schedule(tasks, need, gwoff, gw, CostModel{P, L, SB, gv, tm}, tm.workers);
  for (auto& t : tasks) {
    t.wbeg = gwoff[t.group];
    t.wend = gwoff[t.group + 1];
    t.wg0 = t.wend > t.wbeg ? gw[t.wbeg] : -1;
    t.wn0 = t.wg0 >= 0 ? need[t.wg0] : 0u;
    t.need1 = need[t.group];
    t.need2 = t.group2 >= 0 ? need[t.group2] : 0u;
  }
  {
    // per task for timing traces: the cost model's estimate and, for
    // subtree tasks, the model's inputs (supernodes, front entries, kB)
    const CostModel cm{P, L, SB, gv, tm};
    out.trace_feat.assign(4 * tasks.size(), 0.f);
    for (size_t i = 0; i < tasks.size(); ++i) {
      const dev::DfTask& t = tasks[i];
      float* f = out.trace_feat.data() + 4 * i;
      f[0] = (float)cm.us(t);
      if (t.type == 4 || t.type == 5) {
        for (int q = SB.ptr[t.a]; q < SB.ptr[t.a + 1]; ++q) {
          const Supernode& s = P.nodes[SB.nodes[q]];
          f[1] += 1.f;
          f[2] += (float)(s.ncols + s.R());
          f[3] += (float)((double)s.ncols * (s.ncols + s.R()) * tm.elem_bytes / 1000.0);
        }
      }
    }
  }
  {
    // matrix-vector descriptors parallel to the task list
    dev::GvTask none{};
    none.xpart = -1;
    std::vector<dev::GvTask> gvt(tasks.size(), none);
    for (size_t i = 0; i < tasks.size(); ++i)
      if (tasks[i].type == 2) {
        gvt[i] = gv[tasks[i].a];
        tasks[i].a = (int)i;
      }
    gv.swap(gvt);
  }
  // every group waited on must complete before the waiting task's ticket
  std::vector<int> last_task(need.size(), -1);
  for (size_t i = 0; i < tasks.size(); ++i) {
    last_task[tasks[i].group] = (int)i;
    if (tasks[i].group2 >= 0) last_task[tasks[i].group2] = (int)i;
  }
  for (size_t i = 0; i < tasks.size(); ++i)
    for (int q = gwoff[tasks[i].group]; q < gwoff[tasks[i].group + 1]; ++q) {
      const int g = gw[q];
      if (g < 0 || g >= (int)need.size() || last_task[g] >= (int)i || need[g] == 0u)
        throw std::logic_error("dataflow apply: task waits on a later or empty group");
    }
  {
    // ready-queue form: tasks per group and successor groups, both in list
    // order
    const int ng = (int)need.size();
    out.q_gptr.assign(ng + 1, 0);
    for (const auto& t : tasks) ++out.q_gptr[t.group + 1];
    for (int g = 0; g < ng; ++g) out.q_gptr[g + 1] += out.q_gptr[g];
    out.q_gtasks.assign(tasks.size(), 0);
    std::vector<int> fill(out.q_gptr.begin(), out.q_gptr.end() - 1);
    for (size_t i = 0; i < tasks.size(); ++i) out.q_gtasks[fill[tasks[i].group]++] = (int)i;
    out.q_sptr.assign(ng + 1, 0);
    out.q_pend0.assign(ng, 0);
    for (int h = 0; h < ng; ++h) {
      // a small group's tasks each count down h once, a larger group once
      for (int q = gwoff[h]; q < gwoff[h + 1]; ++q)
        out.q_pend0[h] += need[gw[q]] <= dev::kQDirectMax ? (int)need[gw[q]] : 1;
      for (int q = gwoff[h]; q < gwoff[h + 1]; ++q) ++out.q_sptr[gw[q] + 1];
    }
    for (int g = 0; g < ng; ++g) out.q_sptr[g + 1] += out.q_sptr[g];
    out.q_succ.assign(gw.size(), 0);
    std::vector<int> sf(out.q_sptr.begin(), out.q_sptr.end() - 1);
    for (int h = 0; h < ng; ++h)
      for (int q = gwoff[h]; q < gwoff[h + 1]; ++q) out.q_succ[sf[gw[q]]++] = h;
    // bottom level of every group (estimated time from its start to the end
    // of the apply; groups only wait on groups created before them): the
    // successors of a group and the initial tasks go most critical first (a
    // workgroup runs the first ready successor's first task itself)
    std::vector<double> bl(ng, 0.0), gcost(ng, 0.0);
    for (size_t i = 0; i < tasks.size(); ++i)
      gcost[tasks[i].group] = std::max(gcost[tasks[i].group], (double)out.trace_feat[4 * i]);
    std::vector<int> g2(ng, -1);
    for (const auto& t : tasks)
      if (t.group2 >= 0) g2[t.group] = t.group2;
    for (int g = ng - 1; g >= 0; --g) {
      double below = g2[g] >= 0 ? bl[g2[g]] : 0.0;
      for (int q = out.q_sptr[g]; q < out.q_sptr[g + 1]; ++q)
        below = std::max(below, bl[out.q_succ[q]]);
      bl[g] = gcost[g] + below;
    }
    for (int g = 0; g < ng; ++g)
      std::stable_sort(out.q_succ.begin() + out.q_sptr[g], out.q_succ.begin() + out.q_sptr[g + 1],
                       [&](int x, int y) { return bl[x] > bl[y]; });
    for (auto& t : tasks) {
      t.succ1[0] = out.q_sptr[t.group];
      t.succ1[1] = out.q_sptr[t.group + 1];
      t.succ2[0] = t.group2 >= 0 ? out.q_sptr[t.group2] : 0;
      t.succ2[1] = t.group2 >= 0 ? out.q_sptr[t.group2 + 1] : 0;
      t.sfirst[0] = t.succ1[1] > t.succ1[0] ? out.q_succ[t.succ1[0]] : -1;
      t.sfirst[1] = t.succ2[1] > t.succ2[0] ? out.q_succ[t.succ2[0]] : -1;
    }
    out.q_ginfo.assign(4 * (size_t)ng, 0);
    for (int g = 0; g < ng; ++g) {
      const int c0 = out.q_gptr[g], cn = out.q_gptr[g + 1] - c0;
      out.q_ginfo[4 * g] = c0;
      out.q_ginfo[4 * g + 1] = cn;
      out.q_ginfo[4 * g + 2] = cn > 0 ? out.q_gtasks[c0] : -1;
    }
    out.q_init.clear();
    for (size_t i = 0; i < tasks.size(); ++i)
      if (out.q_pend0[tasks[i].group] == 0) out.q_init.push_back((int)i);
    std::stable_sort(out.q_init.begin(), out.q_init.end(), [&](int x, int y) {
      return bl[tasks[x].group] > bl[tasks[y].group];
    });
    if (out.q_init.empty() && !tasks.empty())
      throw std::logic_error("dataflow apply: no task is ready at the start");
  }
  out.tasks.swap(tasks);
  out.gv.swap(gv);
  out.need.swap(need);
  out.gwoff.swap(gwoff);
  out.gw.swap(gw);
  out.slots = em.slot;
}

}  // namespace rsc
