// Host construction of the apply's task lists (declarations and the
// meaning of each task in solve_tasks.h).
#include "solve_tasks.h"

#include <algorithm>
#include <climits>
#include <stdexcept>

namespace rsc {

void build_solve_layout(const Plan& P, SolveLayout& L) {
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
    d.pad = 0;
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
  L.fv_stride = std::max<int64_t>(fvo, 1);
  L.tv_stride = std::max<int64_t>(tvo, 1);
}

// Level-scheduled form: per level and direction a few phases, each one
// launch of independent tasks (a phase's split column tasks are followed by a
// launch of their reductions).
void build_level_tasks(const Plan& P, const SolveLayout& L, LevelTasks& out) {
  const std::vector<int64_t>& tv_node = L.tv_node;
  const std::vector<std::vector<int64_t>>& tv_h = L.tv_h;
  auto fvoff = [&](int j) { return L.fv_off[j]; };
  std::vector<dev::GvTask> gv;
  std::vector<std::vector<SolvePhase>>& fwd_ = out.fwd;
  std::vector<std::vector<SolvePhase>>& bwd_ = out.bwd;
  fwd_.assign(P.levels.size(), {});
  bwd_.assign(P.levels.size(), {});
  const int W = 0, FV = 1, TV = 2;
  // COL tasks: 32 columns; more than CHUNK rows are split into row chunks
  // writing partial sums, reduced by a second launch (`red`).
  const int CHUNK = 512;
  std::vector<dev::GvTask> red;
  int64_t slot = 0;
  auto col = [&](int64_t mat, int ld, int rows, int cols, int op, int xs, int64_t xo,
                 int ys, int64_t yo, bool upper) {
    for (int j0 = 0; j0 < cols; j0 += 32) {
      const int rr = upper ? std::min(rows, j0 + 32) : rows;
      if (rr <= 0) continue;
      dev::GvTask t{};
      t.ld = ld;
      t.cols = std::min(32, cols - j0);
      t.kind = 1;
      t.op = op;
      t.xsel = xs;
      t.ysel = ys;
      t.yoff = yo + j0;
      if (rr <= CHUNK) {
        t.mat = mat + (int64_t)j0 * ld;
        t.rows = rr;
        t.xoff = xo;
        t.part = -1;
        gv.push_back(t);
        continue;
      }
      const int nsub = (rr + CHUNK - 1) / CHUNK;
      for (int s2 = 0; s2 < nsub; ++s2) {
        dev::GvTask u = t;
        u.mat = mat + (int64_t)j0 * ld + (int64_t)s2 * CHUNK;
        u.rows = std::min(CHUNK, rr - s2 * CHUNK);
        u.xoff = xo + (int64_t)s2 * CHUNK;
        u.part = (int)(slot + s2);
        gv.push_back(u);
      }
      dev::GvTask rd{};
      rd.kind = 2;
      rd.part = (int)slot;
      rd.rows = nsub;
      rd.cols = t.cols;
      rd.op = op;
      rd.ysel = ys;
      rd.yoff = yo + j0;
      red.push_back(rd);
      slot += nsub;
    }
  };
  // ROW tasks: 64 rows, all columns
  auto row = [&](int64_t mat, int ld, int rows, int cols, int op, int xs, int64_t xo,
                 int ys, int64_t yo, bool upper) {
    for (int i0 = 0; i0 < rows; i0 += 64) {
      dev::GvTask t{};
      t.mat = mat + i0 + (upper ? (int64_t)i0 * ld : 0);
      t.ld = ld;
      t.rows = std::min(64, rows - i0);
      t.cols = upper ? cols - i0 : cols;
      t.kind = 0;
      t.op = op;
      t.xsel = xs;
      t.ysel = ys;
      t.part = -1;
      t.xoff = xo + (upper ? i0 : 0);
      t.yoff = yo + i0;
      gv.push_back(t);
    }
  };
  // a phase = small tasks (kind 2 launch) + large tasks (kind 3 launch),
  // then the partial-sum reductions of its split column tasks
  auto is_big = [](const dev::GvTask& t) {
    return t.kind == 0 ? t.cols > 64 : t.rows > 128;
  };
  auto phase = [&](std::vector<SolvePhase>& out, size_t start) {
    if (gv.size() > start) {
      std::stable_partition(gv.begin() + start, gv.end(),
                            [&](const dev::GvTask& t) { return !is_big(t); });
      size_t mid = start;
      while (mid < gv.size() && !is_big(gv[mid])) ++mid;
      if (mid > start) out.push_back({2, (int)start, (int)(mid - start)});
      if (gv.size() > mid) out.push_back({3, (int)mid, (int)(gv.size() - mid)});
    }
    if (!red.empty()) {
      const size_t r0 = gv.size();
      gv.insert(gv.end(), red.begin(), red.end());
      red.clear();
      out.push_back({2, (int)r0, (int)(gv.size() - r0)});
    }
  };
  for (size_t lv = 0; lv < P.levels.size(); ++lv) {
    const std::vector<int>& nodes = P.levels[lv];
    int maxh = 0;
    for (int j : nodes)
      for (auto& h : P.nodes[j].hnodes) maxh = std::max(maxh, h.height);
    // ---- forward
    std::vector<SolvePhase>& F = fwd_[lv];
    F.push_back({0, 0, 0});
    size_t st = gv.size();
    for (int j : nodes) {
      const Supernode& s = P.nodes[j];
      const int64_t fo = 0;
      (void)fo;
      if (!s.compress) {
        col(s.off_panel, s.ncols + s.R(), s.ncols, s.ncols, 0, FV, fvoff(j), W, s.first, true);
      } else {
        for (auto& h : s.hnodes)
          if (h.leaf())
            col(h.off_leaf, h.r1 - h.r0, h.r1 - h.r0, h.r1 - h.r0, 0, FV, fvoff(j) + h.r0, W,
                s.first + h.r0, true);
      }
    }
    phase(F, st);
    for (int hh = 1; hh <= maxh; ++hh) {
      st = gv.size();
      for (int j : nodes) {
        const Supernode& s = P.nodes[j];
        for (size_t q = 0; q < s.hnodes.size(); ++q) {
          const HNode& h = s.hnodes[q];
          if (h.leaf() || h.height != hh) continue;
          const HNode& A = s.hnodes[h.left];
          col(h.off_U, A.r1 - A.r0, A.r1 - A.r0, h.rank, 0, W, s.first + A.r0, TV, tv_h[j][q],
              false);
        }
      }
      phase(F, st);
      st = gv.size();
      for (int j : nodes) {
        const Supernode& s = P.nodes[j];
        for (size_t q = 0; q < s.hnodes.size(); ++q) {
          const HNode& h = s.hnodes[q];
          if (h.leaf() || h.height != hh) continue;
          const HNode& B = s.hnodes[h.right];
          row(h.off_W, B.r1 - B.r0, B.r1 - B.r0, h.rank, 1, TV, tv_h[j][q], W, s.first + B.r0,
              false);
        }
      }
      phase(F, st);
    }
    st = gv.size();
    for (int j : nodes) {
      const Supernode& s = P.nodes[j];
      if (s.compress && s.R() > 0)
        col(s.off_U, s.ncols, s.ncols, s.rank, 0, W, s.first, TV, tv_node[j], false);
    }
    phase(F, st);
    st = gv.size();
    for (int j : nodes) {
      const Supernode& s = P.nodes[j];
      if (s.R() == 0) continue;
      if (!s.compress)
        row(s.off_panel + s.ncols, s.ncols + s.R(), s.R(), s.ncols, 1, W, s.first, FV,
            fvoff(j) + s.ncols, false);
      else
        row(s.off_V, s.R(), s.R(), s.rank, 1, TV, tv_node[j], FV, fvoff(j) + s.ncols, false);
    }
    phase(F, st);
    // ---- backward
    std::vector<SolvePhase>& Bk = bwd_[lv];
    Bk.push_back({1, 0, 0});
    st = gv.size();
    for (int j : nodes) {
      const Supernode& s = P.nodes[j];
      if (s.R() == 0) continue;
      if (!s.compress)
        col(s.off_panel + s.ncols, s.ncols + s.R(), s.R(), s.ncols, 1, FV, fvoff(j) + s.ncols,
            FV, fvoff(j), false);
      else
        col(s.off_V, s.R(), s.R(), s.rank, 0, FV, fvoff(j) + s.ncols, TV, tv_node[j], false);
    }
    phase(Bk, st);
    st = gv.size();
    for (int j : nodes) {
      const Supernode& s = P.nodes[j];
      if (s.compress && s.R() > 0)
        row(s.off_U, s.ncols, s.ncols, s.rank, 1, TV, tv_node[j], FV, fvoff(j), false);
    }
    phase(Bk, st);
    st = gv.size();
    for (int j : nodes) {
      const Supernode& s = P.nodes[j];
      if (!s.compress) {
        row(s.off_panel, s.ncols + s.R(), s.ncols, s.ncols, 0, FV, fvoff(j), W, s.first, true);
      } else {
        for (auto& h : s.hnodes)
          if (h.leaf())
            row(h.off_leaf, h.r1 - h.r0, h.r1 - h.r0, h.r1 - h.r0, 0, FV, fvoff(j) + h.r0, W,
                s.first + h.r0, true);
      }
    }
    phase(Bk, st);
    for (int hh = 1; hh <= maxh; ++hh) {
      st = gv.size();
      for (int j : nodes) {
        const Supernode& s = P.nodes[j];
        for (size_t q = 0; q < s.hnodes.size(); ++q) {
          const HNode& h = s.hnodes[q];
          if (h.leaf() || h.height != hh) continue;
          const HNode& B = s.hnodes[h.right];
          col(h.off_V, B.r1 - B.r0, B.r1 - B.r0, h.rank, 0, W, s.first + B.r0, TV, tv_h[j][q],
              false);
        }
      }
      phase(Bk, st);
      st = gv.size();
      for (int j : nodes) {
        const Supernode& s = P.nodes[j];
        for (size_t q = 0; q < s.hnodes.size(); ++q) {
          const HNode& h = s.hnodes[q];
          if (h.leaf() || h.height != hh) continue;
          const HNode& A = s.hnodes[h.left];
          row(h.off_X, A.r1 - A.r0, A.r1 - A.r0, h.rank, 1, TV, tv_h[j][q], W, s.first + A.r0,
              false);
        }
      }
      phase(Bk, st);
    }
  }
  out.gv.swap(gv);
  out.slots = slot;
}

// Dataflow form of the same solves: per supernode a chain of stages
// (forward: assembly, diagonal leaves, HODLR heights, off-diagonal; backward:
// gather, off-diagonal, diagonal leaves, HODLR heights), each stage one
// group of tasks waiting on the previous stage of the same node.  A
// node's forward assembly waits on its children's last forward stage; its
// backward gather waits on its parent's last backward stage (on its own
// forward for a root).  The scatter waits on every node's last backward
// stage.  Tasks are emitted level by level, so the list is in topological
// order.
void build_dataflow_tasks(const Plan& P, const SolveLayout& L, DataflowTasks& out) {
  const std::vector<int64_t>& tv_node = L.tv_node;
  const std::vector<std::vector<int64_t>>& tv_h = L.tv_h;
  auto fvoff = [&](int j) { return L.fv_off[j]; };
  const int ns = (int)P.nodes.size();
  const int W = 0, FV = 1, TV = 2;
  const int CHUNK = 512, VCH = kSolveChunk, SCH = kScatterChunk;
  std::vector<dev::GvTask> gv, cur, red;
  std::vector<dev::DfTask> tasks;
  std::vector<unsigned> need;
  std::vector<int> gwoff{0}, gw;
  std::vector<int> last_f(ns, -1), last_b(ns, -1);
  int64_t slot = 0;
  auto new_group = [&](const std::vector<int>& waits) {
    need.push_back(0u);
    gw.insert(gw.end(), waits.begin(), waits.end());
    gwoff.push_back((int)gw.size());
    return (int)need.size() - 1;
  };
  auto add = [&](int type, int a, int b, int g) {
    dev::DfTask t{};
    t.type = type;
    t.a = a;
    t.b = b;
    t.group = g;
    t.group2 = -1;
    tasks.push_back(t);
    need[g] += 1u;
  };
  auto col = [&](int64_t mat, int ld, int rows, int cols, int op, int xs, int64_t xo, int ys,
                 int64_t yo, bool upper) {
    for (int j0 = 0; j0 < cols; j0 += 32) {
      const int rr = upper ? std::min(rows, j0 + 32) : rows;
      if (rr <= 0) continue;
      dev::GvTask t{};
      t.ld = ld;
      t.cols = std::min(32, cols - j0);
      t.kind = 1;
      t.op = op;
      t.xsel = xs;
      t.ysel = ys;
      t.yoff = yo + j0;
      if (rr <= CHUNK) {
        t.mat = mat + (int64_t)j0 * ld;
        t.rows = rr;
        t.xoff = xo;
        t.part = -1;
        cur.push_back(t);
        continue;
      }
      const int nsub = (rr + CHUNK - 1) / CHUNK;
      for (int s2 = 0; s2 < nsub; ++s2) {
        dev::GvTask u = t;
        u.mat = mat + (int64_t)j0 * ld + (int64_t)s2 * CHUNK;
        u.rows = std::min(CHUNK, rr - s2 * CHUNK);
        u.xoff = xo + (int64_t)s2 * CHUNK;
        u.part = (int)(slot + s2);
        cur.push_back(u);
      }
      dev::GvTask rd{};
      rd.kind = 2;
      rd.part = (int)slot;
      rd.rows = nsub;
      rd.cols = t.cols;
      rd.op = op;
      rd.ysel = ys;
      rd.yoff = yo + j0;
      red.push_back(rd);
      slot += nsub;
    }
  };
  auto row = [&](int64_t mat, int ld, int rows, int cols, int op, int xs, int64_t xo, int ys,
                 int64_t yo, bool upper) {
    for (int i0 = 0; i0 < rows; i0 += 64) {
      dev::GvTask t{};
      t.mat = mat + i0 + (upper ? (int64_t)i0 * ld : 0);
      t.ld = ld;
      t.rows = std::min(64, rows - i0);
      t.cols = upper ? cols - i0 : cols;
      t.kind = 0;
      t.op = op;
      t.xsel = xs;
      t.ysel = ys;
      t.part = -1;
      t.xoff = xo + (upper ? i0 : 0);
      t.yoff = yo + i0;
      cur.push_back(t);
    }
  };
  // close a matrix-vector stage of node j: its tasks, then its reductions
  auto stage = [&](int j, bool fwd) {
    int& last = fwd ? last_f[j] : last_b[j];
    if (!cur.empty()) {
      const int g = new_group({last});
      for (const dev::GvTask& t : cur) {
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
    cur.clear();
    red.clear();
  };
  auto chunks = [&](int j) {
    return (P.nodes[j].ncols + P.nodes[j].R() + VCH - 1) / VCH;
  };
  // ---- forward, bottom-up
  for (size_t lv = 0; lv < P.levels.size(); ++lv) {
    const std::vector<int>& nodes = P.levels[lv];
    int maxh = 0;
    for (int j : nodes)
      for (auto& h : P.nodes[j].hnodes) maxh = std::max(maxh, h.height);
    for (int j : nodes) {
      std::vector<int> waits;
      for (int c : P.nodes[j].children) waits.push_back(last_f[c]);
      const int g = new_group(waits);
      for (int c = 0; c < chunks(j); ++c) add(0, j, c, g);
      last_f[j] = g;
    }
    for (int j : nodes) {
      const Supernode& s = P.nodes[j];
      if (!s.compress) {
        col(s.off_panel, s.ncols + s.R(), s.ncols, s.ncols, 0, FV, fvoff(j), W, s.first, true);
      } else {
        for (auto& h : s.hnodes)
          if (h.leaf())
            col(h.off_leaf, h.r1 - h.r0, h.r1 - h.r0, h.r1 - h.r0, 0, FV, fvoff(j) + h.r0, W,
                s.first + h.r0, true);
      }
      stage(j, true);
    }
    for (int hh = 1; hh <= maxh; ++hh) {
      for (int j : nodes) {
        const Supernode& s = P.nodes[j];
        for (size_t q = 0; q < s.hnodes.size(); ++q) {
          const HNode& h = s.hnodes[q];
          if (h.leaf() || h.height != hh) continue;
          const HNode& A = s.hnodes[h.left];
          col(h.off_U, A.r1 - A.r0, A.r1 - A.r0, h.rank, 0, W, s.first + A.r0, TV, tv_h[j][q],
              false);
        }
        stage(j, true);
      }
      for (int j : nodes) {
        const Supernode& s = P.nodes[j];
        for (size_t q = 0; q < s.hnodes.size(); ++q) {
          const HNode& h = s.hnodes[q];
          if (h.leaf() || h.height != hh) continue;
          const HNode& B = s.hnodes[h.right];
          row(h.off_W, B.r1 - B.r0, B.r1 - B.r0, h.rank, 1, TV, tv_h[j][q], W, s.first + B.r0,
              false);
        }
        stage(j, true);
      }
    }
    for (int j : nodes) {
      const Supernode& s = P.nodes[j];
      if (s.compress && s.R() > 0)
        col(s.off_U, s.ncols, s.ncols, s.rank, 0, W, s.first, TV, tv_node[j], false);
      stage(j, true);
    }
    for (int j : nodes) {
      const Supernode& s = P.nodes[j];
      if (s.R() > 0) {
        if (!s.compress)
          row(s.off_panel + s.ncols, s.ncols + s.R(), s.R(), s.ncols, 1, W, s.first, FV,
              fvoff(j) + s.ncols, false);
        else
          row(s.off_V, s.R(), s.R(), s.rank, 1, TV, tv_node[j], FV, fvoff(j) + s.ncols, false);
      }
      stage(j, true);
    }
  }
  // ---- backward, top-down
  for (int lv = (int)P.levels.size() - 1; lv >= 0; --lv) {
    const std::vector<int>& nodes = P.levels[lv];
    int maxh = 0;
    for (int j : nodes)
      for (auto& h : P.nodes[j].hnodes) maxh = std::max(maxh, h.height);
    for (int j : nodes) {
      const int par = P.nodes[j].parent;
      const int g = new_group({par >= 0 ? last_b[par] : last_f[j]});
      for (int c = 0; c < chunks(j); ++c) add(1, j, c, g);
      last_b[j] = g;
    }
    for (int j : nodes) {
      const Supernode& s = P.nodes[j];
      if (s.R() > 0) {
        if (!s.compress)
          col(s.off_panel + s.ncols, s.ncols + s.R(), s.R(), s.ncols, 1, FV,
              fvoff(j) + s.ncols, FV, fvoff(j), false);
        else
          col(s.off_V, s.R(), s.R(), s.rank, 0, FV, fvoff(j) + s.ncols, TV, tv_node[j], false);
      }
      stage(j, false);
    }
    for (int j : nodes) {
      const Supernode& s = P.nodes[j];
      if (s.compress && s.R() > 0)
        row(s.off_U, s.ncols, s.ncols, s.rank, 1, TV, tv_node[j], FV, fvoff(j), false);
      stage(j, false);
    }
    for (int j : nodes) {
      const Supernode& s = P.nodes[j];
      if (!s.compress) {
        row(s.off_panel, s.ncols + s.R(), s.ncols, s.ncols, 0, FV, fvoff(j), W, s.first, true);
      } else {
        for (auto& h : s.hnodes)
          if (h.leaf())
            row(h.off_leaf, h.r1 - h.r0, h.r1 - h.r0, h.r1 - h.r0, 0, FV, fvoff(j) + h.r0, W,
                s.first + h.r0, true);
      }
      stage(j, false);
    }
    for (int hh = 1; hh <= maxh; ++hh) {
      for (int j : nodes) {
        const Supernode& s = P.nodes[j];
        for (size_t q = 0; q < s.hnodes.size(); ++q) {
          const HNode& h = s.hnodes[q];
          if (h.leaf() || h.height != hh) continue;
          const HNode& B = s.hnodes[h.right];
          col(h.off_V, B.r1 - B.r0, B.r1 - B.r0, h.rank, 0, W, s.first + B.r0, TV, tv_h[j][q],
              false);
        }
        stage(j, false);
      }
      for (int j : nodes) {
        const Supernode& s = P.nodes[j];
        for (size_t q = 0; q < s.hnodes.size(); ++q) {
          const HNode& h = s.hnodes[q];
          if (h.leaf() || h.height != hh) continue;
          const HNode& A = s.hnodes[h.left];
          row(h.off_X, A.r1 - A.r0, A.r1 - A.r0, h.rank, 1, TV, tv_h[j][q], W, s.first + A.r0,
              false);
        }
        stage(j, false);
      }
    }
  }
  // ---- scatter after every node's last backward stage
  const int ready = new_group({});
  std::vector<char> final_group(need.size(), 0);
  for (int j = 0; j < ns; ++j) final_group[last_b[j]] = 1;
  for (auto& t : tasks)
    if (final_group[t.group]) {
      t.group2 = ready;
      need[ready] += 1u;
    }
  const int gsc = new_group({ready});
  for (int c = 0; c * SCH < P.n; ++c) add(3, c, 0, gsc);
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
  out.tasks.swap(tasks);
  out.gv.swap(gv);
  out.need.swap(need);
  out.gwoff.swap(gwoff);
  out.gw.swap(gw);
  out.slots = slot;
}

}  // namespace rsc
