// Host ordering: nested dissection, elimination tree, column counts,
// relaxed supernodes and separator reordering.
//
// The supernode construction and the separator reordering restate the
// reference's NestedDissection::BuildRelaxedSupernodes
// (src/rschol/ordering/NestedDissection.cpp:1266-1667) and
// NestedDissection::ReorderNestedSeparators (:728-910) with SeparatorTree
// (src/rschol/datastructure/SeparatorTree.cpp:189-589).  Nested dissection
// itself is CHOLMOD/METIS in the reference (CHOLMOD_Environment.cpp:326-432);
// here it is this file's own bisection producing the same (perm, cParent,
// cMember) triple.
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <thread>

#include "rsc_host.h"

namespace rsc {

std::string& last_error() {
  static thread_local std::string msg;
  return msg;
}

// --------------------------------------------------------------------------
// Nested dissection
// --------------------------------------------------------------------------
namespace {

// Per-vertex scratch shared by the dissection of disjoint vertex sets (one
// thread per set near the top of the tree).  A set is marked with a stamp
// value no other set uses, so a neighbour in another set never reads as a
// member; side[], level[] and cmember[] are only touched for members.
struct NdShared {
  std::unique_ptr<std::atomic<int>[]> stamp;
  std::vector<int> side;     // 0 = L, 1 = R side
  std::vector<int> level;    // BFS levels
  std::vector<int> cmember;  // component of each vertex (local, then global)
  std::atomic<int> next_stamp{0};
  explicit NdShared(int n) : stamp(new std::atomic<int>[n]), side(n, 0), level(n, -1),
                             cmember(n, -1) {
    for (int v = 0; v < n; ++v) stamp[v].store(-1, std::memory_order_relaxed);
  }
};

// Dissection of one vertex set.  perm and cparent are this set's own, with
// components numbered from 0 in creation (post)order.
struct Dissector {
  const Csc& A;
  const double* xyz;
  int nd_small;
  NdShared& sh;
  std::vector<int>& side;
  std::vector<int>& level;
  std::vector<int>& cmember;
  std::vector<int> perm;
  std::vector<int> cparent;
  int cur = 0;

  Dissector(const Csc& A_, const double* xyz_, int nd_small_, NdShared& sh_)
      : A(A_), xyz(xyz_), nd_small(std::max(nd_small_, 1)), sh(sh_), side(sh_.side),
        level(sh_.level), cmember(sh_.cmember) {}

  bool member(int u) const { return sh.stamp[u].load(std::memory_order_relaxed) == cur; }

  void mark_set(const std::vector<int>& nodes) {
    cur = sh.next_stamp.fetch_add(1) + 1;
    for (int v : nodes) sh.stamp[v].store(cur, std::memory_order_relaxed);
  }

  // Pseudo-peripheral node of the subgraph induced by `nodes`
  // (George-Liu), BFS levels left in level[] and the level list in lv.
  int bfs(int root, std::vector<int>& order, std::vector<int>& lvstart) {
    order.clear();
    lvstart.clear();
    order.push_back(root);
    level[root] = 0;
    size_t head = 0;
    int curlev = -1;
    while (head < order.size()) {
      int v = order[head++];
      if (level[v] != curlev) {
        curlev = level[v];
        lvstart.push_back((int)head - 1);
      }
      for (int p = A.p[v]; p < A.p[v + 1]; ++p) {
        int u = A.i[p];
        if (u == v || !member(u) || level[u] >= 0) continue;
        level[u] = level[v] + 1;
        order.push_back(u);
      }
    }
    lvstart.push_back((int)order.size());
    return curlev;
  }

  void clear_levels(const std::vector<int>& order) {
    for (int v : order) level[v] = -1;
  }

  // Split `nodes` into L / S / R.  Returns false when no split is possible.
  bool split(const std::vector<int>& nodes, std::vector<int>& L,
             std::vector<int>& S, std::vector<int>& R) {
    L.clear();
    S.clear();
    R.clear();
    mark_set(nodes);
    bool ok = false;
    if (xyz) {
      double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
      double hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
      for (int v : nodes)
        for (int d = 0; d < 3; ++d) {
          lo[d] = std::min(lo[d], xyz[3 * (int64_t)v + d]);
          hi[d] = std::max(hi[d], xyz[3 * (int64_t)v + d]);
        }
      int axis = 0;
      double best = -1.0;
      for (int d = 0; d < 3; ++d)
        if (hi[d] - lo[d] > best) {
          best = hi[d] - lo[d];
          axis = d;
        }
      if (best > 0.0) {
        std::vector<double> c(nodes.size());
        for (size_t k = 0; k < nodes.size(); ++k)
          c[k] = xyz[3 * (int64_t)nodes[k] + axis];
        std::vector<double> tmp(c);
        size_t mid = tmp.size() / 2;
        std::nth_element(tmp.begin(), tmp.begin() + mid, tmp.end());
        double m = tmp[mid];
        // L = coord < m, R side = coord >= m; when L would be empty use <=.
        bool strict = false;
        for (double v : c)
          if (v < m) { strict = true; break; }
        for (size_t k = 0; k < nodes.size(); ++k) {
          bool left = strict ? (c[k] < m) : (c[k] <= m);
          side[nodes[k]] = left ? 0 : 1;
        }
        ok = true;
      }
    }
    if (!ok) {
      // level-structure bisection from a pseudo-peripheral node
      std::vector<int> order, lvstart;
      int root = nodes[0];
      int ecc = bfs(root, order, lvstart);
      for (int it = 0; it < 8; ++it) {
        int best = -1, bestdeg = 1 << 30;
        for (int k = lvstart[lvstart.size() - 2]; k < lvstart.back(); ++k) {
          int v = order[k];
          int deg = A.p[v + 1] - A.p[v];
          if (deg < bestdeg) { bestdeg = deg; best = v; }
        }
        clear_levels(order);
        int e2 = bfs(best, order, lvstart);
        if (e2 <= ecc) {
          if (e2 < ecc) { clear_levels(order); bfs(root, order, lvstart); }
          break;
        }
        ecc = e2;
        root = best;
      }
      int nlev = (int)lvstart.size() - 1;
      if (nlev < 3) {
        clear_levels(order);
        return false;
      }
      int half = (int)nodes.size() / 2, k = 1;
      while (k < nlev - 1 && lvstart[k + 1] <= half) ++k;
      for (int v : nodes) side[v] = 1;  // unreached -> R
      for (int q = 0; q < lvstart[k]; ++q) side[order[q]] = 0;
      clear_levels(order);
    }
    // separator: R-side vertices adjacent to the L side
    for (int v : nodes) {
      if (side[v] == 0) {
        L.push_back(v);
        continue;
      }
      bool adj = false;
      for (int p = A.p[v]; p < A.p[v + 1] && !adj; ++p) {
        int u = A.i[p];
        if (member(u) && side[u] == 0) adj = true;
      }
      if (adj) S.push_back(v);
      else R.push_back(v);
    }
    if (L.empty() || (R.empty() && S.empty())) return false;
    return true;
  }

  // Order a leaf component by recursive bisection without new components.
  void order_leaf(std::vector<int>& nodes) {
    if (nodes.size() <= 8) {
      std::sort(nodes.begin(), nodes.end());
      for (int v : nodes) perm.push_back(v);
      return;
    }
    std::vector<int> L, S, R;
    if (!split(nodes, L, S, R)) {
      std::sort(nodes.begin(), nodes.end());
      for (int v : nodes) perm.push_back(v);
      return;
    }
    order_leaf(L);
    order_leaf(R);
    std::sort(S.begin(), S.end());
    for (int v : S) perm.push_back(v);
  }

  int new_component() {
    cparent.push_back(-1);
    return (int)cparent.size() - 1;
  }

  std::vector<int> dissect(std::vector<int>& nodes) {
    std::vector<int> roots;
    if (nodes.empty()) return roots;
    std::vector<int> L, S, R;
    if ((int)nodes.size() <= nd_small || !split(nodes, L, S, R)) {
      int c = new_component();
      size_t start = perm.size();
      order_leaf(nodes);
      for (size_t k = start; k < perm.size(); ++k) cmember[perm[k]] = c;
      roots.push_back(c);
      return roots;
    }
    std::vector<int>().swap(nodes);
    std::vector<int> rl = dissect(L);
    std::vector<int> rr = dissect(R);
    if (S.empty()) {
      rl.insert(rl.end(), rr.begin(), rr.end());
      return rl;
    }
    int c = new_component();
    std::sort(S.begin(), S.end());
    for (int v : S) {
      perm.push_back(v);
      cmember[v] = c;
    }
    for (int r : rl) cparent[r] = c;
    for (int r : rr) cparent[r] = c;
    roots.push_back(c);
    return roots;
  }
};

// A dissected vertex set: its ordering, its components (local numbers) and
// the roots among them.
struct NdPart {
  std::vector<int> perm, cparent, roots;
};

// The recursion of Dissector::dissect with the two halves of the sets near
// the top run on two threads; the halves' results are stitched exactly as
// the sequential recursion would have produced them (left components, right
// components renumbered after them, then the separator).
NdPart dissect_parallel(const Csc& A, const double* xyz, int nd_small, NdShared& sh,
                        std::vector<int>& nodes, int depth) {
  NdPart out;
  Dissector d(A, xyz, nd_small, sh);
  std::vector<int> L, S, R;
  if (depth <= 0 || nodes.empty() || (int)nodes.size() <= d.nd_small ||
      !d.split(nodes, L, S, R)) {
    // sequential from here (dissect repeats the size and split tests)
    Dissector q(A, xyz, nd_small, sh);
    out.roots = q.dissect(nodes);
    out.perm.swap(q.perm);
    out.cparent.swap(q.cparent);
    return out;
  }
  std::vector<int>().swap(nodes);
  NdPart left;
  std::exception_ptr err;
  std::thread th([&]() {
    try {
      left = dissect_parallel(A, xyz, nd_small, sh, L, depth - 1);
    } catch (...) {
      err = std::current_exception();
    }
  });
  NdPart right = dissect_parallel(A, xyz, nd_small, sh, R, depth - 1);
  th.join();
  if (err) std::rethrow_exception(err);
  const int nl = (int)left.cparent.size(), nr = (int)right.cparent.size();
  out.perm.swap(left.perm);
  out.perm.insert(out.perm.end(), right.perm.begin(), right.perm.end());
  for (int v : right.perm) sh.cmember[v] += nl;
  out.cparent.swap(left.cparent);
  for (int p : right.cparent) out.cparent.push_back(p >= 0 ? p + nl : -1);
  if (S.empty()) {
    out.roots = left.roots;
    for (int r : right.roots) out.roots.push_back(r + nl);
    return out;
  }
  const int c = nl + nr;
  out.cparent.push_back(-1);
  std::sort(S.begin(), S.end());
  for (int v : S) {
    out.perm.push_back(v);
    sh.cmember[v] = c;
  }
  for (int r : left.roots) out.cparent[r] = c;
  for (int r : right.roots) out.cparent[r + nl] = c;
  out.roots.push_back(c);
  return out;
}

}  // namespace

NDResult nested_dissection(const Csc& A, const double* xyz, int nd_small) {
  NdShared sh(A.n);
  std::vector<int> all(A.n);
  std::iota(all.begin(), all.end(), 0);
  // two threads per level for the top levels, up to 16 at once
  unsigned hw = std::max(1u, std::min(std::thread::hardware_concurrency(), 16u));
  if (const char* e = std::getenv("RSC_ND_THREADS")) hw = (unsigned)std::max(1, std::atoi(e));
  int depth = 0;
  while ((2u << depth) <= hw) ++depth;
  NdPart part = dissect_parallel(A, xyz, nd_small, sh, all, depth);
  if ((int)part.perm.size() != A.n) throw std::runtime_error("ND lost vertices");
  NDResult r;
  r.perm = std::move(part.perm);
  r.cparent = std::move(part.cparent);
  r.cmember = std::move(sh.cmember);
  return r;
}

std::vector<Separator> separators_from_nd(const NDResult& nd, int n) {
  int nc = (int)nd.cparent.size();
  std::vector<Separator> seps(nc);
  for (int c = 0; c < nc; ++c)
    if (nd.cparent[c] >= 0) seps.at(nd.cparent[c]).nchildren += 1;
  // components must appear contiguously and in order in perm
  // (CHOLMOD_Environment.cpp:386-415)
  int current = 0;
  for (int k = 0; k < n; ++k) {
    int comp = nd.cmember[nd.perm[k]];
    if (comp == current + 1) current += 1;
    else if (comp != current)
      throw std::runtime_error("nested dissection components not contiguous");
    seps[comp].size += 1;
  }
  int first = 0;
  for (int c = 0; c < nc; ++c) {
    seps[c].first = first;
    seps[c].last = first + seps[c].size - 1;
    first += seps[c].size;
  }
  return seps;
}

// --------------------------------------------------------------------------
// Elimination tree and column counts (textbook algorithms: Liu's etree with
// path compression and the Gilbert-Ng-Peyton skeleton count, as in CSparse
// cs_etree / cs_counts which the reference calls at
// CSparse_Interface.cpp:224-288).
// --------------------------------------------------------------------------
std::vector<int> elimination_tree(const Csc& A, const std::vector<int>& pinv) {
  int n = A.n;
  std::vector<int> perm(n);
  for (int o = 0; o < n; ++o) perm[pinv[o]] = o;
  std::vector<int> parent(n, -1), ancestor(n, -1);
  for (int k = 0; k < n; ++k) {
    int j = perm[k];
    for (int p = A.p[j]; p < A.p[j + 1]; ++p) {
      int i = pinv[A.i[p]];
      while (i != -1 && i < k) {
        int inext = ancestor[i];
        ancestor[i] = k;
        if (inext == -1) parent[i] = k;
        i = inext;
      }
    }
  }
  return parent;
}

std::vector<int> column_counts(const Csc& A, const std::vector<int>& pinv,
                               const std::vector<int>& parent,
                               const std::vector<int>& post) {
  int n = A.n;
  std::vector<int> perm(n);
  for (int o = 0; o < n; ++o) perm[pinv[o]] = o;
  std::vector<int> delta(n), ancestor(n), maxfirst(n, -1), prevleaf(n, -1),
      first(n, -1);
  for (int k = 0; k < n; ++k) {
    int j = post[k];
    delta[j] = (first[j] == -1) ? 1 : 0;
    for (; j != -1 && first[j] == -1; j = parent[j]) first[j] = k;
  }
  for (int i = 0; i < n; ++i) ancestor[i] = i;
  for (int k = 0; k < n; ++k) {
    int j = post[k];
    if (parent[j] != -1) delta[parent[j]]--;
    // column j of the transpose of triu(C): rows i > j with C(j,i) != 0
    int oj = perm[j];
    for (int p = A.p[oj]; p < A.p[oj + 1]; ++p) {
      int i = pinv[A.i[p]];
      // leaf test (cs_leaf)
      if (i <= j || first[j] <= maxfirst[i]) continue;
      maxfirst[i] = first[j];
      int jprev = prevleaf[i];
      prevleaf[i] = j;
      int jleaf = (jprev == -1) ? 1 : 2;
      int q;
      if (jleaf == 1) {
        q = i;
      } else {
        for (q = jprev; q != ancestor[q]; q = ancestor[q]) {
        }
        for (int s = jprev, sp; s != q; s = sp) {
          sp = ancestor[s];
          ancestor[s] = q;
        }
      }
      delta[j]++;
      if (jleaf == 2) delta[q]--;
    }
    if (parent[j] != -1) ancestor[j] = parent[j];
  }
  for (int j = 0; j < n; ++j)
    if (parent[j] != -1) delta[parent[j]] += delta[j];
  return delta;
}

// --------------------------------------------------------------------------
// Relaxed supernodes (the partition rule of NestedDissection.cpp:1266-1667,
// itself CHOLMOD's supernodal analysis with relaxed amalgamation).  Stated
// as a rule:
//
//  1. A nested-dissection component with more than max_block_size columns
//     and at least one child is compressed, and so is every ancestor of a
//     compressed component.  A compressed component is one supernode that
//     nothing merges into or out of.
//  2. Between compressed components, consecutive uncompressed components
//     form one run of columns.  The run's first column starts a fundamental
//     supernode; column j continues the current one exactly when column j-1
//     is its only etree child with one more nonzero in L (parent[j-1] == j,
//     count[j-1] == count[j] + 1, j has at most one etree child).
//  3. Going from the second-to-last fundamental supernode down to the
//     first, supernode s may absorb s+1 when s+1 is (after earlier merges)
//     s's parent and neither is compressed.  With c0, c1 the column counts
//     and k0, k1 the first-column nonzero counts of s and s+1, z1 the
//     explicit zeros s+1 already carries, and t = c0 + c1:
//        t <= nrelax0                             -> merge, no new zeros
//        new = c0 (k1 + c0 - k0) == 0             -> merge, no new zeros
//        otherwise z = (z1 + new) / (t(t+1)/2 + t (k1 - c1)) and merge if
//        (t <= nrelax1 and z < zrelax0) or (t <= nrelax2 and z < zrelax1)
//        or z < zrelax2.
//     The ratio is evaluated in double and the zero count in integers, as
//     the reference does, so the decisions are identical.  After a merge s
//     has c0 + c1 columns, c0 + k1 nonzeros in its first column and z1 + new
//     zeros.
//  4. Every supernode not absorbed is emitted with its column range.
// --------------------------------------------------------------------------
namespace {

// One fundamental supernode while the amalgamation runs.
struct FundNode {
  int first = 0;         // first column
  int width = 0;         // columns, grown by merges
  int64_t nnz_first = 0; // nonzeros of L in its first column (diagonal included)
  int64_t zeros = 0;     // explicit zeros carried from earlier merges
  int parent = -1;       // fundamental supernode of the parent column, -1 = root
  bool whole = false;    // compressed component: never merged
  int taken_by = -1;     // the supernode that absorbed it, -1 while it stands
};

// Components to compress: large separators with children, and their
// ancestors (rule 1).
std::vector<char> components_to_compress(const std::vector<Separator>& seps,
                                         const std::vector<int>& sep_parent,
                                         int max_block_size) {
  std::vector<char> mark(seps.size(), 0);
  for (size_t c = 0; c < seps.size(); ++c) {
    if (seps[c].size <= max_block_size || seps[c].nchildren <= 0) continue;
    for (int a = (int)c; a >= 0 && !mark[a]; a = sep_parent[a]) mark[a] = 1;
    // an ancestor chain already marked above stops the walk
  }
  return mark;
}

// Fundamental supernodes of the column ranges (rule 2).
std::vector<FundNode> fundamental_partition(const std::vector<int>& parent,
                                            const std::vector<int>& count,
                                            const std::vector<Separator>& seps,
                                            const std::vector<char>& whole) {
  const int n = (int)parent.size();
  std::vector<int> kids(n, 0);
  for (int j = 0; j < n; ++j)
    if (parent[j] >= 0) ++kids[parent[j]];
  auto continues = [&](int j) {
    return parent[j - 1] == j && count[j - 1] == count[j] + 1 && kids[j] <= 1;
  };
  std::vector<FundNode> out;
  bool in_run = false;
  for (size_t c = 0; c < seps.size(); ++c) {
    if (whole[c]) {
      FundNode f;
      f.first = seps[c].first;
      f.whole = true;
      out.push_back(f);
      in_run = false;
      continue;
    }
    for (int j = seps[c].first; j <= seps[c].last; ++j) {
      if (in_run && continues(j)) continue;
      FundNode f;
      f.first = j;
      out.push_back(f);
      in_run = true;
    }
  }
  return out;
}

// Representative of fundamental supernode q after the merges so far.
int standing(std::vector<FundNode>& f, int q) {
  int r = q;
  while (f[r].taken_by >= 0) r = f[r].taken_by;
  // point the visited chain at the representative (a shortcut only: the
  // representative is the same either way)
  while (f[q].taken_by >= 0 && f[q].taken_by != r) {
    const int next = f[q].taken_by;
    f[q].taken_by = r;
    q = next;
  }
  return r;
}

struct RelaxParams {
  int nrelax[3];
  double zrelax[3];
};

// Rule 3 for child s (columns first) and its parent p = s + 1; on a merge
// returns true and the zero count of the merged supernode.
bool relaxed_merge(const FundNode& s, const FundNode& p, const RelaxParams& rp,
                   int64_t& zeros_after) {
  const int t = s.width + p.width;
  zeros_after = p.zeros;
  if (t <= rp.nrelax[0]) return true;
  // zeros the merge adds: the child's columns get the parent's rows
  const double added =
      (double)s.width * ((double)p.nnz_first + s.width - (double)s.nnz_first);
  if (added == 0) return true;
  const int64_t added_i = (int64_t)s.width * (p.nnz_first + s.width - s.nnz_first);
  const double tt = (double)t;
  const double entries = tt * (tt + 1) / 2 + tt * ((double)p.nnz_first - p.width);
  const double ratio = ((double)p.zeros + added) / entries;
  zeros_after = p.zeros + added_i;
  return (t <= rp.nrelax[1] && ratio < rp.zrelax[0]) ||
         (t <= rp.nrelax[2] && ratio < rp.zrelax[1]) || ratio < rp.zrelax[2];
}

}  // namespace

std::vector<SupernodeSpec> build_relaxed_supernodes(
    const std::vector<int>& parent, const std::vector<int>& count,
    const std::vector<Separator>& seps, const std::vector<int>& sep_parent,
    int max_block_size, int nrelax0, int nrelax1, int nrelax2, double zrelax0,
    double zrelax1, double zrelax2) {
  const int n = (int)parent.size();
  const std::vector<char> whole = components_to_compress(seps, sep_parent, max_block_size);
  std::vector<FundNode> f = fundamental_partition(parent, count, seps, whole);
  const int nf = (int)f.size();
  // widths, first-column counts and parents; owner[] maps a column to its
  // fundamental supernode
  std::vector<int> owner(n, -1);
  for (int q = 0; q < nf; ++q) {
    const int end = q + 1 < nf ? f[q + 1].first : n;
    f[q].width = end - f[q].first;
    f[q].nnz_first = f[q].first < n ? count[f[q].first] : 0;
    for (int j = f[q].first; j < end; ++j) owner[j] = q;
  }
  for (int q = 0; q < nf; ++q) {
    const int last = (q + 1 < nf ? f[q + 1].first : n) - 1;
    const int pc = last >= 0 ? parent[last] : -1;
    f[q].parent = pc >= 0 ? owner[pc] : -1;
  }
  const RelaxParams rp{{nrelax0, nrelax1, nrelax2}, {zrelax0, zrelax1, zrelax2}};
  for (int s = nf - 2; s >= 0; --s) {
    if (f[s].parent < 0 || f[s].whole || f[s + 1].whole) continue;
    if (standing(f, f[s].parent) != s + 1) continue;
    int64_t zeros = 0;
    if (!relaxed_merge(f[s], f[s + 1], rp, zeros)) continue;
    f[s].zeros = zeros;
    f[s].nnz_first = f[s].width + f[s + 1].nnz_first;
    f[s].width += f[s + 1].width;
    f[s + 1].taken_by = s;
  }
  // rule 4: the standing supernodes in column order
  std::vector<SupernodeSpec> out;
  for (int q = 0; q < nf; ++q) {
    if (f[q].taken_by >= 0) continue;
    SupernodeSpec sp;
    sp.first = f[q].first;
    sp.compress = f[q].whole;
    if (!out.empty()) out.back().last = sp.first - 1;
    out.push_back(sp);
  }
  if (!out.empty()) out.back().last = n - 1;
  return out;
}

// --------------------------------------------------------------------------
// Separator tree (SeparatorTree.cpp:189-367, 470-500, 530-589): longest-axis
// split of the bounding box, std::sort of the points along that axis, split
// at size/2, with the "full tree" rule that keeps both subtrees at equal
// height.
// --------------------------------------------------------------------------
namespace {

struct SepTree {
  const std::vector<double>& pts;  // 3 per point

  explicit SepTree(const std::vector<double>& p) : pts(p) {}

  int longest_axis(const std::vector<int>& data) const {
    double mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    double mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int v : data)
      for (int d = 0; d < 3; ++d) {
        mn[d] = std::min(mn[d], pts[3 * (size_t)v + d]);
        mx[d] = std::max(mx[d], pts[3 * (size_t)v + d]);
      }
    double maxLength = -1.0;
    int maxAxis = 0;
    for (int d = 0; d < 3; ++d)
      if (mx[d] - mn[d] > maxLength) {
        maxAxis = d;
        maxLength = mx[d] - mn[d];
      }
    return maxAxis;
  }

  // Builds the subtree of `data` into (ord, lv) and returns its height.
  // min_height > 0 forces splits below maxLeaf until that height is reached;
  // it is used only when the reference's one-level "full tree" override
  // leaves sibling heights unequal (SeparatorTree.cpp:322-356 asserts equality
  // there, and DiagonalBlock.cpp:106-121 needs a power-of-two leaf count).
  int build(std::vector<int> data, int maxLeaf, int min_height, std::vector<int>& ord,
            std::vector<int>& lv) {
    if ((int)data.size() <= maxLeaf && min_height <= 0) {
      ord.insert(ord.end(), data.begin(), data.end());
      lv.push_back((int)data.size());
      return 0;
    }
    int axis = longest_axis(data);
    const std::vector<double>& P = pts;
    std::sort(data.begin(), data.end(), [&P, axis](int a, int b) {
      return P[3 * (size_t)a + axis] < P[3 * (size_t)b + axis];
    });
    int leftSize = (int)data.size() / 2;
    std::vector<int> L(data.begin(), data.begin() + leftSize);
    std::vector<int> R(data.begin() + leftSize, data.end());
    if ((int)L.size() == (int)data.size() || R.empty()) {
      ord.insert(ord.end(), data.begin(), data.end());
      lv.push_back((int)data.size());
      return 0;
    }
    const int sub = std::max(min_height - 1, 0);
    int lmax = ((int)R.size() > maxLeaf && (int)L.size() <= maxLeaf) ? (int)L.size() - 1
                                                                      : maxLeaf;
    int rmax = ((int)L.size() > maxLeaf && (int)R.size() <= maxLeaf) ? (int)R.size() - 1
                                                                      : maxLeaf;
    std::vector<int> lo, ll, ro, rl;
    int lh = build(L, lmax, sub, lo, ll);
    int rh = build(R, rmax, sub, ro, rl);
    if (lh < rh) {
      lo.clear();
      ll.clear();
      lh = build(L, lmax, rh, lo, ll);
    } else if (rh < lh) {
      ro.clear();
      rl.clear();
      rh = build(R, rmax, lh, ro, rl);
    }
    ord.insert(ord.end(), lo.begin(), lo.end());
    ord.insert(ord.end(), ro.begin(), ro.end());
    lv.insert(lv.end(), ll.begin(), ll.end());
    lv.insert(lv.end(), rl.begin(), rl.end());
    return 1 + std::max(lh, rh);
  }
};

}  // namespace

void separator_tree_order(const std::vector<double>& pts3, int max_leaf,
                          std::vector<int>& order, std::vector<int>& leaf_sizes) {
  SepTree t(pts3);
  std::vector<int> data(pts3.size() / 3);
  std::iota(data.begin(), data.end(), 0);
  order.clear();
  leaf_sizes.clear();
  t.build(std::move(data), max_leaf, 0, order, leaf_sizes);
}

std::vector<std::vector<int>> reorder_nested_separators(
    const std::vector<SupernodeSpec>& nodes, const double* xyz,
    int max_block_size, std::vector<int>& perm, bool reorder) {
  std::vector<std::vector<int>> blocks(nodes.size());
  for (size_t s = 0; s < nodes.size(); ++s) {
    const SupernodeSpec& nd = nodes[s];
    int size = nd.ncols();
    if (size <= max_block_size) continue;
    std::vector<double> pts(3 * (size_t)size, 0.0);
    if (xyz)
      for (int k = 0; k < size; ++k)
        for (int d = 0; d < 3; ++d)
          pts[3 * (size_t)k + d] = xyz[3 * (size_t)perm[nd.first + k] + d];
    std::vector<int> order, leaves;
    separator_tree_order(pts, max_block_size, order, leaves);
    if (reorder && xyz) {
      std::vector<int> old(perm.begin() + nd.first, perm.begin() + nd.last + 1);
      for (int k = 0; k < size; ++k) perm[nd.first + k] = old[order[k]];
    }
    std::vector<int>& b = blocks[s];
    b.push_back(0);
    for (int l : leaves) b.push_back(b.back() + l);
  }
  return blocks;
}

}  // namespace rsc
