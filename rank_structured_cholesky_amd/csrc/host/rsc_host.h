// Host-side data structures of the rank-structured Cholesky: the input
// matrix, the nested-dissection ordering, the supernode partition and the
// symbolic plan that the device factor and solves execute.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <exception>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/rsc.h"

namespace rsc {

// Full symmetric matrix in CSC form (the layout of the reference's
// SparseColumnMatrix, SPARSE_MATRIX.h:187-317).
struct Csc {
  int n = 0;
  std::vector<int> p;     // n+1
  std::vector<int> i;     // nnz
  std::vector<double> x;  // nnz (may be empty for pattern-only use)
  int64_t nnz() const { return p.empty() ? 0 : p[n]; }
};

// Output of nested dissection in CHOLMOD's format
// (CHOLMOD_Environment.cpp:326-432): perm[new] = old, components numbered
// in postorder, cmember[old] = component, cparent[component] (-1 = root).
struct NDResult {
  std::vector<int> perm;
  std::vector<int> cparent;
  std::vector<int> cmember;
};

// One supernode specification (Ordering.h:73-77).
struct SupernodeSpec {
  int first = 0;  // first column (permuted order)
  int last = 0;   // last column, inclusive
  bool compress = false;
  int ncols() const { return last - first + 1; }
};

// ---- ordering.cpp ----
// Geometric nested dissection on the matrix graph guided by coordinates
// (xyz = n*3) or, when xyz is null, level-structure bisection.
NDResult nested_dissection(const Csc& A, const double* xyz, int nd_small);

// Elimination tree of C = P A P^T (cs_etree on the upper triangle).
std::vector<int> elimination_tree(const Csc& A, const std::vector<int>& pinv);
// Column counts of chol(C) with a given postorder (cs_counts, ata = 0),
// diagonal included.
std::vector<int> column_counts(const Csc& A, const std::vector<int>& pinv,
                               const std::vector<int>& parent,
                               const std::vector<int>& post);

struct Separator {
  int size = 0;
  int first = 0, last = -1;
  int nchildren = 0;
};
std::vector<Separator> separators_from_nd(const NDResult& nd, int n);

// NestedDissection::BuildRelaxedSupernodes (NestedDissection.cpp:1266-1667).
std::vector<SupernodeSpec> build_relaxed_supernodes(
    const std::vector<int>& parent, const std::vector<int>& colcount,
    const std::vector<Separator>& seps, const std::vector<int>& sep_parent,
    int max_block_size, int nrelax0 = 4, int nrelax1 = 16, int nrelax2 = 48,
    double zrelax0 = 0.8, double zrelax1 = 0.1, double zrelax2 = 0.05);

// NestedDissection::ReorderNestedSeparators (NestedDissection.cpp:728-910)
// with SeparatorTree (SeparatorTree.cpp:189-589).  Updates perm in place
// (when reorder) and returns, per supernode, its diagonal block leaf ranges
// as relative offsets [0, b1, b2, ..., ncols] (empty when not split).
std::vector<std::vector<int>> reorder_nested_separators(
    const std::vector<SupernodeSpec>& nodes, const double* xyz,
    int max_block_size, std::vector<int>& perm, bool reorder);

// Separator tree leaf partition of a point list (exposed for tests):
// order[] receives the point order and leaf_sizes the leaf sizes.
void separator_tree_order(const std::vector<double>& pts3, int max_leaf,
                          std::vector<int>& order, std::vector<int>& leaf_sizes);

// ---- generators.cpp ----
Csc poisson2d(int N);  // 5-point, Dirichlet eliminated
Csc poisson3d_lognormal(int N, double sigma, uint64_t seed);
Csc elasticity3d(int NE, double sigma, double nu, uint64_t seed);
std::vector<double> grid_positions(int nx, int ny, int nz);
std::vector<double> uniform_rhs(int64_t count, uint64_t seed);

// ---- rank heuristic (FactorManager.cpp:7283-7338) ----
int block_rank(int nrows, int ncols, double constant, int oversampling,
               int min_rank, int max_rank);

// ---- symbolic.cpp ----
// HODLR tree of a compressed separator's diagonal (DiagonalBlock.cpp:92-173,
// 735-790, 1122-1146).  Node 0 is unused; nodes are stored in a flat array.
struct HNode {
  int r0 = 0, r1 = 0;      // full block rows [r0, r1) relative to the supernode
  int left = -1, right = -1, parent = -1;
  bool leaf() const { return left < 0; }
  bool explicit_build = false;
  int rank = 0;            // off-diagonal rank (internal nodes)
  // block rows = right child's range, cols = left child's range
  int64_t off_V = -1, off_U = -1;  // factor storage (elements)
  // solve helpers: W = L_right^{-1} V (rows x r), X = L_left^{-T} U (cols x r)
  int64_t off_W = -1, off_X = -1;
  int height = 0;                  // 0 for leaves
  int64_t off_leaf = -1;           // factor storage of a leaf (b x b)
  int64_t rng_off = -1;            // offset of the sketch in the RNG stream
};

// One entry of the in-order traversal that drives the diagonal
// factorization (FactorManager.cpp:8907-8992).
struct HStep {
  int node;
  enum Kind { EXPLICIT = 0, LEAF = 1, OFFDIAG = 2 } kind;
};

struct Supernode {
  int first = 0, ncols = 0;
  bool compress = false;
  int parent = -1;              // supernodal parent (-1 = root)
  std::vector<int> children;
  std::vector<int> rows;        // off-diagonal rows, sorted, > last column
  int level = 0;                // height: 0 = leaf
  // map of this node's off-diagonal rows into the parent's front
  // (positions relative to the parent's front: cols first, then rows)
  std::vector<int> relmap;
  // storage (elements of the factor array)
  int64_t off_panel = -1;       // (ncols+R) x ncols column-major (uncompressed)
  int64_t off_work_panel = -1;  // workspace panel for compressed nodes
  int64_t off_update = -1;      // R x R update matrix (workspace)
  // compressed
  int rank = 0;
  int64_t off_V = -1, off_U = -1;
  int64_t off_Xoff = -1;        // L_D^{-T} U (ncols x rank), the apply's form of U
  int64_t off_dinv = -1;        // dense L_D^{-T} (ncols x ncols) of a root separator
  int64_t rng_off = -1;
  std::vector<HNode> hnodes;    // HODLR tree (compressed only); [0] = root
  std::vector<HStep> hsteps;
  int R() const { return (int)rows.size(); }
};

struct Plan {
  int n = 0;
  rsc_options opt{};
  std::vector<int> perm, pinv;
  std::vector<SupernodeSpec> specs;
  std::vector<std::vector<int>> diag_blocks;  // per supernode leaf offsets
  std::vector<Supernode> nodes;
  std::vector<int> col2node;
  std::vector<std::vector<int>> levels;       // node ids per level
  int64_t factor_elems = 0;                   // stored factor (elements)
  int64_t work_elems = 0;                     // panel workspace of compressed
  int64_t update_elems = 0;                   // update matrices
  int64_t rng_count = 0;                      // Gaussian draws
  int max_front = 0;
  double factor_flops = 0;
  double apply_bytes_per_elem = 0;            // factor element traversals
};

// Build the symbolic plan (ordering + structure + storage + RNG layout).
// nd may be null (computed here).  diag_rank_constant overrides the option
// (indefinite-diagonal retry).
void build_plan(const Csc& A, const double* xyz, const rsc_options& opt,
                const NDResult* nd_in, Plan& plan);
// Index of the root of a compressed supernode's diagonal tree.
int hodlr_root(const Supernode& s);
// Recompute ranks / storage / RNG layout for a new diagonal rank constant
// without redoing the ordering (SparseSolver.cpp:192-204).
void replan_ranks(Plan& plan, double diagonal_rank_constant);
// Offsets of the update matrices and compressed work panels, reused across
// levels (sets update_elems and work_elems).
void place_transient(Plan& plan);

// Gaussian sketch stream: MT19937 seeded with `seed`, Box-Muller as in
// MERSENNETWISTER::randNorm (MERSENNETWISTER.h:170-177), count draws.
// Runs on up to 16 host threads; the values do not depend on the count.
// The float form rounds each double draw to nearest.
void gaussian_stream(uint32_t seed, int64_t count, double* out);
void gaussian_stream(uint32_t seed, int64_t count, float* out);
void gaussian_stream_threads(uint32_t seed, int64_t count, double* out, int nthreads);

// Permuted lower triangle of A: for each permuted column, (row, value) with
// row >= col (Ordering::reorderMatrix with lowerOnly, Ordering.cpp:162).
struct LowerCsc {
  std::vector<int> p, i;
  std::vector<int64_t> src;  // index of the entry in the original A
};
LowerCsc permuted_lower(const Csc& A, const std::vector<int>& pinv);

std::string& last_error();

// Host threads for the parallel setup loops (at most 16).
int stream_threads();

// f(begin, end, thread) over [0, n) in chunks of `grain`, on up to
// stream_threads() threads (the calling thread included).  Chunks are
// handed out dynamically; f must write only data owned by its range.  An
// exception thrown by f stops the loop and is rethrown here.
template <typename F>
void parallel_for(int64_t n, int64_t grain, F f) {
  if (n <= 0) return;
  grain = std::max<int64_t>(grain, 1);
  const int64_t nchunks = (n + grain - 1) / grain;
  const int nt = (int)std::min<int64_t>(stream_threads(), nchunks);
  if (nt <= 1) {
    f((int64_t)0, n, 0);
    return;
  }
  std::atomic<int64_t> next{0};
  std::vector<std::exception_ptr> err(nt);
  auto work = [&](int t) {
    try {
      for (;;) {
        const int64_t c = next.fetch_add(1);
        if (c >= nchunks) return;
        f(c * grain, std::min(n, (c + 1) * grain), t);
      }
    } catch (...) {
      err[t] = std::current_exception();
      next.store(nchunks);  // the other threads stop at their next chunk
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto& th : pool) th.join();
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
}

}  // namespace rsc
