// ORACLE -- test infrastructure only.  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load this library, and only as the
// checker / CPU baseline; the product (librsc.so) never calls it.
//
// A plain FP64 CPU restatement of the reference's numeric path, left-looking
// like the reference (all paths relative to the reference root):
//   * row structure of each supernode: Supernode::findSystemRows +
//     propagateFillIn (src/rschol/node/Supernode.cpp:11771-11967)
//   * left-looking supernode loop with descendant link lists
//     (src/rschol/solver/FactorManager.cpp:398-804); dense descendant
//     updates (Supernode::buildUpdateMatrix / applyUpdate, Supernode.cpp:
//     667-1114); compressed descendants contribute (V U^T)(V U^T)^T
//     (Supernode::compressedDiagonalUpdate, Supernode.cpp:1266-1429)
//   * compressed diagonal: DiagonalBlock tree (DiagonalBlock.cpp:92-173,
//     735-790, 1122-1146) walked in traversal order
//     (FactorManager::factorCompressedDiagonal, FactorManager.cpp:8907-8992):
//     explicit blocks -> POTRF + recursive sketch of each off-diagonal child
//     (DiagonalBlock.cpp:393-453, 1384-1615); recursive off-diagonal blocks
//     -> sketch of S L_left^{-T} (FactorManager.cpp:8998-9036); ancestor
//     contributions W W^T with W = V U^T (DiagonalBlock::updateDiagonal,
//     DiagonalBlock.cpp:996-1066)
//   * off-diagonal compression: G (rows x r) from the sketch stream,
//     Y = L21^T G, power iteration, Householder QR, V = L21 U
//     (FactorManager.cpp:8193-8297, 8672-8779, 5691-5751; Supernode::
//     lowRankQR, Supernode.cpp:3308-3361)
//   * ranks: FactorManager::blockRank / diagonalBlockRank
//     (FactorManager.cpp:7283-7338)
//   * sketch stream: MT19937 seeded 152 + Box-Muller (MathUtil.cpp:100-112,
//     MERSENNETWISTER.h:170-177), drawn in factorization order
//   * solves: FactorManager::forwardSolve / backwardSolve
//     (FactorManager.cpp:6245-6341), DiagonalBlock::BlockNode::diagonalSolve
//     (DiagonalBlock.cpp:1306-1378); permutation (FactorPreconditioner.cpp:
//     135-162)
//   * PCG: PCG_SOLVER::solvePCG (PCG_SOLVER.cpp:226-363)
// The ordering (permutation, supernode partition, diagonal block ranges) is
// an input.  Dense kernels are plain loops (no BLAS/LAPACK here), blocked
// and spread over OpenMP threads without changing any element's summation
// order, so results do not depend on the thread count.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <random>
#include <vector>

#include <omp.h>

namespace {

typedef std::vector<double> Vec;

// column-major dense matrix view helpers --------------------------------------
inline double& at(double* A, int ld, int i, int j) { return A[i + (int64_t)j * ld]; }

// The dense kernels below are blocked for cache reuse and run on OpenMP
// threads over independent output blocks.  Every output element is
// accumulated by one thread over the summation index in increasing order,
// as the textbook triple loop does, so the results do not depend on the
// number of threads.
constexpr int64_t kParallelWork = 1 << 15;  // flops below which one thread runs

// C(m x n) (+)= alpha * A(m x k) * op(B), op(B)(t, j) = bt ? B[j + t ldb] :
// B[t + j ldb]; lower: only elements with i >= j + diag_off.  Each element
// sums t = 0..k-1 in order; a zero B entry is skipped.
void gemm_core(int m, int n, int k, double alpha, const double* A, int lda,
               const double* B, int ldb, bool bt, double* C, int ldc, bool zero,
               bool lower, int diag_off) {
  if (m <= 0 || n <= 0) return;
  constexpr int MB = 256, NB = 16, KB = 128;
  const int mt = (m + MB - 1) / MB, nt = (n + NB - 1) / NB;
  const bool par = (int64_t)m * n * std::max(k, 1) > kParallelWork;
#pragma omp parallel for collapse(2) schedule(dynamic, 1) if (par)
  for (int jb = 0; jb < nt; ++jb)
    for (int ib = 0; ib < mt; ++ib) {
      const int j0 = jb * NB, j1 = std::min(n, j0 + NB);
      const int i0 = ib * MB, i1 = std::min(m, i0 + MB);
      if (lower && i1 - 1 < j0 + diag_off) continue;
      if (zero)
        for (int j = j0; j < j1; ++j)
          for (int i = i0; i < i1; ++i) C[i + (int64_t)j * ldc] = 0.0;
      for (int t0 = 0; t0 < k; t0 += KB) {
        const int t1 = std::min(k, t0 + KB);
        for (int j = j0; j < j1; ++j) {
          double* c = C + (int64_t)j * ldc;
          const int ilo = lower ? std::max(i0, j + diag_off) : i0;
          for (int t = t0; t < t1; ++t) {
            const double b = alpha * (bt ? B[j + (int64_t)t * ldb] : B[t + (int64_t)j * ldb]);
            if (b == 0.0) continue;
            const double* a = A + (int64_t)t * lda;
            for (int i = ilo; i < i1; ++i) c[i] += a[i] * b;
          }
        }
      }
    }
}

// C(m x n) += alpha * A(m x k) * B(n x k)^T
void gemm_nt(int m, int n, int k, double alpha, const double* A, int lda,
             const double* B, int ldb, double* C, int ldc) {
  gemm_core(m, n, k, alpha, A, lda, B, ldb, true, C, ldc, false, false, 0);
}
// C(m x n) = A(m x k) * B(k x n)
void gemm_nn(int m, int n, int k, const double* A, int lda, const double* B, int ldb,
             double* C, int ldc) {
  gemm_core(m, n, k, 1.0, A, lda, B, ldb, false, C, ldc, true, false, 0);
}
// C(m x n) = A(k x m)^T * B(k x n): A^T is formed once, then each element
// sums t = 0..k-1 in order as the dot product does
void gemm_tn(int m, int n, int k, const double* A, int lda, const double* B, int ldb,
             double* C, int ldc) {
  std::vector<double> At((size_t)m * k);
#pragma omp parallel for schedule(static) if ((int64_t)m * k > kParallelWork)
  for (int i = 0; i < m; ++i)
    for (int t = 0; t < k; ++t) At[i + (size_t)t * m] = A[t + (int64_t)i * lda];
  gemm_core(m, n, k, 1.0, At.data(), m, B, ldb, false, C, ldc, true, false, 0);
}

// In-place lower Cholesky (left-looking by panels of NB columns: the panel
// first takes the products of all columns to its left, then is factored
// column by column; every element subtracts its terms in increasing column
// order).  Returns 0 or the 1-based failing column.
int potrf(int n, double* A, int lda) {
  constexpr int NB = 64;
  for (int j0 = 0; j0 < n; j0 += NB) {
    const int j1 = std::min(n, j0 + NB);
    // A(j0:n, j0:j1) -= A(j0:n, 0:j0) A(j0:j1, 0:j0)^T, lower part
    gemm_core(n - j0, j1 - j0, j0, -1.0, A + j0, lda, A + j0, lda, true,
              A + j0 + (int64_t)j0 * lda, lda, false, true, 0);
    for (int j = j0; j < j1; ++j) {
      double d = at(A, lda, j, j);
      for (int t = j0; t < j; ++t) d -= at(A, lda, j, t) * at(A, lda, j, t);
      if (!(d > 0.0)) return j + 1;
      d = std::sqrt(d);
      at(A, lda, j, j) = d;
      const int rows = n - j - 1;
#pragma omp parallel for schedule(static) if ((int64_t)rows * (j - j0 + 1) > kParallelWork)
      for (int i0 = j + 1; i0 < n; i0 += 256) {
        const int i1 = std::min(n, i0 + 256);
        for (int t = j0; t < j; ++t) {
          const double ljt = at(A, lda, j, t);
          if (ljt == 0.0) continue;
          for (int i = i0; i < i1; ++i) at(A, lda, i, j) -= at(A, lda, i, t) * ljt;
        }
        for (int i = i0; i < i1; ++i) at(A, lda, i, j) /= d;
      }
    }
  }
  return 0;
}

// X(m x n) <- X L^{-T}, L n x n lower (left-looking by panels of columns;
// each element subtracts its terms in increasing column order)
void trsm_rlt(int m, int n, const double* L, int ldl, double* X, int ldx) {
  constexpr int NB = 64;
  for (int j0 = 0; j0 < n; j0 += NB) {
    const int j1 = std::min(n, j0 + NB);
    gemm_core(m, j1 - j0, j0, -1.0, X, ldx, L + j0, ldl, true, X + (int64_t)j0 * ldx, ldx,
              false, false, 0);
#pragma omp parallel for schedule(static) if ((int64_t)m * NB * NB > kParallelWork)
    for (int i0 = 0; i0 < m; i0 += 256) {
      const int i1 = std::min(m, i0 + 256);
      for (int j = j0; j < j1; ++j) {
        double* xj = X + (int64_t)j * ldx;
        for (int t = j0; t < j; ++t) {
          const double l = L[j + (int64_t)t * ldl];
          if (l == 0.0) continue;
          const double* xt = X + (int64_t)t * ldx;
          for (int i = i0; i < i1; ++i) xj[i] -= xt[i] * l;
        }
        const double d = L[j + (int64_t)j * ldl];
        for (int i = i0; i < i1; ++i) xj[i] /= d;
      }
    }
  }
}

// Householder QR of Y (m x r, m >= r) -> explicit Q in place (geqrf+orgqr)
void householder_q(int m, int r, double* Y, int ldy) {
  Vec tau(r), v(m);
  for (int j = 0; j < r; ++j) {
    double* a = Y + (int64_t)j * ldy;
    double nrm = 0;
    for (int i = j; i < m; ++i) nrm += a[i] * a[i];
    nrm = std::sqrt(nrm);
    if (nrm == 0.0) {
      tau[j] = 0.0;
      continue;
    }
    double alpha = a[j];
    double beta = (alpha >= 0) ? -nrm : nrm;
    double t = (beta - alpha) / beta;
    double scale = 1.0 / (alpha - beta);
    for (int i = j + 1; i < m; ++i) a[i] *= scale;
    a[j] = beta;
    tau[j] = t;
#pragma omp parallel for schedule(static) if ((int64_t)(m - j) * (r - j) > kParallelWork)
    for (int c = j + 1; c < r; ++c) {
      double* b = Y + (int64_t)c * ldy;
      double s = b[j];
      for (int i = j + 1; i < m; ++i) s += a[i] * b[i];
      s *= t;
      b[j] -= s;
      for (int i = j + 1; i < m; ++i) b[i] -= s * a[i];
    }
  }
  // form Q: apply reflectors backwards to the first r columns of I
  std::vector<Vec> refl(r);
  for (int j = 0; j < r; ++j) {
    refl[j].assign(m, 0.0);
    refl[j][j] = 1.0;
    for (int i = j + 1; i < m; ++i) refl[j][i] = Y[i + (int64_t)j * ldy];
  }
  for (int c = 0; c < r; ++c) {
    double* q = Y + (int64_t)c * ldy;
    for (int i = 0; i < m; ++i) q[i] = (i == c) ? 1.0 : 0.0;
  }
  for (int j = r - 1; j >= 0; --j) {
    if (tau[j] == 0.0) continue;
    const Vec& hv = refl[j];
#pragma omp parallel for schedule(static) if ((int64_t)(m - j) * (r - j) > kParallelWork)
    for (int c = j; c < r; ++c) {
      double* q = Y + (int64_t)c * ldy;
      double s = 0;
      for (int i = j; i < m; ++i) s += hv[i] * q[i];
      s *= tau[j];
      for (int i = j; i < m; ++i) q[i] -= s * hv[i];
    }
  }
}

struct Opts {
  double rank_constant, diag_rank_constant;
  int oversampling, min_rank, max_rank, power_iterations, explicit_threshold;
  uint32_t seed;
};

int rank_of(int rows, int cols, double c, const Opts& o) {
  double nz = std::sqrt((double)std::min(rows, cols));
  int b = (int)(nz * std::log2(nz) * c) + o.oversampling;
  b = std::max(b, o.min_rank);
  b = std::min(b, rows);
  b = std::min(b, cols);
  if (o.max_rank > 0) b = std::min(b, o.max_rank);
  return b;
}

struct Gauss {
  std::mt19937 g;
  explicit Gauss(uint32_t s) : g(s) {}
  double next() {
    double r = std::sqrt(-2.0 * std::log(1.0 - ((double)g() + 0.5) * (1.0 / 4294967296.0)));
    double phi = 2.0 * 3.14159265358979323846264338328 * ((double)g() * (1.0 / 4294967296.0));
    return r * std::cos(phi);
  }
};

struct TNode {
  int r0, r1, left = -1, right = -1, parent = -1;
  bool explicit_build = false;
  int rank = 0;
  Vec leaf;     // b x b lower factor (leaves)
  Vec V, U;     // rows(right) x r, cols(left) x r
};

struct Node {
  int first, ncols;
  bool compress;
  std::vector<int> rows;
  int parent = -1;
  Vec panel;    // (c+R) x c dense factor (uncompressed)
  std::vector<TNode> tree;
  int root = 0;
  std::vector<std::pair<int, int>> steps;  // (tree node, kind 0 explicit 1 leaf 2 offdiag)
  int rank = 0;
  Vec V, U;     // off-diagonal low-rank pair
};

struct Oracle {
  int n;
  std::vector<int> Ap, Ai;
  Vec Ax;
  std::vector<int> perm, pinv;
  std::vector<Node> nodes;
  std::vector<int> col2node;
  Opts opt;
  int attempts = 0;
};

// Right solve X L^{-T} through the tree rooted at q (DiagonalBlock.cpp:1151-1300)
void tree_solve_right(Node& s, int q, double* X, int m, int ldx) {
  TNode& h = s.tree[q];
  if (h.left < 0) {
    int b = h.r1 - h.r0;
    trsm_rlt(m, b, h.leaf.data(), b, X, ldx);
    return;
  }
  TNode& L = s.tree[h.left];
  TNode& R = s.tree[h.right];
  double* Xl = X + (int64_t)(L.r0 - h.r0) * ldx;
  double* Xr = X + (int64_t)(R.r0 - h.r0) * ldx;
  tree_solve_right(s, h.left, Xl, m, ldx);
  int cl = L.r1 - L.r0, cr = R.r1 - R.r0, r = h.rank;
  Vec T((size_t)m * r);
  gemm_nn(m, r, cl, Xl, ldx, h.U.data(), cl, T.data(), m);
  gemm_nt(m, cr, r, -1.0, T.data(), m, h.V.data(), cr, Xr, ldx);
  tree_solve_right(s, h.right, Xr, m, ldx);
}

// randomized range finder on B (rows x cols), returns V (rows x r), U (cols x r)
void range_finder(const double* B, int ldb, int rows, int cols, int r, Gauss& g,
                  int power, Vec& V, Vec& U) {
  Vec G((size_t)rows * r);
  for (int a = 0; a < rows; ++a)
    for (int b = 0; b < r; ++b) G[a + (size_t)b * rows] = g.next();  // row-major fill
  U.assign((size_t)cols * r, 0.0);
  gemm_tn(cols, r, rows, B, ldb, G.data(), rows, U.data(), cols);
  Vec Z((size_t)rows * r);
  for (int it = 0; it < power; ++it) {
    gemm_nn(rows, r, cols, B, ldb, U.data(), cols, Z.data(), rows);
    gemm_tn(cols, r, rows, B, ldb, Z.data(), rows, U.data(), cols);
  }
  householder_q(cols, r, U.data(), cols);
  V.assign((size_t)rows * r, 0.0);
  gemm_nn(rows, r, cols, B, ldb, U.data(), cols, V.data(), rows);
}

// S(rows x rows block at S, lds) -= W W^T with W = V U^T (updateDiagonal)
void lowrank_syrk(double* S, int lds, int rows, const Vec& V, const Vec& U, int cols, int r) {
  Vec W((size_t)rows * cols, 0.0);
  gemm_nt(rows, cols, r, 1.0, V.data(), rows, U.data(), cols, W.data(), rows);
  gemm_nt(rows, rows, cols, -1.0, W.data(), rows, W.data(), rows, S, lds);
}

int factor_compressed_diag(Node& s, double* P, int ld, Gauss& g, const Opts& o) {
  for (auto& st : s.steps) {
    TNode& h = s.tree[st.first];
    const int b = h.r1 - h.r0;
    if (st.second == 0 || st.second == 1) {
      double* D = P + h.r0 + (int64_t)h.r0 * ld;
      if (potrf(b, D, ld) != 0) return 1;
      std::function<void(int)> pre = [&](int q) {
        TNode& x = s.tree[q];
        if (x.left < 0) {
          int bb = x.r1 - x.r0;
          x.leaf.assign((size_t)bb * bb, 0.0);
          for (int j = 0; j < bb; ++j)
            for (int i = j; i < bb; ++i)
              x.leaf[i + (size_t)j * bb] = P[(x.r0 + i) + (int64_t)(x.r0 + j) * ld];
          return;
        }
        TNode& L = s.tree[x.left];
        TNode& R = s.tree[x.right];
        const double* B = P + R.r0 + (int64_t)L.r0 * ld;
        range_finder(B, ld, R.r1 - R.r0, L.r1 - L.r0, x.rank, g, o.power_iterations, x.V, x.U);
        pre(x.left);
        pre(x.right);
      };
      pre(st.first);
    } else {
      TNode& L = s.tree[h.left];
      TNode& R = s.tree[h.right];
      const int rows = R.r1 - R.r0, cols = L.r1 - L.r0;
      double* B = P + R.r0 + (int64_t)L.r0 * ld;
      tree_solve_right(s, h.left, B, rows, ld);
      range_finder(B, ld, rows, cols, h.rank, g, o.power_iterations, h.V, h.U);
      // ancestors' contribution to the whole right block
      lowrank_syrk(P + R.r0 + (int64_t)R.r0 * ld, ld, rows, h.V, h.U, cols, h.rank);
    }
  }
  return 0;
}

void build_tree(Node& s, const std::vector<int>& offs, int thr) {
  s.tree.clear();
  s.steps.clear();
  int nleaf = offs.empty() ? 0 : (int)offs.size() - 1;
  if (nleaf <= 0) {
    TNode t;
    t.r0 = 0;
    t.r1 = s.ncols;
    s.tree.push_back(t);
    s.root = 0;
    s.steps.push_back({0, 1});
    return;
  }
  std::vector<int> cur;
  for (int b = 0; b < nleaf; ++b) {
    TNode t;
    t.r0 = offs[b];
    t.r1 = offs[b + 1];
    s.tree.push_back(t);
    cur.push_back((int)s.tree.size() - 1);
  }
  while (cur.size() > 1) {
    std::vector<int> nx;
    for (size_t k = 0; k + 1 < cur.size(); k += 2) {
      TNode t;
      t.left = cur[k];
      t.right = cur[k + 1];
      t.r0 = s.tree[t.left].r0;
      t.r1 = s.tree[t.right].r1;
      s.tree.push_back(t);
      int id = (int)s.tree.size() - 1;
      s.tree[cur[k]].parent = id;
      s.tree[cur[k + 1]].parent = id;
      nx.push_back(id);
    }
    cur.swap(nx);
  }
  s.root = cur[0];
  bool use_explicit = nleaf > 1 && thr > 0;
  std::function<void(int)> trav = [&](int id) {
    TNode& h = s.tree[id];
    int rows = h.r1 - h.r0;
    bool big = !use_explicit || rows > thr;
    if (h.left >= 0 && big) trav(h.left);
    if (use_explicit && rows <= thr) s.steps.push_back({id, 0});
    else if (h.left < 0) s.steps.push_back({id, 1});
    else s.steps.push_back({id, 2});
    if (h.left >= 0 && big) trav(h.right);
  };
  trav(s.root);
}

void set_ranks(Node& s, const Opts& o) {
  for (auto& t : s.tree)
    if (t.left >= 0)
      t.rank = rank_of(s.tree[t.right].r1 - s.tree[t.right].r0,
                       s.tree[t.left].r1 - s.tree[t.left].r0, o.diag_rank_constant, o);
  s.rank = s.rows.empty() ? 0 : rank_of((int)s.rows.size(), s.ncols, o.rank_constant, o);
}

int numeric(Oracle& O) {
  const int ns = (int)O.nodes.size();
  Gauss g(O.opt.seed);  // reseeded per factorization (FactorManager.cpp:384)
  std::vector<int> head(ns, -1), next(ns, -1), ptr(ns, 0);
  for (int j = 0; j < ns; ++j) {
    Node& s = O.nodes[j];
    const int c = s.ncols, R = (int)s.rows.size(), ld = c + R, last = s.first + c - 1;
    Vec P((size_t)ld * c, 0.0);
    auto pos = [&](int r) -> int {
      if (r <= last) return r - s.first;
      return c + (int)(std::lower_bound(s.rows.begin(), s.rows.end(), r) - s.rows.begin());
    };
    for (int k = s.first; k <= last; ++k) {
      int oc = O.perm[k];
      for (int p = O.Ap[oc]; p < O.Ap[oc + 1]; ++p) {
        int r = O.pinv[O.Ai[p]];
        if (r >= k) P[pos(r) + (size_t)(k - s.first) * ld] = O.Ax[p];
      }
    }
    // descendants (link lists, FactorManager.cpp:463-620)
    for (int d = head[j], dn; d != -1; d = dn) {
      dn = next[d];
      Node& D = O.nodes[d];
      const int Rd = (int)D.rows.size();
      int p0 = ptr[d], p1 = p0;
      while (p1 < Rd && D.rows[p1] <= last) ++p1;
      const int m = Rd - p0, w = p1 - p0;
      // rows p0.. of the descendant's L21 (explicit or V U^T)
      Vec Ld((size_t)m * D.ncols);
      if (!D.compress) {
        const int ldd = D.ncols + Rd;
        for (int t = 0; t < D.ncols; ++t)
          for (int a = 0; a < m; ++a)
            Ld[a + (size_t)t * m] = D.panel[(D.ncols + p0 + a) + (size_t)t * ldd];
      } else {
        Vec Vs((size_t)m * D.rank);
        for (int q = 0; q < D.rank; ++q)
          for (int a = 0; a < m; ++a) Vs[a + (size_t)q * m] = D.V[(p0 + a) + (size_t)q * Rd];
        std::fill(Ld.begin(), Ld.end(), 0.0);
        gemm_nt(m, D.ncols, D.rank, 1.0, Vs.data(), m, D.U.data(), D.ncols, Ld.data(), m);
      }
      Vec Upd((size_t)m * w, 0.0);
      gemm_nt(m, w, D.ncols, 1.0, Ld.data(), m, Ld.data(), m, Upd.data(), m);
      for (int b = 0; b < w; ++b) {
        int cb = pos(D.rows[p0 + b]);
        for (int a = b; a < m; ++a) P[pos(D.rows[p0 + a]) + (size_t)cb * ld] -= Upd[a + (size_t)b * m];
      }
      ptr[d] = p1;
      if (p1 < Rd) {
        int anc = O.col2node[D.rows[p1]];
        next[d] = head[anc];
        head[anc] = d;
      }
    }
    if (!s.compress) {
      if (potrf(c, P.data(), ld) != 0) return 1;
      if (R > 0) trsm_rlt(R, c, P.data(), ld, P.data() + c, ld);
      s.panel.swap(P);
    } else {
      if (factor_compressed_diag(s, P.data(), ld, g, O.opt) != 0) return 1;
      if (R > 0) {
        tree_solve_right(s, s.root, P.data() + c, R, ld);
        range_finder(P.data() + c, ld, R, c, s.rank, g, O.opt.power_iterations, s.V, s.U);
      }
    }
    if (R > 0) {
      int anc = O.col2node[s.rows[0]];
      next[j] = head[anc];
      head[anc] = j;
    }
  }
  return 0;
}

// HODLR lower solve with a vector (DiagonalBlock.cpp:1306-1378)
void tree_solve_vec(const Node& s, int q, double* y, bool transpose) {
  const TNode& h = s.tree[q];
  if (h.left < 0) {
    int b = h.r1 - h.r0;
    const double* L = h.leaf.data();
    if (!transpose) {
      for (int j = 0; j < b; ++j) {
        y[j] /= L[j + (size_t)j * b];
        for (int i = j + 1; i < b; ++i) y[i] -= L[i + (size_t)j * b] * y[j];
      }
    } else {
      for (int j = b - 1; j >= 0; --j) {
        double v = y[j];
        for (int i = j + 1; i < b; ++i) v -= L[i + (size_t)j * b] * y[i];
        y[j] = v / L[j + (size_t)j * b];
      }
    }
    return;
  }
  const TNode& L = s.tree[h.left];
  const TNode& R = s.tree[h.right];
  double* y1 = y + (L.r0 - h.r0);
  double* y2 = y + (R.r0 - h.r0);
  int cl = L.r1 - L.r0, cr = R.r1 - R.r0, r = h.rank;
  Vec t(r);
  if (!transpose) {
    tree_solve_vec(s, h.left, y1, false);
    for (int q2 = 0; q2 < r; ++q2) {
      double a = 0;
      for (int i = 0; i < cl; ++i) a += h.U[i + (size_t)q2 * cl] * y1[i];
      t[q2] = a;
    }
    for (int i = 0; i < cr; ++i) {
      double a = 0;
      for (int q2 = 0; q2 < r; ++q2) a += h.V[i + (size_t)q2 * cr] * t[q2];
      y2[i] -= a;
    }
    tree_solve_vec(s, h.right, y2, false);
  } else {
    tree_solve_vec(s, h.right, y2, true);
    for (int q2 = 0; q2 < r; ++q2) {
      double a = 0;
      for (int i = 0; i < cr; ++i) a += h.V[i + (size_t)q2 * cr] * y2[i];
      t[q2] = a;
    }
    for (int i = 0; i < cl; ++i) {
      double a = 0;
      for (int q2 = 0; q2 < r; ++q2) a += h.U[i + (size_t)q2 * cl] * t[q2];
      y1[i] -= a;
    }
    tree_solve_vec(s, h.left, y1, true);
  }
}

void apply(const Oracle& O, const double* b, double* x) {
  const int n = O.n;
  Vec w(n);
  for (int i = 0; i < n; ++i) w[i] = b[O.perm[i]];
  // forward: L y = w
  for (const Node& s : O.nodes) {
    const int c = s.ncols, R = (int)s.rows.size();
    double* y = w.data() + s.first;
    if (!s.compress) {
      const int ld = c + R;
      for (int j = 0; j < c; ++j) {
        y[j] /= s.panel[j + (size_t)j * ld];
        for (int i = j + 1; i < c; ++i) y[i] -= s.panel[i + (size_t)j * ld] * y[j];
      }
      for (int a = 0; a < R; ++a) {
        double acc = 0;
        for (int j = 0; j < c; ++j) acc += s.panel[(c + a) + (size_t)j * ld] * y[j];
        w[s.rows[a]] -= acc;
      }
    } else {
      tree_solve_vec(s, s.root, y, false);
      if (R > 0) {
        Vec t(s.rank);
        for (int q = 0; q < s.rank; ++q) {
          double a = 0;
          for (int i = 0; i < c; ++i) a += s.U[i + (size_t)q * c] * y[i];
          t[q] = a;
        }
        for (int a = 0; a < R; ++a) {
          double acc = 0;
          for (int q = 0; q < s.rank; ++q) acc += s.V[a + (size_t)q * R] * t[q];
          w[s.rows[a]] -= acc;
        }
      }
    }
  }
  // backward: L^T x = y
  for (int jj = (int)O.nodes.size() - 1; jj >= 0; --jj) {
    const Node& s = O.nodes[jj];
    const int c = s.ncols, R = (int)s.rows.size();
    double* z = w.data() + s.first;
    if (!s.compress) {
      const int ld = c + R;
      for (int j = 0; j < c; ++j) {
        double acc = 0;
        for (int a = 0; a < R; ++a) acc += s.panel[(c + a) + (size_t)j * ld] * w[s.rows[a]];
        z[j] -= acc;
      }
      for (int j = c - 1; j >= 0; --j) {
        double v = z[j];
        for (int i = j + 1; i < c; ++i) v -= s.panel[i + (size_t)j * ld] * z[i];
        z[j] = v / s.panel[j + (size_t)j * ld];
      }
    } else {
      if (R > 0) {
        Vec t(s.rank);
        for (int q = 0; q < s.rank; ++q) {
          double a = 0;
          for (int i = 0; i < R; ++i) a += s.V[i + (size_t)q * R] * w[s.rows[i]];
          t[q] = a;
        }
        for (int i = 0; i < c; ++i) {
          double a = 0;
          for (int q = 0; q < s.rank; ++q) a += s.U[i + (size_t)q * c] * t[q];
          z[i] -= a;
        }
      }
      tree_solve_vec(s, s.root, z, true);
    }
  }
  for (int i = 0; i < n; ++i) x[O.perm[i]] = w[i];
}

void spmv(const Oracle& O, const double* x, double* y) {
  for (int i = 0; i < O.n; ++i) y[i] = 0;
  for (int j = 0; j < O.n; ++j)
    for (int p = O.Ap[j]; p < O.Ap[j + 1]; ++p) y[O.Ai[p]] += O.Ax[p] * x[j];
}

double dot(const Vec& a, const Vec& b) {
  double s = 0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

}  // namespace

extern "C" {

struct oracle_opts {
  double rank_constant, diagonal_rank_constant;
  int oversampling, min_rank, max_rank, power_iterations, explicit_threshold;
  uint32_t seed;
  int max_attempts;
  double retry_alpha;
};

// diag_ptr[nsuper+1] indexes diag_off: node s has offsets
// diag_off[diag_ptr[s] .. diag_ptr[s+1]) (empty or count+1 entries).
// Returns 0 and a handle, or 1 when every attempt of the retry hit an
// indefinite diagonal block.  attempts_out (may be null) receives the number
// of numeric factorizations run and final_constant_out (may be null) the
// diagonal rank constant of the last one, in both cases.
int oracle_factor(int n, const int* colptr, const int* rowind, const double* values,
                  const int* perm, int nsuper, const int* first, const int* compressed,
                  const int* diag_ptr, const int* diag_off, const oracle_opts* o,
                  void** handle, int* attempts_out, double* final_constant_out) {
  std::unique_ptr<Oracle> O(new Oracle());
  O->n = n;
  O->Ap.assign(colptr, colptr + n + 1);
  O->Ai.assign(rowind, rowind + colptr[n]);
  O->Ax.assign(values, values + colptr[n]);
  O->perm.assign(perm, perm + n);
  O->pinv.assign(n, 0);
  for (int k = 0; k < n; ++k) O->pinv[perm[k]] = k;
  O->opt = {o->rank_constant, o->diagonal_rank_constant, o->oversampling, o->min_rank,
            o->max_rank, o->power_iterations, o->explicit_threshold, o->seed};
  O->nodes.resize(nsuper);
  O->col2node.assign(n, 0);
  for (int s = 0; s < nsuper; ++s) {
    Node& x = O->nodes[s];
    x.first = first[s];
    x.ncols = first[s + 1] - first[s];
    x.compress = compressed[s] != 0;
    for (int k = x.first; k < first[s + 1]; ++k) O->col2node[k] = s;
  }
  // row structure: A's rows below the node plus children's rows
  std::vector<std::vector<int>> kids(nsuper);
  std::vector<int> mark(n, -1);
  for (int s = 0; s < nsuper; ++s) {
    Node& x = O->nodes[s];
    int last = x.first + x.ncols - 1;
    for (int k = x.first; k <= last; ++k) {
      int oc = perm[k];
      for (int p = colptr[oc]; p < colptr[oc + 1]; ++p) {
        int r = O->pinv[rowind[p]];
        if (r > last && mark[r] != s) {
          mark[r] = s;
          x.rows.push_back(r);
        }
      }
    }
    for (int ch : kids[s])
      for (int r : O->nodes[ch].rows)
        if (r > last && mark[r] != s) {
          mark[r] = s;
          x.rows.push_back(r);
        }
    std::sort(x.rows.begin(), x.rows.end());
    if (!x.rows.empty()) {
      x.parent = O->col2node[x.rows[0]];
      kids[x.parent].push_back(s);
    }
  }
  for (int s = 0; s < nsuper; ++s) {
    Node& x = O->nodes[s];
    if (!x.compress) continue;
    std::vector<int> offs(diag_off + diag_ptr[s], diag_off + diag_ptr[s + 1]);
    build_tree(x, offs, o->explicit_threshold);
  }
  // indefinite-diagonal retry (SparseSolver.cpp:174-218): the diagonal rank
  // constant grows by retry_alpha per failed attempt, ordering kept
  double dconst = o->diagonal_rank_constant;
  int attempt = 0;
  for (;;) {
    ++attempt;
    O->opt.diag_rank_constant = dconst;
    for (auto& x : O->nodes)
      if (x.compress) set_ranks(x, O->opt);
    if (numeric(*O) == 0) break;
    if (attempt >= o->max_attempts) {
      if (attempts_out) *attempts_out = attempt;
      if (final_constant_out) *final_constant_out = dconst;
      return 1;
    }
    dconst *= o->retry_alpha;
  }
  if (attempts_out) *attempts_out = attempt;
  if (final_constant_out) *final_constant_out = dconst;
  O->attempts = attempt;
  *handle = O.release();
  return 0;
}

int oracle_attempts(void* h) { return ((Oracle*)h)->attempts; }

void oracle_apply(void* h, const double* b, double* x) { apply(*(Oracle*)h, b, x); }

void oracle_spmv(void* h, const double* x, double* y) { spmv(*(Oracle*)h, x, y); }

// PCG_SOLVER::solvePCG (PCG_SOLVER.cpp:226-363), x holds the initial guess.
int oracle_pcg(void* h, const double* b, double* x, double eps, int iters,
               double* recurrence_relres, double* true_relres) {
  const Oracle& O = *(Oracle*)h;
  const int n = O.n;
  Vec r(n), d(n), q(n), s(n), xv(x, x + n), bv(b, b + n);
  spmv(O, xv.data(), r.data());
  for (int i = 0; i < n; ++i) r[i] = bv[i] - r[i];
  apply(O, r.data(), d.data());
  double deltaNew = dot(r, d);
  double maxR = 2.0 * eps;
  const double bn = std::sqrt(dot(bv, bv));
  int i = 0;
  while (i < iters && maxR > eps) {
    spmv(O, d.data(), q.data());
    double alpha = dot(d, q);
    if (std::fabs(alpha) > 0.0) alpha = deltaNew / alpha;
    for (int k = 0; k < n; ++k) xv[k] += alpha * d[k];
    for (int k = 0; k < n; ++k) r[k] -= alpha * q[k];
    apply(O, r.data(), s.data());
    double deltaOld = deltaNew;
    deltaNew = dot(r, s);
    double beta = deltaNew / deltaOld;
    for (int k = 0; k < n; ++k) d[k] = beta * d[k] + s[k];
    maxR = std::sqrt(dot(r, r)) / bn;
    ++i;
  }
  spmv(O, xv.data(), r.data());
  for (int k = 0; k < n; ++k) r[k] = bv[k] - r[k];
  if (recurrence_relres) *recurrence_relres = maxR;
  if (true_relres) *true_relres = std::sqrt(dot(r, r)) / bn;
  std::copy(xv.begin(), xv.end(), x);
  return i;
}

// k independent PCG solves (columns of B and X, leading dimension n), one
// per OpenMP thread at a time: the CPU counterpart of the device's block
// PCG, whose columns follow solvePCG's arithmetic on their own.
void oracle_block_pcg(void* h, int k, const double* B, double* X, double eps, int iters,
                      int* iterations, double* recurrence_relres, double* true_relres) {
  const int64_t n = ((Oracle*)h)->n;
#pragma omp parallel for schedule(dynamic, 1)
  for (int j = 0; j < k; ++j)
    iterations[j] = oracle_pcg(h, B + j * n, X + j * n, eps, iters,
                               recurrence_relres + j, true_relres + j);
}

// Threads of the OpenMP kernels (the CPU baseline times 1 and all cores).
void oracle_set_threads(int nthreads) { omp_set_num_threads(nthreads > 0 ? nthreads : 1); }
int oracle_max_threads(void) { return omp_get_max_threads(); }

// Off-diagonal factor block of supernode s as a dense R x c matrix
// (explicit L21, or V U^T for a compressed supernode); returns R.
int oracle_offdiag(void* h, int s, double* out, int cap) {
  const Oracle& O = *(Oracle*)h;
  const Node& x = O.nodes[s];
  const int c = x.ncols, R = (int)x.rows.size();
  if (!out || (int64_t)R * c > cap) return R;
  if (!x.compress) {
    for (int j = 0; j < c; ++j)
      for (int a = 0; a < R; ++a) out[a + (size_t)j * R] = x.panel[(c + a) + (size_t)j * (c + R)];
  } else {
    std::fill(out, out + (size_t)R * c, 0.0);
    gemm_nt(R, c, x.rank, 1.0, x.V.data(), R, x.U.data(), c, out, R);
  }
  return R;
}

int oracle_rank(void* h, int s) { return ((Oracle*)h)->nodes[s].rank; }

void oracle_destroy(void* h) { delete (Oracle*)h; }

// first `count` draws of the sketch stream (known-answer checks)
void oracle_gaussian(uint32_t seed, int64_t count, double* out) {
  Gauss g(seed);
  for (int64_t k = 0; k < count; ++k) out[k] = g.next();
}

// MT19937 raw output number `index` (1-based) for seed
uint32_t oracle_mt_output(uint32_t seed, int64_t index) {
  std::mt19937 g(seed);
  uint32_t v = 0;
  for (int64_t k = 0; k < index; ++k) v = g();
  return v;
}

}  // extern "C"
