// extern "C" boundary of librsc (declared in include/rsc.h).
//
// rsc_factor_create mirrors FactorPreconditioner::initialize
// (src/rschol/solver/FactorPreconditioner.cpp:70-131): deep copy of A,
// symbolic analysis on the host, then the numeric factor on the device with
// the indefinite-diagonal retry of SparseSolver::factorNumeric
// (SparseSolver.cpp:174-218): diagonal rank constant *= retry_alpha, ranks
// re-planned (ordering kept), at most max_attempts factorizations.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>

#include "capi_internal.h"


struct rsc_symbolic {
  rsc::Csc A;
  rsc::Plan plan;
};

struct rsc_matrix {
  rsc::Csc A;
  std::vector<double> xyz;
};

namespace {

using clk = std::chrono::steady_clock;

rsc_status fail(rsc_status st, const std::string& msg) {
  rsc::last_error() = msg;
  return st;
}

template <typename F>
rsc_status guard(F&& f) {
  try {
    return f();
  } catch (const rsc::IndefiniteError& e) {
    return fail(RSC_INDEFINITE, e.what());
  } catch (const rsc::DeviceError& e) {
    return fail(RSC_DEVICE_ERROR, e.what());
  } catch (const std::bad_alloc& e) {
    return fail(RSC_OOM, std::string("allocation failed: ") + e.what());
  } catch (const std::invalid_argument& e) {
    return fail(RSC_BAD_INPUT, e.what());
  } catch (const std::exception& e) {
    return fail(RSC_BAD_INPUT, e.what());
  }
}

rsc_status check_csc(int n, const int* colptr, const int* rowind) {
  if (n <= 0 || !colptr || !rowind) return fail(RSC_BAD_INPUT, "null matrix or n <= 0");
  if (colptr[0] != 0) return fail(RSC_BAD_INPUT, "colptr[0] != 0");
  for (int j = 0; j < n; ++j) {
    if (colptr[j + 1] < colptr[j]) return fail(RSC_BAD_INPUT, "colptr not monotone");
    for (int p = colptr[j]; p < colptr[j + 1]; ++p)
      if (rowind[p] < 0 || rowind[p] >= n) return fail(RSC_BAD_INPUT, "row index out of range");
  }
  return RSC_OK;
}

rsc_status check_options(const rsc_options& o) {
  if (o.extended_factor_type != RSC_CHOLESKY || o.compression_type != RSC_IN_PLACE ||
      o.decomposition_block != RSC_DECOMPOSE_TRANSPOSE ||
      o.decomposition_type != RSC_DECOMPOSE_FACTOR)
    return fail(RSC_UNSUPPORTED_MODE,
                "only CHOLESKY / IN_PLACE / DECOMPOSE_TRANSPOSE / DECOMPOSE_FACTOR is built");
  if (o.factor_precision != RSC_FP64 && o.factor_precision != RSC_FP32)
    return fail(RSC_UNSUPPORTED_MODE, "factor_precision must be RSC_FP64 or RSC_FP32");
  // sketch products run in the factor's precision
  if (o.sketch_precision != RSC_FP64 && o.sketch_precision != RSC_FP32 &&
      o.sketch_precision != RSC_BF16)
    return fail(RSC_UNSUPPORTED_MODE, "bad sketch_precision");
  if (o.sketch_precision != RSC_FP64 && o.factor_precision != RSC_FP32)
    return fail(RSC_UNSUPPORTED_MODE, "FP32 and BF16 sketches need an FP32 factor");
  if (o.max_dense_block < 0 || o.max_diagonal_block < 1 || o.power_iterations < 0 ||
      o.min_rank < 1 || o.max_attempts < 1)
    return fail(RSC_BAD_INPUT, "bad option value");
  return RSC_OK;
}

rsc::Csc copy_csc(int n, const int* colptr, const int* rowind, const double* values) {
  rsc::Csc A;
  A.n = n;
  A.p.assign(colptr, colptr + n + 1);
  A.i.assign(rowind, rowind + colptr[n]);
  if (values) A.x.assign(values, values + colptr[n]);
  return A;
}

rsc_status check_ranks(const rsc::Plan& plan) {
  for (auto& s : plan.nodes) {
    if (s.rank > 1024) return fail(RSC_UNSUPPORTED_MODE, "rank above 1024");
    for (auto& h : s.hnodes)
      if (h.rank > 1024) return fail(RSC_UNSUPPORTED_MODE, "rank above 1024");
  }
  return RSC_OK;
}

// Numeric factorization with the indefinite-diagonal retry of
// SparseSolver::factorNumeric (SparseSolver.cpp:174-218).  Every call starts
// from the constant the plan currently holds (f->diag_const): a constant
// raised by an earlier retry is kept for later refactorizations, as the
// reference keeps _diagonalRankConstant in its SparseSolver.
rsc_status factor_numeric(rsc_factor* f, const double* values) {
  auto t0 = clk::now();
  int attempt = 0;
  for (;;) {
    ++attempt;
    try {
      f->eng->numeric(values);
      break;
    } catch (const rsc::IndefiniteError& e) {
      if (attempt >= f->plan.opt.max_attempts) {
        f->info.attempts = attempt;
        f->info.final_diagonal_rank_constant = f->diag_const;
        return fail(RSC_INDEFINITE, e.what());
      }
      f->diag_const *= f->plan.opt.retry_alpha;
      if (f->plan.opt.verbose)
        std::printf("Increasing rank constant to %f\n", f->diag_const);
      f->eng.reset();
      rsc::replan_ranks(f->plan, f->diag_const);
      rsc_status st = check_ranks(f->plan);
      if (st != RSC_OK) return st;
      f->eng = rsc::make_engine(f->plan, f->A, f->stream);
    }
  }
  f->info.attempts = attempt;
  f->info.final_diagonal_rank_constant = f->diag_const;
  f->info.numeric_seconds = std::chrono::duration<double>(clk::now() - t0).count();
  f->info.factor_bytes = f->eng->factor_bytes();
  f->info.workspace_bytes = f->eng->workspace_bytes();
  f->info.pivot_replacements = f->eng->pivot_replacements();
  f->info.factor_flops = f->plan.factor_flops;
  const double es = f->plan.opt.factor_precision == RSC_FP32 ? 4.0 : 8.0;
  f->info.apply_bytes = f->plan.apply_bytes_per_elem * es + 24.0 * f->plan.n +
                        32.0 * f->plan.n;
  f->info.spmv_bytes = 12.0 * f->A.nnz() + 4.0 * (f->A.n + 1) + 16.0 * f->A.n;
  return RSC_OK;
}

rsc_status create_common(int n, const int* colptr, const int* rowind,
                         const double* values, const double* xyz,
                         const rsc::NDResult* nd, const rsc_options* opt,
                         rsc_factor** out) {
  if (!out) return fail(RSC_BAD_INPUT, "out is null");
  *out = nullptr;
  rsc_status st = check_csc(n, colptr, rowind);
  if (st != RSC_OK) return st;
  if (!values) return fail(RSC_BAD_INPUT, "values is null");
  rsc_options o;
  if (opt) o = *opt;
  else rsc_options_default(&o);
  st = check_options(o);
  if (st != RSC_OK) return st;
  return guard([&]() -> rsc_status {
    std::unique_ptr<rsc_factor> f(new rsc_factor());
    f->A = copy_csc(n, colptr, rowind, values);
    if (xyz) {
      f->xyz.assign(xyz, xyz + 3 * (size_t)n);
      f->has_xyz = true;
    }
    auto t0 = clk::now();
    rsc::build_plan(f->A, f->has_xyz ? f->xyz.data() : nullptr, o, nd, f->plan);
    f->diag_const = o.diagonal_rank_constant;
    f->info.symbolic_seconds = std::chrono::duration<double>(clk::now() - t0).count();
    rsc_status s2 = check_ranks(f->plan);
    if (s2 != RSC_OK) return s2;
    RSC_HIP(hipStreamCreate(&f->stream));
    f->dA.upload(f->A, f->stream);
    f->eng = rsc::make_engine(f->plan, f->A, f->stream);
    f->info.n = n;
    f->info.nsuper = (int)f->plan.nodes.size();
    int nc = 0;
    for (auto& s : f->plan.nodes) nc += s.compress ? 1 : 0;
    f->info.ncompressed = nc;
    f->info.nlevels = (int)f->plan.levels.size();
    s2 = factor_numeric(f.get(), f->A.x.data());
    if (s2 != RSC_OK) {
      rsc_factor* raw = f.release();
      rsc_factor_destroy(raw);
      return s2;
    }
    *out = f.release();
    return RSC_OK;
  });
}

}  // namespace

extern "C" {

rsc_status rsc_options_default(rsc_options* o) {
  if (!o) return fail(RSC_BAD_INPUT, "null options");
  std::memset(o, 0, sizeof(*o));
  o->max_dense_block = 200;
  o->max_diagonal_block = 400;
  o->extended_factor_type = RSC_CHOLESKY;
  o->compression_type = RSC_IN_PLACE;
  o->decomposition_block = RSC_DECOMPOSE_TRANSPOSE;
  o->decomposition_type = RSC_DECOMPOSE_FACTOR;
  o->reorder_separators = 1;
  o->explicit_block_threshold = 4000;
  o->rank_constant = 0.4;
  o->diagonal_rank_constant = 0.5;
  o->oversampling = 5;
  o->min_rank = 16;
  o->max_rank = 0;
  o->power_iterations = 1;
  o->seed = 152;
  o->max_attempts = 10;
  o->retry_alpha = 1.25;
  o->nd_small = 200;
  o->factor_precision = RSC_FP64;
  o->sketch_precision = RSC_FP64;
  o->verbose = 0;
  return RSC_OK;
}

rsc_status rsc_factor_create(int n, const int* colptr, const int* rowind,
                             const double* values, const double* xyz,
                             const rsc_options* opt, rsc_factor** out) {
  return create_common(n, colptr, rowind, values, xyz, nullptr, opt, out);
}

rsc_status rsc_factor_create_with_ordering(int n, const int* colptr,
                                           const int* rowind, const double* values,
                                           const double* xyz, const int* perm,
                                           int ncomponents, const int* cparent,
                                           const int* cmember, const rsc_options* opt,
                                           rsc_factor** out) {
  if (!perm || !cparent || !cmember || ncomponents <= 0)
    return fail(RSC_BAD_INPUT, "null ordering");
  rsc::NDResult nd;
  nd.perm.assign(perm, perm + n);
  nd.cparent.assign(cparent, cparent + ncomponents);
  nd.cmember.assign(cmember, cmember + n);
  return create_common(n, colptr, rowind, values, xyz, &nd, opt, out);
}

rsc_status rsc_factor_refactor(rsc_factor* f, const double* values) {
  if (!f || !values) return fail(RSC_BAD_INPUT, "null argument");
  return guard([&]() -> rsc_status {
    std::copy(values, values + f->A.nnz(), f->A.x.begin());
    f->dA.update_values(values, f->stream);
    return factor_numeric(f, f->A.x.data());
  });
}

rsc_status rsc_apply(rsc_factor* f, int nrhs, const double* B, int ldb, double* X,
                     int ldx, void* stream) {
  if (!f || !B || !X || nrhs < 0 || ldb < f->plan.n || ldx < f->plan.n)
    return fail(RSC_BAD_INPUT, "bad apply arguments");
  return guard([&]() -> rsc_status {
    hipStream_t s = stream ? (hipStream_t)stream : f->stream;
    if (nrhs > 0) f->eng->apply(nrhs, B, ldb, X, ldx, s);
    RSC_HIP(hipGetLastError());
    return RSC_OK;
  });
}

rsc_status rsc_apply_status(rsc_factor* f, void* stream) {
  if (!f) return fail(RSC_BAD_INPUT, "null factor");
  return guard([&]() -> rsc_status {
    hipStream_t s = stream ? (hipStream_t)stream : f->stream;
    f->eng->check_apply(s);
    return RSC_OK;
  });
}

rsc_status rsc_spmv(rsc_factor* f, const double* x, double* y, void* stream) {
  if (!f || !x || !y) return fail(RSC_BAD_INPUT, "null argument");
  return guard([&]() -> rsc_status {
    hipStream_t s = stream ? (hipStream_t)stream : f->stream;
    f->dA.spmv(x, f->plan.n, y, f->plan.n, 1, nullptr, s);
    RSC_HIP(hipGetLastError());
    return RSC_OK;
  });
}

rsc_status rsc_block_pcg_solve(rsc_factor* f, int k, const double* B, int ldb,
                               double* X, int ldx, double eps, int max_iters,
                               rsc_pcg_result* res, void* stream) {
  if (!f || !B || !X || k <= 0 || ldb < f->plan.n || ldx < f->plan.n || max_iters < 0)
    return fail(RSC_BAD_INPUT, "bad PCG arguments");
  return guard([&]() -> rsc_status {
    hipStream_t s = stream ? (hipStream_t)stream : f->stream;
    rsc::pcg_solve_device(*f->eng, f->dA, f->pcgw, k, B, ldb, X, ldx, eps, max_iters, res, s);
    if (res)
      for (int c = 0; c < k; ++c)
        if (!(res[c].recurrence_relres <= eps)) return RSC_NOT_CONVERGED;
    return RSC_OK;
  });
}

rsc_status rsc_pcg_solve(rsc_factor* f, const double* b, double* x, double eps,
                         int max_iters, rsc_pcg_result* res, void* stream) {
  if (!f) return fail(RSC_BAD_INPUT, "null factor");
  return rsc_block_pcg_solve(f, 1, b, f->plan.n, x, f->plan.n, eps, max_iters, res,
                             stream);
}

rsc_status rsc_apply_host(rsc_factor* f, int nrhs, const double* B, int ldb,
                          double* X, int ldx) {
  if (!f || !B || !X || nrhs < 0 || ldb < f->plan.n || ldx < f->plan.n)
    return fail(RSC_BAD_INPUT, "bad apply arguments");
  return guard([&]() -> rsc_status {
    const int n = f->plan.n;
    rsc::DBuf dB, dX;
    dB.alloc((size_t)n * nrhs * sizeof(double));
    dX.alloc((size_t)n * nrhs * sizeof(double));
    for (int j = 0; j < nrhs; ++j)
      RSC_HIP(hipMemcpyAsync(dB.as<double>() + (size_t)j * n, B + (size_t)j * ldb,
                             n * sizeof(double), hipMemcpyHostToDevice, f->stream));
    if (nrhs > 0) f->eng->apply(nrhs, dB.as<double>(), n, dX.as<double>(), n, f->stream);
    for (int j = 0; j < nrhs; ++j)
      RSC_HIP(hipMemcpyAsync(X + (size_t)j * ldx, dX.as<double>() + (size_t)j * n,
                             n * sizeof(double), hipMemcpyDeviceToHost, f->stream));
    RSC_HIP(hipStreamSynchronize(f->stream));
    f->eng->check_apply(f->stream);
    return RSC_OK;
  });
}

rsc_status rsc_block_pcg_solve_host(rsc_factor* f, int k, const double* B, int ldb,
                                    double* X, int ldx, double eps, int max_iters,
                                    rsc_pcg_result* res) {
  if (!f || !B || !X || k <= 0 || ldb < f->plan.n || ldx < f->plan.n)
    return fail(RSC_BAD_INPUT, "bad PCG arguments");
  rsc_status out = RSC_OK;
  rsc_status st = guard([&]() -> rsc_status {
    const int n = f->plan.n;
    rsc::DBuf dB, dX;
    dB.alloc((size_t)n * k * sizeof(double));
    dX.alloc((size_t)n * k * sizeof(double));
    for (int j = 0; j < k; ++j) {
      RSC_HIP(hipMemcpyAsync(dB.as<double>() + (size_t)j * n, B + (size_t)j * ldb,
                             n * sizeof(double), hipMemcpyHostToDevice, f->stream));
      RSC_HIP(hipMemcpyAsync(dX.as<double>() + (size_t)j * n, X + (size_t)j * ldx,
                             n * sizeof(double), hipMemcpyHostToDevice, f->stream));
    }
    out = rsc_block_pcg_solve(f, k, dB.as<double>(), n, dX.as<double>(), n, eps,
                              max_iters, res, f->stream);
    if (out != RSC_OK && out != RSC_NOT_CONVERGED) return out;
    for (int j = 0; j < k; ++j)
      RSC_HIP(hipMemcpyAsync(X + (size_t)j * ldx, dX.as<double>() + (size_t)j * n,
                             n * sizeof(double), hipMemcpyDeviceToHost, f->stream));
    RSC_HIP(hipStreamSynchronize(f->stream));
    return RSC_OK;
  });
  return st != RSC_OK ? st : out;
}

rsc_status rsc_pcg_solve_host(rsc_factor* f, const double* b, double* x, double eps,
                              int max_iters, rsc_pcg_result* res) {
  if (!f) return fail(RSC_BAD_INPUT, "null factor");
  return rsc_block_pcg_solve_host(f, 1, b, f->plan.n, x, f->plan.n, eps, max_iters, res);
}

rsc_status rsc_factor_get_info(const rsc_factor* f, rsc_factor_info* info) {
  if (!f || !info) return fail(RSC_BAD_INPUT, "null argument");
  *info = f->info;
  return RSC_OK;
}

rsc_status rsc_factor_release_workspace(rsc_factor* f) {
  if (!f) return fail(RSC_BAD_INPUT, "null factor");
  return guard([&]() -> rsc_status {
    f->eng->release_workspace();
    f->info.workspace_bytes = f->eng->workspace_bytes();
    return RSC_OK;
  });
}

rsc_status rsc_factor_get_permutation(const rsc_factor* f, int* perm) {
  if (!f || !perm) return fail(RSC_BAD_INPUT, "null argument");
  std::copy(f->plan.perm.begin(), f->plan.perm.end(), perm);
  return RSC_OK;
}

rsc_status rsc_factor_get_supernodes(const rsc_factor* f, int* first, int* compressed) {
  if (!f) return fail(RSC_BAD_INPUT, "null argument");
  const auto& nodes = f->plan.nodes;
  for (size_t s = 0; s < nodes.size(); ++s) {
    if (first) first[s] = nodes[s].first;
    if (compressed) compressed[s] = nodes[s].compress ? 1 : 0;
  }
  if (first) first[nodes.size()] = f->plan.n;
  return RSC_OK;
}

rsc_status rsc_factor_destroy(rsc_factor* f) {
  if (!f) return RSC_OK;
  if (f->stream) (void)hipStreamSynchronize(f->stream);
  f->eng.reset();
  f->dA.p.reset();
  f->dA.i.reset();
  f->dA.x.reset();
  if (f->stream) (void)hipStreamDestroy(f->stream);
  delete f;
  return RSC_OK;
}

const char* rsc_last_error(void) { return rsc::last_error().c_str(); }

// ---- symbolic ------------------------------------------------------------

static rsc_status symbolic_common(int n, const int* colptr, const int* rowind,
                                  const double* xyz, const rsc::NDResult* nd,
                                  const rsc_options* opt, rsc_symbolic** out) {
  if (!out) return fail(RSC_BAD_INPUT, "out is null");
  *out = nullptr;
  rsc_status st = check_csc(n, colptr, rowind);
  if (st != RSC_OK) return st;
  rsc_options o;
  if (opt) o = *opt;
  else rsc_options_default(&o);
  st = check_options(o);
  if (st != RSC_OK) return st;
  return guard([&]() -> rsc_status {
    std::unique_ptr<rsc_symbolic> s(new rsc_symbolic());
    s->A = copy_csc(n, colptr, rowind, nullptr);
    rsc::build_plan(s->A, xyz, o, nd, s->plan);
    *out = s.release();
    return RSC_OK;
  });
}

rsc_status rsc_symbolic_create(int n, const int* colptr, const int* rowind,
                               const double* xyz, const rsc_options* opt,
                               rsc_symbolic** out) {
  return symbolic_common(n, colptr, rowind, xyz, nullptr, opt, out);
}

rsc_status rsc_symbolic_create_with_ordering(int n, const int* colptr,
                                             const int* rowind, const double* xyz,
                                             const int* perm, int ncomponents,
                                             const int* cparent, const int* cmember,
                                             const rsc_options* opt,
                                             rsc_symbolic** out) {
  if (!perm || !cparent || !cmember || ncomponents <= 0)
    return fail(RSC_BAD_INPUT, "null ordering");
  rsc::NDResult nd;
  nd.perm.assign(perm, perm + n);
  nd.cparent.assign(cparent, cparent + ncomponents);
  nd.cmember.assign(cmember, cmember + n);
  return symbolic_common(n, colptr, rowind, xyz, &nd, opt, out);
}

rsc_status rsc_symbolic_sizes(const rsc_symbolic* s, int* nsuper, int64_t* total_rows,
                              int* nlevels, int64_t* factor_elems, int64_t* rng_count) {
  if (!s) return fail(RSC_BAD_INPUT, "null symbolic");
  int64_t tr = 0;
  for (auto& x : s->plan.nodes) tr += x.R();
  if (nsuper) *nsuper = (int)s->plan.nodes.size();
  if (total_rows) *total_rows = tr;
  if (nlevels) *nlevels = (int)s->plan.levels.size();
  if (factor_elems) *factor_elems = s->plan.factor_elems;
  if (rng_count) *rng_count = s->plan.rng_count;
  return RSC_OK;
}

rsc_status rsc_symbolic_get(const rsc_symbolic* s, int* perm, int* first,
                            int* compressed, int* parent, int64_t* rows_ptr, int* rows,
                            int* rank, int* level) {
  if (!s) return fail(RSC_BAD_INPUT, "null symbolic");
  const auto& P = s->plan;
  if (perm) std::copy(P.perm.begin(), P.perm.end(), perm);
  int64_t off = 0;
  for (size_t j = 0; j < P.nodes.size(); ++j) {
    const auto& x = P.nodes[j];
    if (first) first[j] = x.first;
    if (compressed) compressed[j] = x.compress ? 1 : 0;
    if (parent) parent[j] = x.parent;
    if (rows_ptr) rows_ptr[j] = off;
    if (rows) std::copy(x.rows.begin(), x.rows.end(), rows + off);
    if (rank) rank[j] = x.rank;
    if (level) level[j] = x.level;
    off += x.R();
  }
  if (first) first[P.nodes.size()] = P.n;
  if (rows_ptr) rows_ptr[P.nodes.size()] = off;
  return RSC_OK;
}

rsc_status rsc_symbolic_diag_blocks(const rsc_symbolic* s, int node, int* count,
                                    int* offsets) {
  if (!s || !count || node < 0 || node >= (int)s->plan.diag_blocks.size())
    return fail(RSC_BAD_INPUT, "bad node");
  const auto& b = s->plan.diag_blocks[node];
  *count = b.empty() ? 0 : (int)b.size() - 1;
  if (offsets) std::copy(b.begin(), b.end(), offsets);
  return RSC_OK;
}

void rsc_symbolic_destroy(rsc_symbolic* s) { delete s; }

rsc_status rsc_nested_dissection(int n, const int* colptr, const int* rowind,
                                 const double* xyz, int nd_small, int* perm,
                                 int* ncomponents, int* cparent, int* cmember) {
  rsc_status st = check_csc(n, colptr, rowind);
  if (st != RSC_OK) return st;
  if (!perm || !ncomponents || !cparent || !cmember) return fail(RSC_BAD_INPUT, "null output");
  return guard([&]() -> rsc_status {
    rsc::Csc A = copy_csc(n, colptr, rowind, nullptr);
    rsc::NDResult r = rsc::nested_dissection(A, xyz, nd_small);
    std::copy(r.perm.begin(), r.perm.end(), perm);
    std::copy(r.cparent.begin(), r.cparent.end(), cparent);
    std::copy(r.cmember.begin(), r.cmember.end(), cmember);
    *ncomponents = (int)r.cparent.size();
    return RSC_OK;
  });
}

rsc_status rsc_etree_colcounts(int n, const int* colptr, const int* rowind,
                               const int* perm, int* parent, int* colcount) {
  rsc_status st = check_csc(n, colptr, rowind);
  if (st != RSC_OK) return st;
  if (!perm || !parent || !colcount) return fail(RSC_BAD_INPUT, "null argument");
  return guard([&]() -> rsc_status {
    rsc::Csc A = copy_csc(n, colptr, rowind, nullptr);
    std::vector<int> pinv(n);
    for (int k = 0; k < n; ++k) pinv[perm[k]] = k;
    std::vector<int> par = rsc::elimination_tree(A, pinv);
    std::vector<int> post(n);
    for (int k = 0; k < n; ++k) post[k] = k;
    std::vector<int> cc = rsc::column_counts(A, pinv, par, post);
    std::copy(par.begin(), par.end(), parent);
    std::copy(cc.begin(), cc.end(), colcount);
    return RSC_OK;
  });
}

rsc_status rsc_relaxed_supernodes(int n, const int* parent, const int* colcount, int ncomp,
                                  const int* comp_size, const int* comp_first,
                                  const int* comp_last, const int* comp_nchildren,
                                  const int* comp_parent, int max_dense_block, int* first,
                                  int* compressed, int* nsuper) {
  if (n < 0 || ncomp < 0 || !parent || !colcount || !comp_size || !comp_first ||
      !comp_last || !comp_nchildren || !comp_parent || !first || !compressed || !nsuper)
    return fail(RSC_BAD_INPUT, "bad argument");
  for (int j = 0; j < n; ++j)
    if (parent[j] < -1 || parent[j] >= n || parent[j] == j)
      return fail(RSC_BAD_INPUT, "parent out of range");
  int next = 0;
  for (int c = 0; c < ncomp; ++c) {
    if (comp_first[c] != next || comp_last[c] - comp_first[c] + 1 != comp_size[c] ||
        comp_parent[c] < -1 || comp_parent[c] >= ncomp)
      return fail(RSC_BAD_INPUT, "components do not tile the columns");
    next = comp_last[c] + 1;
  }
  if (next != n) return fail(RSC_BAD_INPUT, "components do not cover the columns");
  return guard([&]() -> rsc_status {
    std::vector<int> par(parent, parent + n), cc(colcount, colcount + n);
    std::vector<rsc::Separator> seps(ncomp);
    for (int c = 0; c < ncomp; ++c) {
      seps[c].size = comp_size[c];
      seps[c].first = comp_first[c];
      seps[c].last = comp_last[c];
      seps[c].nchildren = comp_nchildren[c];
    }
    std::vector<int> sp(comp_parent, comp_parent + ncomp);
    std::vector<rsc::SupernodeSpec> out =
        rsc::build_relaxed_supernodes(par, cc, seps, sp, max_dense_block);
    for (size_t q = 0; q < out.size(); ++q) {
      first[q] = out[q].first;
      compressed[q] = out[q].compress ? 1 : 0;
    }
    first[out.size()] = n;
    *nsuper = (int)out.size();
    return RSC_OK;
  });
}

rsc_status rsc_separator_tree(int npts, const double* xyz, int max_leaf, int* order,
                              int* nleaves, int* leaf_sizes) {
  if (npts < 0 || !xyz || !order || !nleaves || max_leaf < 0)
    return fail(RSC_BAD_INPUT, "bad argument");
  return guard([&]() -> rsc_status {
    std::vector<double> pts(xyz, xyz + 3 * (size_t)npts);
    std::vector<int> o, l;
    rsc::separator_tree_order(pts, max_leaf, o, l);
    std::copy(o.begin(), o.end(), order);
    *nleaves = (int)l.size();
    if (leaf_sizes) std::copy(l.begin(), l.end(), leaf_sizes);
    return RSC_OK;
  });
}

rsc_status rsc_gaussian_stream(uint32_t seed, int64_t count, double* out) {
  if (count < 0 || (count > 0 && !out)) return fail(RSC_BAD_INPUT, "bad argument");
  rsc::gaussian_stream(seed, count, out);
  return RSC_OK;
}

int rsc_block_rank(int nrows, int ncols, double constant, int oversampling,
                   int min_rank, int max_rank) {
  return rsc::block_rank(nrows, ncols, constant, oversampling, min_rank, max_rank);
}

// ---- synthetic systems ------------------------------------------------------

static rsc_status make_matrix(rsc::Csc&& A, std::vector<double>&& xyz, rsc_matrix** out) {
  rsc_matrix* m = new rsc_matrix();
  m->A = std::move(A);
  m->xyz = std::move(xyz);
  *out = m;
  return RSC_OK;
}

rsc_status rsc_matrix_poisson2d(int N, rsc_matrix** out) {
  if (!out || N < 1) return fail(RSC_BAD_INPUT, "bad argument");
  return guard([&]() {
    return make_matrix(rsc::poisson2d(N), rsc::grid_positions(N, N, 1), out);
  });
}

rsc_status rsc_matrix_poisson3d(int N, double sigma, uint64_t seed, rsc_matrix** out) {
  if (!out || N < 1) return fail(RSC_BAD_INPUT, "bad argument");
  return guard([&]() {
    return make_matrix(rsc::poisson3d_lognormal(N, sigma, seed),
                       rsc::grid_positions(N, N, N), out);
  });
}

rsc_status rsc_matrix_elasticity3d(int NE, double sigma, double nu, uint64_t seed,
                                   rsc_matrix** out) {
  if (!out || NE < 1) return fail(RSC_BAD_INPUT, "bad argument");
  return guard([&]() {
    rsc::Csc A = rsc::elasticity3d(NE, sigma, nu, seed);
    const int NN = NE + 1;
    std::vector<double> xyz;
    xyz.reserve(3 * (size_t)A.n);
    for (int z = 0; z < NN; ++z)
      for (int y = 0; y < NN; ++y)
        for (int x = 1; x < NN; ++x)
          for (int d = 0; d < 3; ++d) {
            xyz.push_back(x);
            xyz.push_back(y);
            xyz.push_back(z);
          }
    return make_matrix(std::move(A), std::move(xyz), out);
  });
}

rsc_status rsc_matrix_info(const rsc_matrix* m, int* n, int64_t* nnz) {
  if (!m) return fail(RSC_BAD_INPUT, "null matrix");
  if (n) *n = m->A.n;
  if (nnz) *nnz = m->A.nnz();
  return RSC_OK;
}

rsc_status rsc_matrix_export(const rsc_matrix* m, int* colptr, int* rowind,
                             double* values, double* xyz) {
  if (!m) return fail(RSC_BAD_INPUT, "null matrix");
  if (colptr) std::copy(m->A.p.begin(), m->A.p.end(), colptr);
  if (rowind) std::copy(m->A.i.begin(), m->A.i.end(), rowind);
  if (values) std::copy(m->A.x.begin(), m->A.x.end(), values);
  if (xyz) std::copy(m->xyz.begin(), m->xyz.end(), xyz);
  return RSC_OK;
}

void rsc_matrix_destroy(rsc_matrix* m) { delete m; }

rsc_status rsc_uniform_rhs(int64_t count, uint64_t seed, double* out) {
  if (count < 0 || (count > 0 && !out)) return fail(RSC_BAD_INPUT, "bad argument");
  std::vector<double> v = rsc::uniform_rhs(count, seed);
  std::copy(v.begin(), v.end(), out);
  return RSC_OK;
}

}  // extern "C"

template <typename T>
static rsc_status selftest_gemm_t(int ta, int tb, int m, int n, int k, double alpha,
                                  const double* A, int lda, const double* B, int ldb,
                                  double beta, double* Cm, int ldc, bool bf16 = false) {
  const int64_t asz = (int64_t)lda * (ta ? m : k), bsz = (int64_t)ldb * (tb ? k : n),
                csz = (int64_t)ldc * n;
  std::vector<T> ha(A, A + asz), hb(B, B + bsz), hc(Cm, Cm + csz);
  rsc::DBuf da, db, dc, dj, dt;
  hipStream_t s = nullptr;
  da.upload(ha, s);
  db.upload(hb, s);
  dc.upload(hc, s);
  rsc::dev::GemmJob J{da.p, db.p, dc.p, m, n, k, lda, ldb, ldc, alpha, beta};
  std::vector<rsc::dev::GemmJob> jobs{J};
  std::vector<rsc::dev::GemmTile> tiles;
  int ncnt = 0;
  const int parts = rsc::dev::plan_gemm_tiles(jobs.data(), 1, false, tiles, &ncnt);
  rsc::DBuf dpart, dcnt;
  dpart.alloc((size_t)std::max(parts, 1) * rsc::dev::TILE * rsc::dev::TILE * sizeof(T));
  dcnt.alloc((size_t)std::max(ncnt, 1) * sizeof(unsigned));
  RSC_HIP(hipMemset(dcnt.p, 0, dcnt.bytes));
  dj.upload(jobs, s);
  dt.upload(tiles, s);
  if (bf16)
    rsc::dev::launch_gemm_bf16(ta != 0, tb != 0, dj.as<rsc::dev::GemmJob>(),
                               dt.as<rsc::dev::GemmTile>(), (int)tiles.size(), dpart.p,
                               dcnt.as<unsigned>(), s);
  else
    rsc::dev::launch_gemm<T>(ta != 0, tb != 0, dj.as<rsc::dev::GemmJob>(),
                             dt.as<rsc::dev::GemmTile>(), (int)tiles.size(), dpart.p,
                             dcnt.as<unsigned>(), s);
  RSC_HIP(hipGetLastError());
  RSC_HIP(hipMemcpy(hc.data(), dc.p, csz * sizeof(T), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < csz; ++i) Cm[i] = (double)hc[i];
  return RSC_OK;
}

extern "C" {

rsc_status rsc_selftest_gemm(int precision, int transa, int transb, int m, int n, int k,
                             double alpha, const double* A, int lda, const double* B,
                             int ldb, double beta, double* C, int ldc) {
  if (!A || !B || !C || m < 0 || n < 0 || k < 0 || ldc < m ||
      lda < (transa ? k : m) || ldb < (transb ? n : k))
    return fail(RSC_BAD_INPUT, "bad gemm arguments");
  return guard([&]() -> rsc_status {
    if (precision == RSC_FP32 || precision == RSC_BF16)
      return selftest_gemm_t<float>(transa, transb, m, n, k, alpha, A, lda, B, ldb, beta,
                                    C, ldc, precision == RSC_BF16);
    return selftest_gemm_t<double>(transa, transb, m, n, k, alpha, A, lda, B, ldb, beta,
                                   C, ldc);
  });
}

// Exact reference entry point (pcg_pysolver.cpp:83-101): no positions,
// maxDenseBlock 200, maxDiagonalBlock 400, PCG to 1e-6, 10000 iterations,
// starting from zero.
void pcg_pysolve(int nrow, const int* jc, const int* ir, const double* pr,
                 const double* rhs, double* solution) {
  rsc_options o;
  rsc_options_default(&o);
  o.reorder_separators = 0;
  rsc_factor* f = nullptr;
  rsc_status st = rsc_factor_create(nrow, jc, ir, pr, nullptr, &o, &f);
  if (st != RSC_OK) {
    std::fprintf(stderr, "pcg_pysolve: factorization failed: %s\n", rsc_last_error());
    for (int i = 0; i < nrow; ++i) solution[i] = 0.0;
    return;
  }
  for (int i = 0; i < nrow; ++i) solution[i] = 0.0;
  rsc_pcg_result res{};
  st = rsc_pcg_solve_host(f, rhs, solution, 1e-6, 10000, &res);
  if (st != RSC_OK && st != RSC_NOT_CONVERGED)
    std::fprintf(stderr, "pcg_pysolve: PCG failed: %s\n", rsc_last_error());
  rsc_factor_destroy(f);
}

}  // extern "C"
