// Exercises the C++ mirror (include/rschol_amd.hpp) the way a reference
// caller does (src/pcg_solver.cpp:70-120): build A, FactorPreconditioner::
// initialize, PCG_SOLVER::solvePCG, then refactor after clearFactor().
//
//   test_cpp_api cpu   error paths only (no device work)
//   test_cpp_api gpu   factor + PCG on a 16^3 log-normal Poisson system
//
// Prints one "key=value ..." line per check; exit status 0 on success.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "rschol_amd.hpp"

using namespace rschol_amd;

static SparseColumnMatrix poisson3d(int N, double sigma, uint64_t seed, Vector3Array& pos) {
  rsc_matrix* m = nullptr;
  check(rsc_matrix_poisson3d(N, sigma, seed, &m), "rsc_matrix_poisson3d");
  int n = 0;
  int64_t nnz = 0;
  rsc_matrix_info(m, &n, &nnz);
  SparseColumnMatrix A;
  A._nrow = A._ncol = n;
  A._p.resize(n + 1);
  A._i.resize(nnz);
  A._x.resize(nnz);
  std::vector<double> xyz(3 * (size_t)n);
  check(rsc_matrix_export(m, A._p.data(), A._i.data(), A._x.data(), xyz.data()),
        "rsc_matrix_export");
  rsc_matrix_destroy(m);
  pos.resize(n);
  for (int k = 0; k < n; ++k) pos[k] = {xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2]};
  return A;
}

static double true_relres(const SparseColumnMatrix& A, const VECTOR& x, const VECTOR& b) {
  VECTOR r(b);
  for (int j = 0; j < A._ncol; ++j)
    for (int k = A._p[j]; k < A._p[j + 1]; ++k) r[A._i[k]] -= A._x[k] * x[j];
  double rn = 0, bn = 0;
  for (size_t i = 0; i < b.size(); ++i) {
    rn += r[i] * r[i];
    bn += b[i] * b[i];
  }
  return std::sqrt(rn / bn);
}

static int cpu_checks() {
  Vector3Array pos;
  SparseColumnMatrix A = poisson3d(6, 1.0, 1234, pos);
  FactorPreconditioner M;
  // unsupported enum combination: an error, not an abort
  try {
    M.initialize(A, pos, 63, 400, LDL, IN_PLACE, DECOMPOSE_TRANSPOSE, DECOMPOSE_FACTOR, true);
    std::printf("unsupported_mode=not_rejected\n");
    return 1;
  } catch (const Error& e) {
    if (e.status != RSC_UNSUPPORTED_MODE) return 1;
    std::printf("unsupported_mode=rejected\n");
  }
  // malformed matrix
  SparseColumnMatrix bad = A;
  bad._i[2] = 1 << 28;
  try {
    M.initialize(bad, pos, 63, 400, CHOLESKY, IN_PLACE, DECOMPOSE_TRANSPOSE, DECOMPOSE_FACTOR,
                 true);
    std::printf("bad_input=not_rejected\n");
    return 1;
  } catch (const Error& e) {
    if (e.status != RSC_BAD_INPUT) return 1;
    std::printf("bad_input=rejected\n");
  }
  // solve before initialize
  try {
    VECTOR x, b(A._nrow, 1.0);
    M.solve(x, b);
    return 1;
  } catch (const Error& e) {
    std::printf("solve_before_init=rejected\n");
  }
  return 0;
}

static int gpu_checks() {
  Vector3Array pos;
  SparseColumnMatrix A = poisson3d(16, 1.0, 1234, pos);
  const int n = A._nrow;
  VECTOR b(n);
  check(rsc_uniform_rhs(n, 5678, b.data()), "rsc_uniform_rhs");
  FactorPreconditioner M;
  M.initialize(A, pos, 63, 400, CHOLESKY, IN_PLACE, DECOMPOSE_TRANSPOSE, DECOMPOSE_FACTOR,
               true);
  PCG_SOLVER solver(A, 1e-8, 1000);
  VECTOR x(n, 0.0);
  solver.solvePCG(x, b, &M);
  const double rr = true_relres(A, x, b);
  std::printf("solve iterations=%d recurrence=%.3e true=%.3e\n", solver.lastIterations(),
              solver.lastResidual(), rr);
  if (!(solver.lastResidual() <= 1e-8) || !(rr < 1e-7)) return 1;

  // x = M^{-1} b through the PRECONDITIONER interface
  VECTOR y;
  M.solve(y, b);
  std::printf("apply ok=%d\n", (int)(y.size() == (size_t)n && std::isfinite(y[0])));

  // refactor: same pattern, values doubled -> solution halves
  for (double& v : A._x) v *= 2.0;
  M.clearFactor();
  M.initialize(A, pos, 63, 400, CHOLESKY, IN_PLACE, DECOMPOSE_TRANSPOSE, DECOMPOSE_FACTOR,
               true);
  PCG_SOLVER solver2(A, 1e-8, 1000);
  VECTOR x2(n, 0.0);
  solver2.solvePCG(x2, b, &M);
  double d = 0, s = 0;
  for (int i = 0; i < n; ++i) {
    d += (2.0 * x2[i] - x[i]) * (2.0 * x2[i] - x[i]);
    s += x[i] * x[i];
  }
  std::printf("refactor iterations=%d scaled_diff=%.3e\n", solver2.lastIterations(),
              std::sqrt(d / s));
  if (std::sqrt(d / s) > 1e-7) return 1;
  if (solver2.lastIterations() != solver.lastIterations()) return 1;
  std::printf("mean_iterations=%.1f\n", solver2.meanIterations());
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s cpu|gpu\n", argv[0]);
    return 2;
  }
  try {
    return std::strcmp(argv[1], "gpu") == 0 ? gpu_checks() : cpu_checks();
  } catch (const Error& e) {
    std::printf("error status=%d what=%s\n", (int)e.status, e.what());
    return 1;
  }
}
