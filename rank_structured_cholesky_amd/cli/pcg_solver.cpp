// pcg_solver -- the reference's command-line driver (src/pcg_solver.cpp:
// 50-122) on the MI355X factor, with the same file contract:
//
//   pcg_solver <base> [ordering_dir]
//
//   <base>_system.bcsm.gz    full symmetric CSC matrix (readFromBinaryGZ)
//   <base>_rhs.vector        right-hand side (VECTOR::read)
//   <base>_meshPositions.dat degree-of-freedom positions; when absent, zero
//                            positions and no separator reordering, as the
//                            reference does (:85-96)
//
// Factors with maxDenseBlock 200 and maxDiagonalBlock 400 (CHOLESKY,
// IN_PLACE, DECOMPOSE_TRANSPOSE, DECOMPOSE_FACTOR; :100-110), solves with
// PCG to a relative residual of 1e-6 in at most 10000 iterations from x = 0
// (:113-115) and prints the wall time of factor + solve (:119).  Also
// writes <base>_solution.vector.
//
// The optional second argument (not in the reference driver, which ignores
// extra arguments) names a directory holding a reference run's ordering
// dumps (SparseSolver_permutation.vector, separators.vector,
// separatorParents.vector): the factor then uses that run's nested
// dissection and separator sub-ordering instead of this library's own.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <string>

#include "rschol_amd.hpp"
#include "rschol_amd_io.hpp"

using namespace rschol_amd;

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "Usage %s <matrix file>\n", argv[0]);
    return 1;
  }
  const std::string base = argv[1];
  try {
    SparseColumnMatrix A = readFromBinaryGZ(base + "_system.bcsm.gz");
    std::printf("_nzmax = %ld _nrow = %d _ncol = %d\n", A.nnz(), A._nrow, A._ncol);
    VECTOR rhs = readVectorFile(base + "_rhs.vector");
    if ((int)rhs.size() != A._nrow)
      throw Error(RSC_BAD_INPUT, "right-hand side length differs from the matrix");
    VECTOR solution(rhs.size(), 0.0);
    Vector3Array meshPositions;
    const bool positionsFound = readPositions(base + "_meshPositions.dat", meshPositions);
    if (!positionsFound) {
      // the reference fills zero positions, which only its separator
      // reordering reads; here no positions means the graph dissection and
      // no reordering
      meshPositions.clear();
      std::fprintf(stderr,
                   "Warning: no system degree of freedom positions found - "
                   "using default separator ordering\n");
    } else if ((int)meshPositions.size() != A._nrow) {
      throw Error(RSC_BAD_INPUT, "positions count differs from the matrix");
    }

    const auto t0 = std::chrono::steady_clock::now();
    FactorPreconditioner factorPrecond;
    if (argc >= 3) {
      Ordering ord = readReferenceOrdering(argv[2]);
      factorPrecond.initializeWithOrdering(A, meshPositions, ord.perm, ord.cparent,
                                           ord.cmember, 200, 400, false);
    } else {
      factorPrecond.initialize(A, meshPositions, 200, 400, CHOLESKY, IN_PLACE,
                               DECOMPOSE_TRANSPOSE, DECOMPOSE_FACTOR,
                               // only reorder separators with real positions
                               positionsFound);
    }
    const auto t1 = std::chrono::steady_clock::now();
    // relative residual 1e-6, at most 10000 iterations
    PCG_SOLVER solver(A, 1e-6, 10000);
    solver.solvePCG(solution, rhs, &factorPrecond);
    const auto t2 = std::chrono::steady_clock::now();

    const rsc_factor_info inf = factorPrecond.info();
    std::printf("PCG iterations = %d, relative residual = %.6e\n", solver.lastIterations(),
                solver.lastResidual());
    std::printf("factor: %d supernodes, %d compressed, %.1f MB, %d attempt(s); "
                "%.3f s, solve %.3f s\n",
                inf.nsuper, inf.ncompressed, inf.factor_bytes / 1e6, inf.attempts,
                std::chrono::duration<double>(t1 - t0).count(),
                std::chrono::duration<double>(t2 - t1).count());
    std::printf("\nPCG solution took %f seconds\n",
                std::chrono::duration<double>(t2 - t0).count());
    writeVectorFile(base + "_solution.vector", solution);
    return solver.lastResidual() <= 1e-6 ? 0 : 3;
  } catch (const Error& e) {
    std::fprintf(stderr, "pcg_solver: %s (status %d)\n", e.what(), (int)e.status);
    return 2;
  }
}
