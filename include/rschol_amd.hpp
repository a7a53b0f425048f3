// rschol_amd.hpp -- header-only C++ mirror of the reference's preconditioner
// interface over the C ABI in rsc.h, for C++ callers of rschol:
//
//   PRECONDITIONER          src/rschol/linearalgebra/PRECONDITIONER.h:48-57
//   FactorPreconditioner    src/rschol/solver/FactorPreconditioner.h:55-139
//   PCG_SOLVER              src/rschol/linearalgebra/PCG_SOLVER.h:45-122
//   SparseColumnMatrix      src/rschol/linearalgebra/SPARSE_MATRIX.h:187-317
//
// Names, argument order and meaning follow the reference.  Differences:
//   * failures throw rschol_amd::Error (the reference abort()s);
//   * VECTOR is std::vector<double>, Vector3Array a vector of 3-arrays;
//   * the numeric factor and the PCG iterations run on the current HIP
//     device: PCG_SOLVER::solvePCG runs the whole recurrence there
//     (rsc_pcg_solve_host) and takes a FactorPreconditioner initialized on
//     the same matrix (other preconditioners are not supported).
#ifndef RSCHOL_AMD_HPP
#define RSCHOL_AMD_HPP

#include <array>
#include <stdexcept>
#include <string>
#include <vector>

#include "rsc.h"

namespace rschol_amd {

typedef double Real;
typedef std::vector<Real> VECTOR;
typedef std::array<Real, 3> VEC3F;
typedef std::vector<VEC3F> Vector3Array;

enum ExtendedFactorType { CHOLESKY = RSC_CHOLESKY, LDL = RSC_LDL };
enum CompressionType { EXTENDED_VARIABLE = RSC_EXTENDED_VARIABLE, IN_PLACE = RSC_IN_PLACE };
enum DecompositionBlock {
  DECOMPOSE_STANDARD = RSC_DECOMPOSE_STANDARD,
  DECOMPOSE_TRANSPOSE = RSC_DECOMPOSE_TRANSPOSE
};
enum DecompositionType { DECOMPOSE_SCHUR = RSC_DECOMPOSE_SCHUR, DECOMPOSE_FACTOR = RSC_DECOMPOSE_FACTOR };

class Error : public std::runtime_error {
 public:
  Error(rsc_status s, const std::string& what) : std::runtime_error(what), status(s) {}
  rsc_status status;
};

inline void check(rsc_status s, const char* where) {
  if (s != RSC_OK) {
    const char* msg = rsc_last_error();
    throw Error(s, std::string(where) + ": " + (msg ? msg : "error"));
  }
}

// Full symmetric matrix in CSC form: _p[n+1], _i[nnz], _x[nnz].
struct SparseColumnMatrix {
  int _nrow = 0, _ncol = 0;
  std::vector<int> _p, _i;
  std::vector<Real> _x;
  int rows() const { return _nrow; }
  int cols() const { return _ncol; }
  long nnz() const { return _p.empty() ? 0 : _p.back(); }
};

class PRECONDITIONER {
 public:
  virtual ~PRECONDITIONER() {}
  virtual void init() {}
  virtual void solve(VECTOR& x, VECTOR& b) = 0;
};

class FactorPreconditioner : public PRECONDITIONER {
 public:
  FactorPreconditioner() { rsc_options_default(&_options); }
  ~FactorPreconditioner() override { clearSystem(); }
  FactorPreconditioner(const FactorPreconditioner&) = delete;
  FactorPreconditioner& operator=(const FactorPreconditioner&) = delete;

  void clearSystem() {
    if (_factor) rsc_factor_destroy(_factor);
    _factor = nullptr;
    _symbolicFactorNeeded = true;
    _numericFactorNeeded = true;
  }
  void clearFactor() {
    if (!_factor) _symbolicFactorNeeded = true;
    _numericFactorNeeded = true;
  }
  // The factor is symmetric: M^{-T} = M^{-1}.
  void setUseTranspose(bool useTranspose) { _useTranspose = useTranspose; }

  // Extra knobs the reference hard-codes (rank constants, explicit block
  // threshold, precision); set before initialize().
  rsc_options& options() { return _options; }

  void initialize(const SparseColumnMatrix& matrix, const Vector3Array& meshPositions,
                  int maxDenseBlock, int maxDiagonalBlock, ExtendedFactorType extFactorType,
                  CompressionType compressionType, DecompositionBlock blockType,
                  DecompositionType decompType, bool reorderSeparators) {
    _options.max_dense_block = maxDenseBlock;
    _options.max_diagonal_block = maxDiagonalBlock;
    _options.extended_factor_type = extFactorType;
    _options.compression_type = compressionType;
    _options.decomposition_block = blockType;
    _options.decomposition_type = decompType;
    _options.reorder_separators = reorderSeparators ? 1 : 0;
    if (_factor && !_symbolicFactorNeeded && _numericFactorNeeded &&
        matrix._p == _pattern_p && matrix._i == _pattern_i) {
      check(rsc_factor_refactor(_factor, matrix._x.data()), "rsc_factor_refactor");
      _values = matrix._x;
      _numericFactorNeeded = false;
      return;
    }
    if (!_factor || _symbolicFactorNeeded || _numericFactorNeeded) {
      clearSystem();
      std::vector<Real> xyz;
      if (!meshPositions.empty()) {
        xyz.resize(3 * meshPositions.size());
        for (size_t k = 0; k < meshPositions.size(); ++k)
          for (int d = 0; d < 3; ++d) xyz[3 * k + d] = meshPositions[k][d];
      }
      check(rsc_factor_create(matrix._nrow, matrix._p.data(), matrix._i.data(),
                              matrix._x.data(), xyz.empty() ? nullptr : xyz.data(),
                              &_options, &_factor),
            "rsc_factor_create");
      _pattern_p = matrix._p;
      _pattern_i = matrix._i;
      _values = matrix._x;
      _symbolicFactorNeeded = false;
      _numericFactorNeeded = false;
    }
  }

  // initialize() with a caller-supplied nested dissection in CHOLMOD's
  // format (perm[new] = old, cparent per component, cmember per original
  // index), e.g. a reference run's dumps read by readReferenceOrdering
  // (rschol_amd_io.hpp).  The modes are the reference drivers' (CHOLESKY,
  // IN_PLACE, DECOMPOSE_TRANSPOSE, DECOMPOSE_FACTOR).
  void initializeWithOrdering(const SparseColumnMatrix& matrix,
                              const Vector3Array& meshPositions, const std::vector<int>& perm,
                              const std::vector<int>& cparent,
                              const std::vector<int>& cmember, int maxDenseBlock,
                              int maxDiagonalBlock, bool reorderSeparators) {
    clearSystem();
    _options.max_dense_block = maxDenseBlock;
    _options.max_diagonal_block = maxDiagonalBlock;
    _options.reorder_separators = reorderSeparators ? 1 : 0;
    std::vector<Real> xyz;
    if (!meshPositions.empty()) {
      xyz.resize(3 * meshPositions.size());
      for (size_t k = 0; k < meshPositions.size(); ++k)
        for (int d = 0; d < 3; ++d) xyz[3 * k + d] = meshPositions[k][d];
    }
    if ((int)perm.size() != matrix._nrow || (int)cmember.size() != matrix._nrow)
      throw Error(RSC_BAD_INPUT, "ordering size differs from the matrix");
    check(rsc_factor_create_with_ordering(
              matrix._nrow, matrix._p.data(), matrix._i.data(), matrix._x.data(),
              xyz.empty() ? nullptr : xyz.data(), perm.data(), (int)cparent.size(),
              cparent.data(), cmember.data(), &_options, &_factor),
          "rsc_factor_create_with_ordering");
    _pattern_p = matrix._p;
    _pattern_i = matrix._i;
    _values = matrix._x;
    _symbolicFactorNeeded = false;
    _numericFactorNeeded = false;
  }

  // x = M^{-1} b
  void solve(VECTOR& x, VECTOR& b) override {
    if (!_factor) throw Error(RSC_BAD_INPUT, "FactorPreconditioner::solve before initialize");
    x.resize(b.size());
    check(rsc_apply_host(_factor, 1, b.data(), (int)b.size(), x.data(), (int)x.size()),
          "rsc_apply_host");
  }

  rsc_factor* handle() const { return _factor; }
  bool sameMatrix(const SparseColumnMatrix& A) const {
    return _factor && A._p == _pattern_p && A._i == _pattern_i && A._x == _values;
  }
  rsc_factor_info info() const {
    rsc_factor_info inf{};
    if (_factor) check(rsc_factor_get_info(_factor, &inf), "rsc_factor_get_info");
    return inf;
  }

 private:
  rsc_options _options{};
  rsc_factor* _factor = nullptr;
  std::vector<int> _pattern_p, _pattern_i;
  std::vector<Real> _values;
  bool _useTranspose = false;
  bool _symbolicFactorNeeded = true;
  bool _numericFactorNeeded = true;
};

class PCG_SOLVER {
 public:
  PCG_SOLVER(SparseColumnMatrix& A, Real eps = 1e-4, int iterations = 1000)
      : _A(A), _iterations(iterations), _eps(eps) {}

  // Solves A x = b with preconditioner M; x holds the initial guess.
  void solvePCG(VECTOR& x, VECTOR& b, PRECONDITIONER* M) {
    x.resize(b.size(), 0.0);
    FactorPreconditioner* F = dynamic_cast<FactorPreconditioner*>(M);
    if (F && F->sameMatrix(_A)) {
      rsc_pcg_result res{};
      rsc_status s = rsc_pcg_solve_host(F->handle(), b.data(), x.data(), _eps, _iterations,
                                        &res);
      if (s != RSC_OK && s != RSC_NOT_CONVERGED) check(s, "rsc_pcg_solve_host");
      record(res.iterations, res.recurrence_relres, res.seconds);
      return;
    }
    throw Error(RSC_UNSUPPORTED_MODE,
                "PCG_SOLVER::solvePCG needs a FactorPreconditioner initialized on this matrix");
  }

  int& maxIterations() { return _iterations; }
  Real& eps() { return _eps; }
  Real meanIterations() { return (Real)_totalIterations / _totalSolves; }
  Real meanResidual() { return _totalResidual / _totalSolves; }
  double& totalTime() { return _totalTime; }
  int lastIterations() const { return _lastIterations; }
  Real lastResidual() const { return _lastResidual; }

 private:
  void record(int it, Real res, double seconds) {
    ++_totalSolves;
    _totalIterations += it;
    _totalResidual += res;
    _totalTime += seconds;
    _lastIterations = it;
    _lastResidual = res;
  }

  SparseColumnMatrix& _A;
  int _iterations;
  Real _eps;
  int _totalSolves = 0;
  int _totalIterations = 0;
  Real _totalResidual = 0;
  double _totalTime = 0;
  int _lastIterations = 0;
  Real _lastResidual = 0;
};

}  // namespace rschol_amd

#endif
