/*
 * rsc.h -- C ABI of the MI355X (gfx950) rank-structured supernodal Cholesky
 * preconditioner and its PCG solve.
 *
 * Every entry point is extern "C", takes plain pointers and sizes and returns
 * an rsc_status.  Matrices are passed as the FULL symmetric matrix in
 * compressed-sparse-column form with int32 indices and float64 values, the
 * layout of SPARSE_MATRIX::SparseColumnMatrix (src/rschol/linearalgebra/
 * SPARSE_MATRIX.h:187-317, fields _p/_i/_x).
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   rsc_factor_create   FactorPreconditioner::initialize
 *                       (src/rschol/solver/FactorPreconditioner.h:88-99,
 *                        FactorPreconditioner.cpp:70-131) -> SparseSolver::
 *                       factorSymbolic/factorNumeric (SparseSolver.cpp:95-218)
 *   rsc_factor_refactor FactorPreconditioner::clearFactor + initialize with
 *                       the same pattern (FactorPreconditioner.h:69-75,
 *                       FactorPreconditioner.cpp:115-131)
 *   rsc_apply           FactorPreconditioner::solve
 *                       (FactorPreconditioner.cpp:135-162)
 *   rsc_pcg_solve       PCG_SOLVER::solvePCG (PCG_SOLVER.cpp:226-363)
 *   rsc_block_pcg_solve solvePCG applied to k right-hand sides at once (no
 *                       reference counterpart for the batching; per-column
 *                       arithmetic is that of solvePCG)
 *   rsc_factor_destroy  FactorPreconditioner::~FactorPreconditioner /
 *                       clearSystem (FactorPreconditioner.h:63-68)
 *   pcg_pysolve         pcg_pysolve (src/python/pcg_pysolver.h:35-36,
 *                       pcg_pysolver.cpp:83-101)
 *
 * Error behaviour: the reference abort()s on everything except an indefinite
 * diagonal, which it retries (SparseSolver.cpp:174-218).  Here the same retry
 * runs inside rsc_factor_create/refactor and every other failure is returned
 * as a status code; rsc_last_error() gives a message.
 */
#ifndef RSC_H
#define RSC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  RSC_OK = 0,
  RSC_INDEFINITE = 1,        /* indefinite diagonal after all retries */
  RSC_OOM = 2,               /* device or host allocation failed */
  RSC_BAD_INPUT = 3,         /* malformed matrix / arguments */
  RSC_UNSUPPORTED_MODE = 4,  /* enum combination not built (see options) */
  RSC_DEVICE_ERROR = 5,      /* HIP runtime error */
  RSC_NOT_CONVERGED = 6      /* PCG hit max iterations (result still valid) */
} rsc_status;

/* Enum values follow src/rschol/TYPES.h:134-165. */
enum { RSC_CHOLESKY = 0, RSC_LDL = 1 };                       /* ExtendedFactorType */
enum { RSC_EXTENDED_VARIABLE = 0, RSC_IN_PLACE = 1 };         /* CompressionType */
enum { RSC_DECOMPOSE_STANDARD = 0, RSC_DECOMPOSE_TRANSPOSE = 1 }; /* DecompositionBlock */
enum { RSC_DECOMPOSE_SCHUR = 0, RSC_DECOMPOSE_FACTOR = 1 };   /* DecompositionType */

/* Arithmetic of the numeric factor and of the randomized sketches. */
enum { RSC_FP64 = 0, RSC_FP32 = 1, RSC_BF16 = 2 };

typedef struct {
  /* FactorPreconditioner::initialize arguments (pcg_solver.cpp:100-110) */
  int max_dense_block;          /* separators larger than this are compressed; 200 */
  int max_diagonal_block;       /* diagonal leaf size of compressed separators; 400 */
  int extended_factor_type;     /* RSC_CHOLESKY only */
  int compression_type;         /* RSC_IN_PLACE only */
  int decomposition_block;      /* RSC_DECOMPOSE_TRANSPOSE only */
  int decomposition_type;       /* RSC_DECOMPOSE_FACTOR only */
  int reorder_separators;       /* use positions to reorder separators; 1 */
  /* constants hard-coded in the reference, exposed here */
  int explicit_block_threshold; /* FactorManager.cpp:312 -> 4000 */
  double rank_constant;         /* FactorManager.cpp:58 -> 0.4 */
  double diagonal_rank_constant;/* FactorManager.cpp:63 -> 0.5 */
  int oversampling;             /* FactorManager.cpp:68 -> 5 */
  int min_rank;                 /* FactorManager.cpp:7301 -> 16 */
  int max_rank;                 /* 0 = reference heuristic, >0 caps every rank */
  int power_iterations;         /* FactorManager.cpp:71 -> 1 */
  uint32_t seed;                /* FactorManager.cpp:384 -> 152 */
  int max_attempts;             /* SparseSolver.cpp:63 -> 10 */
  double retry_alpha;           /* SparseSolver.cpp:62 -> 1.25 */
  int nd_small;                 /* leaf size of the nested dissection -> 200 */
  /* MI355X build options */
  int factor_precision;         /* RSC_FP64 or RSC_FP32 */
  int sketch_precision;         /* products of the randomized range finder:
                                   RSC_FP64 (default: the factor precision),
                                   RSC_FP32 or RSC_BF16 (BF16 products with
                                   FP32 accumulation) with an FP32 factor;
                                   V = L21 U stays in the factor precision */
  int verbose;                  /* 0 = silent */
} rsc_options;

typedef struct {
  int iterations;               /* PCG iterations taken */
  double recurrence_relres;     /* ||r_k||/||b|| from the recurrence (stop test) */
  double true_relres;           /* ||b - A x||/||b|| recomputed after the loop */
  double seconds;               /* wall time of the solve on the device */
} rsc_pcg_result;

typedef struct {
  int n;
  int nsuper;                   /* supernodes */
  int ncompressed;              /* compressed separator supernodes */
  int nlevels;                  /* etree levels (batched launches) */
  int attempts;                 /* numeric factorizations run (retries + 1) */
  double final_diagonal_rank_constant;
  int64_t factor_bytes;         /* device bytes of the stored factor */
  int64_t workspace_bytes;      /* device bytes of factor workspace */
  double symbolic_seconds;
  double numeric_seconds;
  double factor_flops;          /* algorithmic flops of the numeric factor */
  double apply_bytes;           /* algorithmic bytes of one preconditioner apply */
  double spmv_bytes;            /* algorithmic bytes of one SpMV */
  int pivot_replacements;       /* FP32 factor: pivots raised to the floor */
} rsc_factor_info;

typedef struct rsc_factor rsc_factor;

/* Fill *opt with the reference's defaults (pcg_solver.cpp:100-113 plus the
 * FactorManager constants).  Returns RSC_OK. */
rsc_status rsc_options_default(rsc_options* opt);

/* Symbolic + numeric factorization of the full symmetric CSC matrix
 * (host pointers).  xyz: n*3 float64 positions (x,y,z per unknown) or NULL.
 * With positions, nested dissection is geometric and separators are
 * reordered; without, graph-based.  Ordering stays on the host; the numeric
 * factor runs on the current HIP device. */
rsc_status rsc_factor_create(int n, const int* colptr, const int* rowind,
                             const double* values, const double* xyz,
                             const rsc_options* opt, rsc_factor** out);

/* Same as rsc_factor_create but with a caller-supplied nested dissection in
 * CHOLMOD's format: perm[new] = old, cparent[component], cmember[old]. */
rsc_status rsc_factor_create_with_ordering(int n, const int* colptr,
                             const int* rowind, const double* values,
                             const double* xyz, const int* perm,
                             int ncomponents, const int* cparent,
                             const int* cmember, const rsc_options* opt,
                             rsc_factor** out);

/* New numeric values with the identical sparsity pattern (host pointer,
 * nnz entries in the original CSC order).  Reuses the symbolic plan. */
rsc_status rsc_factor_refactor(rsc_factor* f, const double* values);

/* X = M^{-1} B for nrhs columns; B and X are DEVICE pointers (float64,
 * column-major, leading dimensions ldb/ldx >= n).  stream: hipStream_t or
 * NULL for the factor's own stream.  The apply is queued asynchronously, so
 * RSC_OK means "queued": a device-side failure of the apply (a dependency
 * wait of the dataflow launch that ran out of time) is reported by
 * rsc_apply_status, which the caller runs before trusting X. */
rsc_status rsc_apply(rsc_factor* f, int nrhs, const double* B, int ldb,
                     double* X, int ldx, void* stream);

/* Synchronizes stream (NULL: the factor's own) and returns RSC_DEVICE_ERROR
 * if an apply queued on it since the last check failed on the device, else
 * RSC_OK.  Counterpart of the reference's synchronous solve returning
 * (FactorPreconditioner.cpp:135-162): the result of rsc_apply is valid once
 * this returns RSC_OK.  The PCG and host-pointer entry points check this
 * themselves. */
rsc_status rsc_apply_status(rsc_factor* f, void* stream);

/* rsc_apply without any allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory during the call (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and results (bit for bit) as rsc_apply; the
 * per-call scratch (permuted and front vectors, rank vectors, partial sums)
 * comes from a fixed 64 MiB device array per GPU, held until the process
 * ends, instead of buffers the library grows.  Applies whose scratch exceeds
 * it (about 64 MiB / (8 (2n + sum of front sizes)) right-hand sides: e.g.
 * one right-hand side of a 64^3 Poisson system, not a 128^3 one) return
 * RSC_BAD_INPUT before any work.  All calls on one GPU share the array:
 * calls that overlap on one GPU race, so the caller orders them.  Check the
 * result with rsc_apply_status as for rsc_apply. */
rsc_status rsc_apply_noalloc(rsc_factor* f, int nrhs, const double* B, int ldb,
                             double* X, int ldx, void* stream);

/* PCG on A x = b with the factor as preconditioner.  b, x: DEVICE pointers;
 * x holds the initial guess on entry (the reference driver passes zero).
 * Stops when ||r||/||b|| <= eps or after max_iters. */
rsc_status rsc_pcg_solve(rsc_factor* f, const double* b, double* x,
                         double eps, int max_iters, rsc_pcg_result* res,
                         void* stream);

/* k independent PCG recurrences advanced together (one SpMV with k vectors
 * and one preconditioner apply with k right-hand sides per iteration).
 * Column j follows the solvePCG arithmetic exactly and stops on its own
 * test; res has k entries. */
rsc_status rsc_block_pcg_solve(rsc_factor* f, int k, const double* B, int ldb,
                               double* X, int ldx, double eps, int max_iters,
                               rsc_pcg_result* res, void* stream);

/* Host-pointer convenience forms (copy in, solve on the device, copy out). */
rsc_status rsc_apply_host(rsc_factor* f, int nrhs, const double* B, int ldb,
                          double* X, int ldx);
rsc_status rsc_pcg_solve_host(rsc_factor* f, const double* b, double* x,
                              double eps, int max_iters, rsc_pcg_result* res);
rsc_status rsc_block_pcg_solve_host(rsc_factor* f, int k, const double* B,
                                    int ldb, double* X, int ldx, double eps,
                                    int max_iters, rsc_pcg_result* res);

/* ---- request server: continuous batching over resident factors ----------
 * No counterpart in the reference (its drivers solve one system at a time,
 * src/pcg_solver.cpp:70-120 and python/pcg_pysolver.cpp:83-101); this is a
 * caller of the preconditioner and PCG path for serving loops.
 *
 * The server owns a pool of column slots per factor (slot memory: b, x and
 * five PCG vectors of n doubles, plus the apply's per-column buffers).  A
 * request carries nrhs right-hand sides for one factor; its columns are
 * admitted into free slots in arrival order as slots free up, every busy
 * slot advances one PCG iteration per step, and a converged column is
 * retired (x copied out, true residual computed) while the others go on.
 * Factors on different streams run concurrently.  Each column's arithmetic
 * is that of rsc_block_pcg_solve with x0 = 0, so a result does not depend
 * on which requests share the device with it.
 *
 * Factors passed to a server must not be used by any other call until
 * rsc_server_destroy returns. */
typedef struct rsc_server rsc_server;
typedef struct {
  double eps;          /* relative residual target of every column */
  int max_iters;       /* per column */
  int slots;           /* most column slots of one factor (1..64) */
  /* Device memory for the column slots of all factors together (each
   * factor's slot vectors and the apply buffers of its columns).  The
   * budget is divided among the factors by water-filling over their bytes
   * per column, up to `slots` each and at least one; 0 = half of the device
   * memory free when the server is created. */
  int64_t memory_bytes;
  /* 1: release every factor's numeric workspace at creation (as
   * rsc_factor_release_workspace), since served factors are not
   * refactorized while the server exists */
  int release_workspace;
  int reserved;
} rsc_server_options;
/* 1e-8, 10000, 64 slots, memory_bytes 0, release_workspace 1 */
rsc_status rsc_server_options_default(rsc_server_options* o);
rsc_status rsc_server_create(rsc_factor* const* factors, int nfactors,
                             const rsc_server_options* opt, rsc_server** out);
/* B (n x nrhs, ld ldb) and X (n x nrhs, ld ldx) are DEVICE pointers that
 * stay valid until rsc_server_wait returns for the ticket. */
rsc_status rsc_server_submit(rsc_server* s, int factor, int nrhs, const double* B,
                             int ldb, double* X, int ldx, int64_t* ticket);
/* Blocks until every column of the request is retired; res (nrhs entries,
 * may be NULL) receives each column's result, seconds = the request's
 * latency from submission to completion. */
rsc_status rsc_server_wait(rsc_server* s, int64_t ticket, rsc_pcg_result* res);
/* Steps run and the largest number of busy slots of any step, so far. */
rsc_status rsc_server_stats(const rsc_server* s, int64_t* steps, int* max_busy);
/* Column slots the memory budget gave factor `factor` (its index at
 * creation). */
rsc_status rsc_server_slots(const rsc_server* s, int factor, int* slots);
void rsc_server_destroy(rsc_server* s);

/* y = A x on the device (the PCG operator; DEVICE pointers). */
rsc_status rsc_spmv(rsc_factor* f, const double* x, double* y, void* stream);

rsc_status rsc_factor_get_info(const rsc_factor* f, rsc_factor_info* info);

/* Frees the numeric factorization's workspace (update matrices, work
 * panels, sketch and QR scratch: rsc_factor_info.workspace_bytes, then 0).
 * The factor, its apply and PCG keep working; the next rsc_factor_refactor
 * allocates the workspace and builds the numeric program again.  No
 * counterpart in the reference (its factor frees its transient storage as
 * it goes); for callers that keep many factors resident. */
rsc_status rsc_factor_release_workspace(rsc_factor* f);

/* Permutation used by the factor: perm[new] = old (n entries, host). */
rsc_status rsc_factor_get_permutation(const rsc_factor* f, int* perm);

/* Supernode partition: first column of each supernode (nsuper+1 entries,
 * last = n) and compress flags (nsuper entries). */
rsc_status rsc_factor_get_supernodes(const rsc_factor* f, int* first,
                                     int* compressed);

rsc_status rsc_factor_destroy(rsc_factor* f);

/* Message for the last failing call on this thread. */
const char* rsc_last_error(void);

/* ---- Symbolic analysis on the host (no device needed) ------------------
 * The ordering + structure half of rsc_factor_create, for inspection and for
 * host-only checks: SparseSolver::factorSymbolic (SparseSolver.cpp:95-154)
 * and FactorManager::factorSymbolic (FactorManager.cpp:280-349). */
typedef struct rsc_symbolic rsc_symbolic;
rsc_status rsc_symbolic_create(int n, const int* colptr, const int* rowind,
                               const double* xyz, const rsc_options* opt,
                               rsc_symbolic** out);
rsc_status rsc_symbolic_create_with_ordering(int n, const int* colptr,
                               const int* rowind, const double* xyz,
                               const int* perm, int ncomponents,
                               const int* cparent, const int* cmember,
                               const rsc_options* opt, rsc_symbolic** out);
/* counts: nsuper, total off-diagonal row entries, levels, stored factor
 * elements, Gaussian draws of the sketches */
rsc_status rsc_symbolic_sizes(const rsc_symbolic* s, int* nsuper,
                              int64_t* total_rows, int* nlevels,
                              int64_t* factor_elems, int64_t* rng_count);
/* perm[n]; first[nsuper+1]; compressed[nsuper]; parent[nsuper];
 * rows_ptr[nsuper+1]; rows[total_rows]; rank[nsuper] (off-diagonal rank, 0
 * if uncompressed or no rows); level[nsuper].  Any pointer may be NULL. */
rsc_status rsc_symbolic_get(const rsc_symbolic* s, int* perm, int* first,
                            int* compressed, int* parent, int64_t* rows_ptr,
                            int* rows, int* rank, int* level);
/* Diagonal blocks of supernode `node` (ReorderNestedSeparators leaf ranges):
 * *count = number of leaf blocks (0 if not split); offsets gets count+1
 * relative offsets when non-NULL. */
rsc_status rsc_symbolic_diag_blocks(const rsc_symbolic* s, int node, int* count,
                                    int* offsets);
void rsc_symbolic_destroy(rsc_symbolic* s);

/* ---- Ordering building blocks (host) ------------------------------------ */
/* Nested dissection triple in CHOLMOD's format (CHOLMOD_Environment.cpp:
 * 326-432); cparent/cmember need n entries of space; *ncomponents is set. */
rsc_status rsc_nested_dissection(int n, const int* colptr, const int* rowind,
                                 const double* xyz, int nd_small, int* perm,
                                 int* ncomponents, int* cparent, int* cmember);
/* Elimination tree and column counts of P A P^T (perm[new] = old) with the
 * identity postorder (NestedDissection.cpp:418-429). */
rsc_status rsc_etree_colcounts(int n, const int* colptr, const int* rowind,
                               const int* perm, int* parent, int* colcount);
/* Relaxed supernode partition (NestedDissection::BuildRelaxedSupernodes,
 * NestedDissection.cpp:1266-1667) of an elimination tree with column counts
 * (parent[n], colcount[n]) and ncomp nested-dissection components given by
 * their sizes, first/last columns, child counts and parents, with the
 * reference's relaxation constants.  first[] (n+1 entries of space)
 * receives each supernode's first column, compressed[] (n entries) its flag,
 * *nsuper the count. */
rsc_status rsc_relaxed_supernodes(int n, const int* parent, const int* colcount,
                                  int ncomp, const int* comp_size,
                                  const int* comp_first, const int* comp_last,
                                  const int* comp_nchildren, const int* comp_parent,
                                  int max_dense_block, int* first, int* compressed,
                                  int* nsuper);
/* SeparatorTree leaf order of npts points (xyz), SeparatorTree.cpp:189-589:
 * order[npts] receives the point order, leaf_sizes (<= npts entries) the
 * leaf sizes. */
rsc_status rsc_separator_tree(int npts, const double* xyz, int max_leaf,
                              int* order, int* nleaves, int* leaf_sizes);
/* MT19937 + Box-Muller sketch stream (MathUtil.cpp:100-112,
 * MERSENNETWISTER.h:170-177). */
rsc_status rsc_gaussian_stream(uint32_t seed, int64_t count, double* out);
/* FactorManager::blockRank (FactorManager.cpp:7283-7308) with optional cap. */
int rsc_block_rank(int nrows, int ncols, double constant, int oversampling,
                   int min_rank, int max_rank);

/* ---- Synthetic test systems (no reference counterpart: the benchmark
 * inputs; full symmetric CSC + positions) --------------------------------- */
typedef struct rsc_matrix rsc_matrix;
rsc_status rsc_matrix_poisson2d(int N, rsc_matrix** out);
rsc_status rsc_matrix_poisson3d(int N, double sigma, uint64_t seed,
                                rsc_matrix** out);
rsc_status rsc_matrix_elasticity3d(int NE, double sigma, double nu,
                                   uint64_t seed, rsc_matrix** out);
rsc_status rsc_matrix_info(const rsc_matrix* m, int* n, int64_t* nnz);
/* colptr[n+1], rowind[nnz], values[nnz], xyz[3n]; any may be NULL */
rsc_status rsc_matrix_export(const rsc_matrix* m, int* colptr, int* rowind,
                             double* values, double* xyz);
void rsc_matrix_destroy(rsc_matrix* m);
/* count uniform[-1,1] values from std::mt19937_64(seed) */
rsc_status rsc_uniform_rhs(int64_t count, uint64_t seed, double* out);

/* ---- Kernel self-test (no reference counterpart: exercises the batched
 * MFMA GEMM of the factor on one job, host pointers, column-major):
 * C(m x n) = beta*C + alpha*op(A)*op(B), precision RSC_FP64, RSC_FP32
 * (inputs rounded to FP32) or RSC_BF16 (FP32 operands, BF16 products with
 * FP32 accumulation). */
rsc_status rsc_selftest_gemm(int precision, int transa, int transb, int m, int n,
                             int k, double alpha, const double* A, int lda,
                             const double* B, int ldb, double beta, double* C,
                             int ldc);

/* Exact signature of src/python/pcg_pysolver.h:35-36: CSC arrays (jc = column
 * pointers, ir = row indices, pr = values) of the full symmetric matrix,
 * host pointers.  Factors with maxDenseBlock=200, maxDiagonalBlock=400 and no
 * positions, then runs PCG to 1e-6 relative residual, at most 10000
 * iterations, from x = 0 (pcg_pysolver.cpp:49-78). */
void pcg_pysolve(int nrow, const int* jc, const int* ir, const double* pr,
                 const double* rhs, double* solution);

#ifdef __cplusplus
}
#endif
#endif /* RSC_H */
