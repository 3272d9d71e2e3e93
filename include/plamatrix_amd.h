/*
 * plamatrix_amd.h -- C ABI of the MI355X (gfx950) sparse solve library.
 *
 * This is the drop-in boundary under PlaMatrix's C++ API (COOMatrix /
 * CSRMatrix, cooToCsr, spmv, cg / pcg / blockPcg, the fixed-iteration async
 * solves and the damped Schur-complement solve).  The C++ headers in
 * include/plamatrix/ implement those templates on top of these entry points;
 * any other host language (ctypes, cgo, JNI) binds the same symbols.
 *
 * Conventions
 *   - All array arguments are device pointers unless stated otherwise.
 *   - Every entry returns a pm_status.  On failure pm_last_error() holds the
 *     message for the calling thread.  Codes map onto the reference's
 *     exception classes (std::invalid_argument, std::out_of_range, ...).
 *   - Work is enqueued on `stream` (NULL = the legacy default stream) and is
 *     stream ordered; entries that must read results on the host say so.
 *   - A pm_workspace is not thread-safe and binds to the first stream it is
 *     used on, as SparseOpsWorkspace / IterativeSolverWorkspace do
 *     (sparse_ops.h:13-16, iterative_solver.h:29-33).
 */
#ifndef PLAMATRIX_AMD_H
#define PLAMATRIX_AMD_H

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum pm_status {
    PM_OK = 0,
    PM_INVALID_ARGUMENT = 1, /* std::invalid_argument */
    PM_OUT_OF_RANGE = 2,     /* std::out_of_range */
    PM_OVERFLOW = 3,         /* std::overflow_error */
    PM_LOGIC_ERROR = 4,      /* std::logic_error (stream / ownership misuse, pending work) */
    PM_RUNTIME_ERROR = 5,    /* std::runtime_error (breakdown, non-convergence, non-SPD) */
    PM_DEVICE_ERROR = 6      /* HIP runtime error (std::runtime_error in C++) */
} pm_status;

typedef enum pm_dtype { PM_F32 = 0, PM_F64 = 1 } pm_dtype;

/* Message of the last failing call on this thread ("" when none). */
const char* pm_last_error(void);
int pm_version(void);

/* ---- CSR descriptor ---------------------------------------------------------
 * Mirrors CSRMatrix<Scalar, Device::GPU> (sparse/csr_matrix.h:22-825): int64
 * row offsets and column indices, values of `dtype`.  The optional 32-bit
 * mirrors (filled by pm_csr_compact_indices) are what the kernels stream when
 * present: 4 instead of 8 bytes per index on the bandwidth-bound path. */
typedef struct pm_csr {
    int64_t rows;
    int64_t cols;
    int64_t nnz;
    const int64_t* row_offsets;
    const int64_t* col_indices;
    const void* values;
    int32_t dtype;
    const int32_t* row_offsets32;
    const int32_t* col_indices32;
} pm_csr;

/* Workspace: device scratch, reduction partials, solver state, sort buffers
 * and cached launch plans.  Replaces SparseOpsWorkspace (sparse_ops.h:13-77)
 * and IterativeSolverWorkspace<S> (iterative_solver.h:29-62). */
typedef struct pm_workspace pm_workspace;
int pm_workspace_create(pm_workspace** ws);
/* Synchronises the bound stream, then frees everything. */
int pm_workspace_destroy(pm_workspace* ws);
/* closeAsyncAllocation(): releases storage and unbinds the stream;
 * PM_LOGIC_ERROR while a COO conversion's status is unchecked or while the
 * bound stream still has work queued (sparse_ops.cu:644-660). */
int pm_workspace_close(pm_workspace* ws);
size_t pm_workspace_capacity_bytes(const pm_workspace* ws);
/* Vector capacity (rows) of the iterative-solver storage (capacitySize()). */
int64_t pm_workspace_capacity_rows(const pm_workspace* ws);
/* No reference counterpart: the number of diagonals of the diagonal layout
 * the last adaptive solve streamed A in (0: the CSR kernels).  Square matrices
 * of at least 65,536 rows whose entries lie on at most 16 diagonals, with
 * strictly increasing columns in every row, are re-laid diagonal-major at the
 * start of each solve (PLAMATRIX_AMD_DIA=0 disables it); results are bitwise
 * those of the CSR kernels. */
int32_t pm_workspace_solver_diagonals(const pm_workspace* ws);
/* No reference counterpart: the value arrays the last adaptive solve's
 * diagonal-layout row sums read per row (0: the CSR kernels).  A matrix that is
 * symmetric bit for bit reads its lower diagonals from their upper mirrors,
 * and a main diagonal of exact ones from no array, so a unit-diagonal 7-point
 * stencil reads 3 arrays instead of 7 (PLAMATRIX_AMD_DIA_MIRROR=0 reads every
 * diagonal); results are bitwise those of the CSR kernels. */
int32_t pm_workspace_solver_value_streams(const pm_workspace* ws);
/* No reference counterpart: the iteration form the last adaptive solve
 * started in.  3: the classic form (direction, SpMV, update; block-Jacobi,
 * factored factors off the diagonal layout, or
 * PLAMATRIX_AMD_SINGLE_REDUCTION=0).  2: the single-reduction iteration of CG
 * and Jacobi-PCG (step, SpMV).  1: that iteration as one kernel on the
 * diagonal layout (PLAMATRIX_AMD_CG_FUSED=0 keeps 2).  4: the factored
 * single-reduction iteration with A, G and G^T on the diagonal layout (step
 * and G r, G^T t, SpMV).  5: the same with a four-diagonal lower G, G r and
 * G^T t in one kernel (PLAMATRIX_AMD_CG_FUSED=0 keeps 4).  A solve whose
 * single-reduction denominator loses positivity continues in the classic
 * form. */
int32_t pm_workspace_solver_form(const pm_workspace* ws);
/* No reference counterpart: the preconditioner the last adaptive solve ran
 * (0: none, 1: Jacobi, 2: block-Jacobi, 3: factored G^T G).  A Jacobi solve of
 * a matrix of at least 65,536 rows whose every Jacobi inverse is exactly 1 (a
 * diagonally scaled matrix) runs as the unpreconditioned iteration, which is
 * the same computation bit for bit (z = 1 * r), and reports 0
 * (PLAMATRIX_AMD_UNIT_DIAGONAL=0 keeps the Jacobi kernels). */
int32_t pm_workspace_solver_preconditioner(const pm_workspace* ws);

/* Structural validation on the device (CSRMatrix::validateStructure,
 * csr_matrix.h:197-237, 633-667): offsets start at 0, end at nnz and are
 * monotone, columns in [0, cols), values finite.  Synchronises `stream`;
 * PM_INVALID_ARGUMENT names the first offending row or position. */
int pm_csr_validate(const pm_csr* a, pm_workspace* ws, hipStream_t stream);

/* Fill 32-bit index mirrors (rows+1 and nnz entries).  PM_OVERFLOW when
 * nnz or cols do not fit in int32. */
int pm_csr_compact_indices(const pm_csr* a, int32_t* row_offsets32, int32_t* col_indices32,
                           hipStream_t stream);

/* ---- COO -> CSR ---------------------------------------------------------------
 * Replaces cooToCsr / cooToCsrAsync for device triplets (sparse_ops.h:94-120,
 * coo_to_csr.cu:430-576).  Stable (row, col) order and left fold of
 * duplicates in input order starting from the first value, bit-identical to
 * the CPU cooToCsr (sparse_ops_cpu.cpp:24-134).  Triplets are not modified.
 * Limits: count < 2^31, ceil(log2 rows) + ceil(log2 cols) <= 64. */
enum { PM_COO_CHECK_FINITE = 1 };

/* Validate, sort and count (asynchronous). */
int pm_coo_to_csr_prepare(int64_t rows, int64_t cols, int64_t count, const int64_t* row_indices,
                          const int64_t* col_indices, const void* values, int32_t dtype, int32_t flags,
                          pm_workspace* ws, hipStream_t stream);
/* Synchronise and return the combined nnz of the prepared conversion; raises
 * the deferred coordinate / value errors (lowest triplet position first). */
int pm_coo_to_csr_nnz(pm_workspace* ws, int64_t* nnz);
/* Fold duplicates and write CSR arrays of exactly expected_nnz entries
 * (asynchronous; a mismatch is reported by pm_workspace_check_status). */
int pm_coo_to_csr_finish(int64_t* row_offsets, int64_t* col_indices, void* values, int64_t expected_nnz,
                         pm_workspace* ws, hipStream_t stream);
/* prepare + finish in one asynchronous call (cooToCsrAsync). */
int pm_coo_to_csr_async(int64_t rows, int64_t cols, int64_t count, const int64_t* row_indices,
                        const int64_t* col_indices, const void* values, int32_t dtype, int32_t flags,
                        int64_t* row_offsets, int64_t* out_col_indices, void* out_values, int64_t expected_nnz,
                        pm_workspace* ws, hipStream_t stream);
/* SparseOpsWorkspace::checkStatus (sparse_ops.h:27): PM_LOGIC_ERROR while the
 * conversion is still running, else the deferred error, else PM_OK. */
int pm_workspace_check_status(pm_workspace* ws);

/* ---- SpMV -------------------------------------------------------------------
 * y = A x (spmvAsync, sparse_ops.h:158-166; sparse_ops.cu:687-740).  x and y
 * have a->dtype; accumulation in that precision.  x and y must not alias. */
int pm_spmv(const pm_csr* a, const void* x, void* y, pm_workspace* ws, hipStream_t stream);
/* spmmAsync (sparse_ops.h:180-188; CPU rule sparse_ops_cpu.cpp:244-268):
 * Y = A B with B (a->cols x k) and Y (a->rows x k) column-major, leading
 * dimensions a->cols and a->rows, in a's dtype.  Each (row, column) sum runs
 * left to right over the row in the scalar type.  B == Y is rejected
 * (invalid_argument); k == 0 or a->rows == 0 does nothing; nnz == 0 writes
 * zeros.  ws is required (it holds a packed copy of 8 columns of B at a
 * time).  Stream-ordered, no host synchronisation. */
int pm_spmm(const pm_csr* a, const void* b, int64_t k, void* y, pm_workspace* ws, hipStream_t stream);
/* pm_spmm without a workspace and without any allocation: no counterpart in
 * the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The packed tile of B lives in a 64 MiB device array of the library that each
 * GPU holds from the library's load until the process ends, so a->cols * 8 *
 * sizeof(value) must not exceed 64 MiB (2,097,152 columns in FP32, 1,048,576
 * in FP64); larger matrices get PM_OVERFLOW.  Same results as pm_spmm, bit for
 * bit.  Every call on one GPU shares that array: calls that overlap on one GPU
 * (other streams, other threads) race, so serialise them. */
int pm_spmm_noalloc(const pm_csr* a, const void* b, int64_t k, void* y, hipStream_t stream);
/* Same with an explicit lanes-per-row choice (1..64, power of two); used by
 * the autotuner. */
int pm_spmv_lanes(const pm_csr* a, const void* x, void* y, int32_t lanes_per_row, pm_workspace* ws,
                  hipStream_t stream);
/* Autotuner (no reference counterpart): times every lanes-per-row variant of
 * y = A x on this matrix (median of `repeats` launches) and records the
 * fastest for its shape (rows, nnz, dtype, index width) in memory and in the
 * cache file named by PLAMATRIX_AMD_TUNE_CACHE.  Synchronises `stream`. */
int pm_autotune_spmv(const pm_csr* a, const void* x, void* y, int32_t repeats, int32_t* best_lanes,
                     float* best_ms, hipStream_t stream);
/* The lanes-per-row the kernels will use for this matrix (-1 on error). */
int pm_tuned_lanes(const pm_csr* a);
/* Record a lanes-per-row choice (1..64, power of two) for this matrix shape in
 * the same table the autotuner fills. */
int pm_tune_record(const pm_csr* a, int32_t lanes_per_row);
/* FP64 true relative residual ||b - A x|| / ||b|| (||b - A x|| if b = 0).
 * vec_dtype is the dtype of b and x (PM_F64 allowed with a float matrix).
 * Synchronises `stream`. */
int pm_true_relres(const pm_csr* a, const void* b, const void* x, int32_t vec_dtype, double* relres,
                   pm_workspace* ws, hipStream_t stream);

/* ---- CG / PCG -----------------------------------------------------------------
 * Mirrors IterativeSolverOptions / IterativeSolverReport
 * (iterative_solver.h:10-26). */
typedef struct pm_solver_options {
    int32_t max_iterations;
    double relative_tolerance;
    double absolute_tolerance;
    int32_t use_jacobi;
    int32_t require_convergence;
} pm_solver_options;

typedef struct pm_solver_report {
    int32_t converged;
    int32_t iterations;
    double initial_residual;
    double final_residual;
} pm_solver_report;

/* Execution policy (no counterpart in the reference). Zero-initialised means
 * the defaults. */
typedef struct pm_solve_policy {
    int32_t chunk_iterations;  /* iterations launched per device-side chunk (0: 16 with a refresh
                                  interval; without one 4, 4, then doubling up to 64) */
    int32_t use_graph;         /* 1: replay chunks as HIP graphs, 0: direct launches, -1: default
                                  (graphs below 262,144 rows, direct launches above) */
    int32_t refresh_interval;  /* recompute r = b - A x (FP64 accumulation) every k iterations; 0: never */
    int32_t lanes_per_row;     /* SpMV lanes per row; 0: autotuned / heuristic */
} pm_solve_policy;

/* Adaptive solves (cg / pcg / blockPcg on the GPU, iterative_solver.h:110-141;
 * iterative_solver.cu:262-402).  Convergence is decided on the device; the
 * host polls one small status record per chunk instead of every iteration.
 * x holds the initial guess and receives the solution.  Arithmetic: vectors
 * and SpMV in a->dtype, every dot product and alpha/beta in FP64. */
int pm_cg(const pm_csr* a, const void* rhs, void* x, const pm_solver_options* options,
          const pm_solve_policy* policy, pm_solver_report* report, pm_workspace* ws, hipStream_t stream);
int pm_pcg(const pm_csr* a, const void* rhs, void* x, const pm_solver_options* options,
           const pm_solve_policy* policy, pm_solver_report* report, pm_workspace* ws, hipStream_t stream);
/* inverse_blocks: rows/block_size row-major block_size^2 blocks (rows*block_size
 * entries), block_size <= 16. */
int pm_block_pcg(const pm_csr* a, const void* rhs, void* x, const void* inverse_blocks, int64_t block_size,
                 const pm_solver_options* options, const pm_solve_policy* policy, pm_solver_report* report,
                 pm_workspace* ws, hipStream_t stream);

/* Fixed-iteration asynchronous solves (cgFixedIterationsAsync /
 * pcgFixedIterationsAsync / finalizeIterativeSolverReport,
 * iterative_solver.h:143-171).  Exactly `iterations` steps are enqueued with
 * no host synchronisation; the state records the initial and final squared
 * residuals on the device and an event. */
typedef struct pm_fixed_state pm_fixed_state;
int pm_solve_fixed_async(const pm_csr* a, const void* rhs, void* x, int32_t iterations, int32_t preconditioned,
                         pm_fixed_state** state, pm_workspace* ws, hipStream_t stream);
/* PM_LOGIC_ERROR if the stream has not completed the state's work. */
int pm_fixed_finalize(const pm_fixed_state* state, const pm_solver_options* options, pm_solver_report* report);
int pm_fixed_state_destroy(pm_fixed_state* state);

/* Mixed precision: FP32 matrix and inner Jacobi-PCG (FP64 dots, periodic
 * FP64-accumulated residual refresh), FP64 solution and outer iterative
 * refinement until the FP64 true relative residual reaches target_relres.
 * rhs and x are FP64 device vectors.  Each inner solve runs to `inner`'s
 * relative tolerance, except that a step whose relres times that tolerance is
 * within twice the target runs to 0.9 * target / relres, so that it ends the
 * refinement although the inner solve stops on its recursive residual. */
typedef struct pm_refine_report {
    int32_t converged;
    int32_t outer_iterations;
    int32_t inner_iterations;   /* total over all outer steps */
    double initial_relres;
    double final_relres;        /* FP64 ||b - A x|| / ||b|| */
} pm_refine_report;
int pm_pcg_refine(const pm_csr* a, const double* rhs, double* x, const pm_solver_options* inner,
                  double target_relres, int32_t max_outer, const pm_solve_policy* policy,
                  pm_refine_report* report, pm_workspace* ws, hipStream_t stream);

/* ---- Batched PCG ---------------------------------------------------------------
 * Many independent SPD systems solved together (no counterpart in the
 * reference, which solves one system per call).  The systems are the diagonal
 * blocks of one square CSR matrix `a` (FP32 or FP64): system s owns rows
 * [row_begin[s], row_end[s]) (host arrays), every row_begin a multiple of 64,
 * ranges ascending and disjoint; columns of system s point into its own rows.
 * Rows outside every range are ignored.  Each system stops on its own
 * tolerance max(abs, rel * ||r0||), iteration limit or breakdown, with the
 * rules of the single-system solver; the call returns when all have stopped.
 * options->use_jacobi selects Jacobi-PCG or CG; policy->refresh_interval
 * recomputes r = b - A x with FP64 accumulation.  Synchronises `stream` at the
 * end to fill the host reports. */
typedef struct pm_batch_report {
    int32_t converged;
    int32_t iterations;
    double initial_residual;
    double final_residual;
    int32_t status;   /* 1 converged, 2 iteration limit, 3 breakdown, 4 non-finite,
                         5 non-finite initial residual, 6 bad Jacobi diagonal */
    int32_t pad;
    int64_t bad_row;  /* first rejected diagonal row (status 6), else -1 */
} pm_batch_report;
/* Host helper: row_begin[s] for systems of the given sizes packed at 64-row
 * boundaries, and the total row count of the block-diagonal matrix. */
int pm_batch_layout(int64_t systems, const int64_t* sizes, int64_t* row_begin, int64_t* total_rows);
int pm_batch_pcg(const pm_csr* a, int64_t systems, const int64_t* row_begin, const int64_t* row_end,
                 const void* rhs, void* x, const pm_solver_options* options, const pm_solve_policy* policy,
                 pm_batch_report* reports, pm_workspace* ws, hipStream_t stream);

/* Batched FP64 iterative refinement (no reference counterpart): for every
 * system, r = b - A x with FP64 accumulation over FP64 x; systems above
 * target_relres (||r|| / ||b||) solve A d = r / ||r|| with the FP32 batched
 * PCG under `inner` (its relative tolerance, tightened per system as in
 * pm_pcg_refine, and iteration limit; absolute tolerance ignored) and take
 * x += ||r|| d, for at most max_outer steps.
 * a is FP32; rhs and x are FP64 and x is the initial guess.  Synchronises
 * `stream` once per outer step. */
typedef struct pm_batch_refine_report {
    int32_t converged;          /* final relres <= target */
    int32_t outer_iterations;   /* outer steps taken */
    int64_t inner_iterations;   /* FP32 PCG iterations over all outer steps */
    double initial_relres;
    double final_relres;        /* FP64 ||b - A x|| / ||b|| at exit */
    int32_t last_inner_status;  /* status of the last inner solve (pm_batch_report codes), 0 if none */
    int32_t pad;
} pm_batch_refine_report;
int pm_batch_pcg_refine(const pm_csr* a, int64_t systems, const int64_t* row_begin, const int64_t* row_end,
                        const double* rhs, double* x, const pm_solver_options* inner, double target_relres,
                        int32_t max_outer, const pm_solve_policy* policy, pm_batch_refine_report* reports,
                        pm_workspace* ws, hipStream_t stream);

/* Batched PCG with a factored preconditioner M_s^-1 = G_s^T G_s per system (no
 * reference counterpart; it replaces one pm_pcg_factored call per system, the
 * shape of the reference's one-system-per-call solver,
 * iterative_solver.cu:305-340).  g is lower-triangular and gt its transpose,
 * both block-diagonal over the same row ranges as `a` (pm_fsai_factors of a
 * graph built on the batched matrix gives exactly that), with a's dtype and
 * index widths.  jacobi_systems (device, one byte per system, or null) selects
 * Jacobi-PCG for the systems flagged nonzero: their z = D^-1 r, as in
 * pm_batch_pcg with use_jacobi.  Per iteration: q = A p, t = G (r - alpha q)
 * with x += alpha p, then r -= alpha q and z = G^T t.  Same stop rules,
 * reports and errors as pm_batch_pcg; options->use_jacobi is ignored. */
int pm_batch_pcg_factored(const pm_csr* a, int64_t systems, const int64_t* row_begin, const int64_t* row_end,
                          const void* rhs, void* x, const pm_csr* g, const pm_csr* gt, const uint8_t* jacobi_systems,
                          const pm_solver_options* options, const pm_solve_policy* policy, pm_batch_report* reports,
                          pm_workspace* ws, hipStream_t stream);
/* pm_batch_pcg_refine with the factored preconditioner (and its per-system
 * Jacobi flags) in the FP32 inner batched solves. */
int pm_batch_pcg_refine_factored(const pm_csr* a, int64_t systems, const int64_t* row_begin, const int64_t* row_end,
                                 const double* rhs, double* x, const pm_csr* g, const pm_csr* gt,
                                 const uint8_t* jacobi_systems, const pm_solver_options* inner, double target_relres,
                                 int32_t max_outer, const pm_solve_policy* policy, pm_batch_refine_report* reports,
                                 pm_workspace* ws, hipStream_t stream);

/* Batched block-Jacobi PCG: blockPcg (iterative_solver.h:130-141) per system
 * in one batch.  block_sizes (host, one per system, 1 to 8) gives each
 * system's block size; its rows [row_begin + k bs, row_begin + (k + 1) bs)
 * form diagonal block k, whose inverse the device computes from A (FP64
 * Gauss-Jordan, rounded to A's dtype) at every solve; block size 1 is
 * Jacobi.  A system's size must be a multiple of its block size
 * (PM_INVALID_ARGUMENT otherwise).  A block with a pivot that is not finite
 * or not above 16 eps |block|max stops its system as a bad diagonal (status
 * and bad_row as pm_batch_pcg's Jacobi rules).  Same stop rules, reports and
 * errors as pm_batch_pcg; options->use_jacobi is ignored. */
int pm_batch_block_pcg(const pm_csr* a, int64_t systems, const int64_t* row_begin, const int64_t* row_end,
                       const int32_t* block_sizes, const void* rhs, void* x, const pm_solver_options* options,
                       const pm_solve_policy* policy, pm_batch_report* reports, pm_workspace* ws, hipStream_t stream);
/* pm_batch_pcg_refine with per-system block-Jacobi in the FP32 inner solves. */
int pm_batch_block_pcg_refine(const pm_csr* a, int64_t systems, const int64_t* row_begin, const int64_t* row_end,
                              const int32_t* block_sizes, const double* rhs, double* x,
                              const pm_solver_options* inner, double target_relres, int32_t max_outer,
                              const pm_solve_policy* policy, pm_batch_refine_report* reports, pm_workspace* ws,
                              hipStream_t stream);

/* ---- Factored preconditioner M^-1 = G^T G ----------------------------------------
 * PCG with z = G^T (G r) (no reference counterpart).  g is lower-triangular
 * CSR, gt its transpose; both need 32-bit index mirrors and a's dtype.  Same
 * stop rules, reports and errors as pm_pcg. */
int pm_pcg_factored(const pm_csr* a, const void* rhs, void* x, const pm_csr* g, const pm_csr* gt,
                    const pm_solver_options* options, const pm_solve_policy* policy, pm_solver_report* report,
                    pm_workspace* ws, hipStream_t stream);
/* pm_pcg_refine with the factored preconditioner in the FP32 inner solves. */
int pm_pcg_refine_factored(const pm_csr* a, const double* rhs, double* x, const pm_csr* g, const pm_csr* gt,
                           const pm_solver_options* inner, double target_relres, int32_t max_outer,
                           const pm_solve_policy* policy, pm_refine_report* report, pm_workspace* ws,
                           hipStream_t stream);

/* ---- Learned factorized sparse approximate inverse -------------------------------
 * A 3-layer message-passing network (hidden 64, BF16 weights, FP32
 * accumulation, MFMA dense layers) over the graph of an SPD matrix emits G on
 * the lower pattern of A with a positive diagonal (M^-1 = G^T G is SPD by
 * construction); trained on the GPU to minimise
 * (1/P) sum_p ||(I - G A G^T) z_p||^2 / ||z_p||^2 over Gaussian probes.
 * No reference counterpart.  A freshly created network starts from
 * pm_fsai_initial_parameters(seed): random hidden layers and a decoder whose
 * G is exactly D^-1/2, i.e. the Jacobi preconditioner. */
typedef struct pm_fsai_net pm_fsai_net;
typedef struct pm_fsai_graph pm_fsai_graph;
int32_t pm_fsai_parameter_count(void);
/* Host initial parameters (count = pm_fsai_parameter_count()). */
int pm_fsai_initial_parameters(uint64_t seed, float* parameters, int64_t count);
int pm_fsai_net_create(uint64_t seed, pm_fsai_net** net);
int pm_fsai_net_destroy(pm_fsai_net* net);
/* Host FP32 parameters in (Adam state reset) / out; count must equal
 * pm_fsai_parameter_count(). */
int pm_fsai_net_set_parameters(pm_fsai_net* net, const float* parameters, int64_t count);
int pm_fsai_net_get_parameters(const pm_fsai_net* net, float* parameters, int64_t count);
/* Graph of an SPD matrix (square, 32-bit index mirrors, positive diagonal):
 * node features, normalised couplings, lower and upper patterns, storage for
 * G and the activations.  The matrix must outlive the graph. */
int pm_fsai_graph_create(const pm_csr* a, pm_fsai_graph** graph, hipStream_t stream);
/* Waits for every stream of the graph's device (forward, training and
 * factored solves may use the graph on any stream), then keeps its device
 * buffers for the next graph created on the same device: at most 4 sets of at
 * most 1 GiB each, held until pm_fsai_graph_pool_release or the end of the
 * process; larger or surplus buffers are freed.  G / G^T descriptors from
 * pm_fsai_factors are invalid once their graph is destroyed.  The environment
 * variable PLAMATRIX_AMD_GRAPH_POOL=0 frees every destroyed graph's storage
 * instead. */
int pm_fsai_graph_destroy(pm_fsai_graph* graph);
/* Frees the pooled storage of destroyed graphs on the current device. */
int pm_fsai_graph_pool_release(void);
/* Bytes the pool holds on the current device. */
int pm_fsai_graph_pool_bytes(int64_t* bytes);
/* Evaluate the network: G and G^T values in the graph. */
int pm_fsai_forward(pm_fsai_net* net, pm_fsai_graph* graph, hipStream_t stream);
/* CSR descriptors of G (lower) and G^T (upper), owned by the graph. */
int pm_fsai_factors(const pm_fsai_graph* graph, pm_csr* g, pm_csr* gt);
/* Per-system choice between the learned factor and Jacobi for a graph built on
 * a batched block-diagonal matrix (no reference counterpart).  With `probes`
 * (1..8) Gaussian probes z drawn from `seed`, system s keeps G when
 * sum_p ||(I - G A G^T) z_p||^2 over its rows [row_begin[s], row_end[s]) is
 * below the same loss of Jacobi's factor D^-1/2, else jacobi_systems[s] = 1
 * (device array of `systems` bytes, the input of pm_batch_pcg_factored).
 * losses (host, 2 * systems: learned, Jacobi per system) may be null.  Needs a
 * forward pass first; synchronises `stream`. */
int pm_fsai_batch_select(pm_fsai_graph* graph, int64_t systems, const int64_t* row_begin, const int64_t* row_end,
                         int32_t probes, uint64_t seed, uint8_t* jacobi_systems, double* losses, hipStream_t stream);
/* Probe loss and full parameter gradient (host) for `probes` (1..8) probes
 * drawn from `seed`; synchronises `stream`. */
int pm_fsai_loss_grad(pm_fsai_net* net, pm_fsai_graph* graph, int32_t probes, uint64_t seed, double* loss,
                      float* gradient, hipStream_t stream);
/* One Adam step (lr) on that loss; *loss (if not null) is the loss before the
 * step and synchronises `stream`. */
int pm_fsai_train_step(pm_fsai_net* net, pm_fsai_graph* graph, int32_t probes, uint64_t seed, float lr,
                       double* loss, hipStream_t stream);

/* ---- Damped Schur complement (FP64) --------------------------------------------
 * Block normal equations in HBM, the device form of BlockNormalEquations<double>
 * (block_schur.h:197-319): B (pc blocks ps x ps), C (ec blocks es x es),
 * gradients, cross blocks W (ps x es, row-major, creation order) with their
 * (primary, eliminated) ids, and primary cross blocks (row < col).  Block
 * sizes up to 16. */
typedef struct pm_normal_equations {
    int64_t primary_count;
    int64_t eliminated_count;
    int64_t primary_size;
    int64_t eliminated_size;
    const double* primary_diagonal;
    const double* eliminated_diagonal;
    const double* primary_gradient;
    const double* eliminated_gradient;
    int64_t cross_count;
    const int64_t* cross_primary;
    const int64_t* cross_eliminated;
    const double* cross_values;
    int64_t primary_cross_count;
    const int64_t* primary_cross_rows;
    const int64_t* primary_cross_cols;
    const double* primary_cross_values;
} pm_normal_equations;

/* Residuals that each link one primary and one eliminated block
 * (addResidualBlock, block_schur.h:223-262), row-major Jacobians
 * residual_size x block size.  weights == NULL takes the Huber IRLS weight
 * evaluateHuberLoss(||r||^2, huber_delta).weight (robust_loss.h:27-51). */
typedef struct pm_residual_batch {
    int64_t count;
    int64_t residual_size;
    const int64_t* primary;
    const int64_t* eliminated;
    const double* primary_jacobians;
    const double* eliminated_jacobians;
    const double* residuals;
    const double* weights;
    double huber_delta;
} pm_residual_batch;

/* SchurComplementSolverOptions<double> (block_schur.h:31-48) for the device
 * backends.  linear_backend selects the reduced solve: PM_SCHUR_BLOCK_PCG
 * (the accelerated backend: block-Jacobi PCG on the Schur CSR) or
 * PM_SCHUR_DENSE_CHOLESKY (the DenseCpu backend, block_schur.cpp:329-359 and
 * block_schur_dense_solver.h:300-398: dense FP64 Cholesky of the lower
 * triangle, one "iteration", converged when ||r|| <= max(abs + rel ||b||,
 * 100 eps max(1, ||b||)); use_initial_guess and max_iterations are ignored). */
enum { PM_SCHUR_BLOCK_PCG = 0, PM_SCHUR_DENSE_CHOLESKY = 1 };
/* flags of pm_schur_options (block-PCG backend):
 * PM_SCHUR_TOLERANCE_SUM: stop at ||r|| <= abs + rel ||r0||, the reference's
 *   Cpu backend rule (block_schur_cpu_solver.h:34-35), instead of
 *   max(abs, rel ||r0||);
 * PM_SCHUR_MIXED_PRECISION: useMixedPrecision (block_schur.h:43-44,
 *   block_schur_accelerated.cpp:131-193): an FP32 block-PCG pass on the
 *   rounded system (at most 200 iterations, relative tolerance at least 1e-3,
 *   absolute at least 1e-7) seeds the FP64 solve when it converges. */
enum { PM_SCHUR_TOLERANCE_SUM = 1, PM_SCHUR_MIXED_PRECISION = 2 };
typedef struct pm_schur_options {
    int32_t max_iterations;
    int32_t use_initial_guess;
    double relative_tolerance;
    double absolute_tolerance;
    double minimum_diagonal;
    int32_t linear_backend;
    int32_t flags;
} pm_schur_options;

/* SchurComplementSolverReport (block_schur.h:51-80).  status: 0 converged,
 * 1 eliminated block not positive definite, 2 preconditioner block not
 * positive definite, 4 iteration limit, 8 reduced Schur complement not
 * positive definite (dense backend), 16 dense Cholesky residual above the
 * tolerance. */
typedef struct pm_schur_report {
    int32_t converged;
    int32_t iterations;
    double initial_residual_norm;
    double final_residual_norm;
    int32_t status;
    int32_t pattern_reused;
    int32_t assembly_on_device;
    int32_t mixed_precision_used; /* the FP32 seed was converged and used */
    int64_t pattern_build_count;
    char message[96];
} pm_schur_report;

/* SchurComplementSolverWorkspace<double> (block_schur.h:83-169): the cached
 * device pattern, normal-equation grouping and per-solve numerics.  Binds to
 * its first stream like pm_workspace. */
typedef struct pm_schur_workspace pm_schur_workspace;
int pm_schur_workspace_create(pm_schur_workspace** ws);
int pm_schur_workspace_destroy(pm_schur_workspace* ws);
/* patternBuildCount() (block_schur.h:124-127). */
int64_t pm_schur_pattern_build_count(const pm_schur_workspace* ws);
int pm_schur_pattern_size(const pm_schur_workspace* ws, int64_t* slots, int64_t* values, int64_t* terms);

/* Device assembly of the normal equations of a residual batch
 * (BlockNormalEquations::addResidualBlocks, block_normal_equations.cpp:
 * 135-312, for one primary and one eliminated block per residual).  Output
 * arrays live in the workspace until the next call.  Synchronises `stream`
 * once to report invalid ids or values like the reference's exceptions. */
int pm_normal_equations_assemble(pm_schur_workspace* ws, int64_t primary_count, int64_t eliminated_count,
                                 int64_t primary_size, int64_t eliminated_size, const pm_residual_batch* batch,
                                 pm_normal_equations* out, hipStream_t stream);

/* Residuals that each link k distinct primary blocks and no eliminated block
 * (addPrimaryTerms, block_normal_equations.cpp:135-222: pose priors,
 * relative-pose or other primary-only constraints).  primary: count x k ids;
 * jacobians: count x k row-major residual_size x primary_size blocks (block
 * a of residual i at (i k + a) residual_size primary_size); weights == NULL
 * takes the Huber IRLS weight of ||r||^2 as pm_residual_batch does. */
typedef struct pm_primary_residual_batch {
    int64_t count;
    int64_t residual_size;
    int64_t blocks_per_residual;
    const int64_t* primary;
    const double* jacobians;
    const double* residuals;
    const double* weights;
    double huber_delta;
} pm_primary_residual_batch;

/* Device assembly of a primary-only residual batch on top of `in` (for
 * example the output of pm_normal_equations_assemble): out holds in's arrays
 * with the primary diagonal blocks, primary gradient and primary cross blocks
 * replaced by workspace arrays that add the batch's terms, bit for bit as
 * addPrimaryTerms called for every residual in order after the equations in
 * `in` were accumulated (primary cross blocks in their order of creation,
 * in's first).  in is not modified and may equal out; out's new arrays live
 * in the workspace until the next call.  Invalid ids or values fail like the
 * reference's exceptions (the lowest offending residual decides) and leave
 * out unchanged.  Synchronises `stream`. */
int pm_normal_equations_add_primary(pm_schur_workspace* ws, const pm_normal_equations* in,
                                    const pm_primary_residual_batch* batch, pm_normal_equations* out,
                                    hipStream_t stream);

/* solveDampedSchurComplement(equations, damping, options, workspace,
 * primary_step, eliminated_step) with the device backend
 * (block_schur.cpp:53-459): damping, eliminated inverses, reduced rhs, Schur
 * CSR on the cached pattern, block-Jacobi PCG with tolerance
 * max(abs, rel * ||r0||) (block_schur_accelerated.cpp:53-65) and
 * back-substitution.  primary_step / eliminated_step are device vectors. */
int pm_schur_solve(pm_schur_workspace* ws, const pm_normal_equations* equations, double damping,
                   const pm_schur_options* options, double* primary_step, double* eliminated_step,
                   pm_schur_report* report, pm_workspace* solver_ws, hipStream_t stream);

/* Assembly only (assembleReducedSchurCsr, block_schur_sparse_assembly.h:
 * 679-911): pattern (built or reused), damping, eliminated inverses, reduced
 * rhs, Schur values and inverted diagonal blocks, read with pm_schur_system.
 * report->status 1 / 2 as for pm_schur_solve.  Synchronises `stream`. */
int pm_schur_assemble(pm_schur_workspace* ws, const pm_normal_equations* equations, double damping,
                      double minimum_diagonal, pm_schur_report* report, hipStream_t stream);

/* The reduced system of the last assembly or solve: scalar CSR (FP64, int64 and int32
 * indices), right-hand side and inverted diagonal blocks (device pointers
 * owned by the workspace). */
int pm_schur_system(const pm_schur_workspace* ws, pm_csr* reduced, const double** rhs,
                    const double** preconditioner);

/* ---- dense direct solve (no reference counterpart) -------------------------
 * The dense Cholesky solve behind PM_SCHUR_DENSE_CHOLESKY, for callers that
 * hold a symmetric positive definite FP64 CSR matrix: the lower triangle
 * (duplicates summed) is factored in a rows x rows dense array in the
 * workspace (rows <= 65536), then b is solved into x (distinct device
 * vectors).  report->not_positive_definite is set when a pivot is not positive
 * and finite (x is then meaningless); residual_norm is ||b - A x|| with the
 * CSR values.  Synchronises `stream`. */
typedef struct pm_dense_report {
    int32_t not_positive_definite;
    int32_t pad;
    double rhs_norm;
    double residual_norm;
} pm_dense_report;
int pm_dense_cholesky_solve(const pm_csr* a, const double* b, double* x, pm_dense_report* report, pm_workspace* ws,
                            hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif
