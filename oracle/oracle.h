/*
 * Oracle: a plain-C restatement of PlaMatrix's CPU algorithms for the sparse
 * solve path (COO->CSR, CSR SpMV, CG / Jacobi-PCG, block-Jacobi PCG and the
 * damped Schur-complement solve).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (plamatrix_amd/, include/)
 * may link or call this code.  Only tests/, the smoke() check and the CPU
 * baseline leg of bench.py load it, and only as the checker or as the timed
 * CPU baseline, never as the thing measured for the GPU path.
 *
 * Every function cites the reference file:line it follows (paths relative to
 * the PlaMatrix 0.1.0 source tree).  The operation order of each loop is kept
 * as in the reference so that results agree bit for bit with the reference's
 * own CPU build on the same inputs (checked by tests/test_oracle_vs_ref.py
 * against oracle/_ref where that library was built).
 *
 * Pinning (which check pins which function):
 *   or_coo_to_csr_*   reference KATs (sparse_ops_test.cpp:22-45, :49-71,
 *                     :73-118, :276-293, :320-369, :371-430), brute-force
 *                     permutation check on tiny inputs, oracle/_ref
 *   or_spmv_*         KATs sparse_ops_test.cpp:120-193, row-sum closed form
 *   or_spmm_*         KATs sparse_ops_test.cpp:196-268, columns = SpMV, reference build
 *   or_solve_*        closed forms iterative_solver_test.cpp:103-335, 782-820,
 *                     scipy.sparse.linalg.spsolve, oracle/_ref
 *   or_block_pcg_*    2x2 block system in one iteration
 *                     (iterative_solver_test.cpp:78-101, 394-424), equivalence
 *                     with or_solve_* when block size is 1, scipy spsolve
 *   Schur functions   see schur.c header
 */
#ifndef PLAMATRIX_ORACLE_H
#define PLAMATRIX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes, mapped onto the reference's exception classes. */
enum {
    OR_OK = 0,
    OR_INVALID_ARGUMENT = 1, /* std::invalid_argument */
    OR_OUT_OF_RANGE = 2,     /* std::out_of_range */
    OR_OVERFLOW = 3,         /* std::overflow_error */
    OR_LOGIC = 4,            /* std::logic_error */
    OR_RUNTIME = 5           /* std::runtime_error */
};

const char* or_last_error(void);

/* ---- COO -> CSR (sparse_ops_cpu.cpp:24-134, 180-203) ------------------ */
/* Output buffers need room for `count` entries (the combined nnz is <= count)
 * and rows+1 offsets.  *nnz receives the combined count. */
int or_coo_to_csr_f64(int64_t rows, int64_t cols, int64_t count,
                      const int64_t* row_indices, const int64_t* col_indices,
                      const double* values, int64_t* row_offsets,
                      int64_t* col_out, double* val_out, int64_t* nnz);
int or_coo_to_csr_f32(int64_t rows, int64_t cols, int64_t count,
                      const int64_t* row_indices, const int64_t* col_indices,
                      const float* values, int64_t* row_offsets,
                      int64_t* col_out, float* val_out, int64_t* nnz);

/* ---- SpMV (sparse_ops_cpu.cpp:214-233): Scalar accumulation, left to right */
int or_spmv_f64(int64_t rows, const int64_t* row_offsets,
                const int64_t* col_indices, const double* values,
                const double* x, double* y);
int or_spmv_f32(int64_t rows, const int64_t* row_offsets,
                const int64_t* col_indices, const float* values,
                const float* x, float* y);

/* ---- SpMM (sparse_ops_cpu.cpp:244-268): per column as SpMV, column-major */
int or_spmm_f64(int64_t rows, int64_t cols, int64_t k, const int64_t* row_offsets,
                const int64_t* col_indices, const double* values, const double* b, double* y);
int or_spmm_f32(int64_t rows, int64_t cols, int64_t k, const int64_t* row_offsets,
                const int64_t* col_indices, const float* values, const float* b, float* y);

/* ---- CG / Jacobi-PCG (iterative_solver_cpu.cpp:15-304) ----------------- */
typedef struct or_options {
    int max_iterations;
    double relative_tolerance;
    double absolute_tolerance;
    int use_jacobi;
    int require_convergence;
} or_options;

typedef struct or_report {
    int converged;
    int iterations;
    double initial_residual;
    double final_residual;
} or_report;

/* preconditioned = 0 for cg(), 1 for pcg().  x is the initial guess on entry
 * and the solution (cast to Scalar) on exit. */
int or_solve_f64(int64_t rows, int64_t cols, int64_t nnz,
                 const int64_t* row_offsets, const int64_t* col_indices,
                 const double* values, const double* rhs, double* x,
                 const or_options* options, int preconditioned,
                 or_report* report);
int or_solve_f32(int64_t rows, int64_t cols, int64_t nnz,
                 const int64_t* row_offsets, const int64_t* col_indices,
                 const float* values, const float* rhs, float* x,
                 const or_options* options, int preconditioned,
                 or_report* report);

/* Block-Jacobi PCG: the reference's blockPcg (iterative_solver.cu:386-402,
 * applyBlockJacobiKernel iterative_solver_kernels.cuh:144-167) has no CPU
 * counterpart; this restates it in the loop shape and FP64 arithmetic of the
 * CPU solver.  inverse_blocks holds rows/b row-major b x b blocks. */
int or_block_pcg_f64(int64_t rows, int64_t nnz, const int64_t* row_offsets,
                     const int64_t* col_indices, const double* values,
                     const double* rhs, double* x,
                     const double* inverse_blocks, int64_t block_size,
                     const or_options* options, or_report* report);
int or_block_pcg_f32(int64_t rows, int64_t nnz, const int64_t* row_offsets,
                     const int64_t* col_indices, const float* values,
                     const float* rhs, float* x,
                     const float* inverse_blocks, int64_t block_size,
                     const or_options* options, or_report* report);

/* FP64 true relative residual ||b - A x|| / ||b|| (0 when b == 0 and x
 * solves exactly; ||b - A x|| when b == 0). */
double or_true_relres_f64(int64_t rows, const int64_t* row_offsets,
                          const int64_t* col_indices, const double* values,
                          const double* rhs, const double* x);
double or_true_relres_f32(int64_t rows, const int64_t* row_offsets,
                          const int64_t* col_indices, const float* values,
                          const float* rhs, const float* x);
double or_true_relres_f32x64(int64_t rows, const int64_t* row_offsets,
                             const int64_t* col_indices, const float* values,
                             const double* rhs, const double* x);

/* ---- Huber loss (optimization/robust_loss.h:27-51) ---------------------- */
int or_huber_f64(double squared_residual_norm, double delta, double* cost,
                 double* weight);

/* ---- Block normal equations + damped Schur solve (schur.c) ------------- */
typedef struct or_ne or_ne;

or_ne* or_ne_create(int64_t primary_count, int64_t eliminated_count,
                    int64_t primary_size, int64_t eliminated_size);
void or_ne_destroy(or_ne* eq);
/* addResidualBlocks (block_normal_equations.cpp:243-312) */
int or_ne_add_residual_blocks(or_ne* eq, int64_t primary_block_count,
                              const int64_t* primary_blocks,
                              const double* const* primary_jacobians,
                              int64_t eliminated_block,
                              const double* eliminated_jacobian,
                              const double* residual, int64_t residual_size,
                              double weight);
/* Many single-primary residuals at once: residual i links primary[i] and
 * eliminated[i]; jp is R x (rs*ps), je is R x (rs*es), r is R x rs.  Adds them
 * in index order, exactly as repeated addResidualBlock calls would. */
int or_ne_add_residual_batch(or_ne* eq, int64_t count, const int64_t* primary,
                             const int64_t* eliminated, const double* jp,
                             const double* je, const double* r,
                             int64_t residual_size, const double* weights);
/* addPrimaryResidualBlocks (block_normal_equations.cpp:326-335) */
int or_ne_add_primary_residual_blocks(or_ne* eq, int64_t primary_block_count,
                                      const int64_t* primary_blocks,
                                      const double* const* primary_jacobians,
                                      const double* residual,
                                      int64_t residual_size, double weight);
/* addEliminatedResidualBlock (block_normal_equations.cpp:337-373) */
int or_ne_add_eliminated_residual_block(or_ne* eq, int64_t eliminated_block,
                                        const double* eliminated_jacobian,
                                        const double* residual,
                                        int64_t residual_size, double weight);
/* mergeFrom (block_normal_equations.cpp:375-411) */
int or_ne_merge(or_ne* eq, const or_ne* other);

/* Accessors for the accumulated blocks (row-major, as the reference stores
 * them).  cross blocks are in creation order. */
int64_t or_ne_cross_count(const or_ne* eq);
int64_t or_ne_primary_cross_count(const or_ne* eq);
void or_ne_export(const or_ne* eq, double* primary_diagonal,
                  double* eliminated_diagonal, double* primary_gradient,
                  double* eliminated_gradient, int64_t* cross_primary,
                  int64_t* cross_eliminated, double* cross_values,
                  int64_t* primary_cross_rows, int64_t* primary_cross_cols,
                  double* primary_cross_values);

typedef struct or_schur_options {
    int max_iterations;
    double relative_tolerance;
    double absolute_tolerance;
    double minimum_diagonal;
    int use_initial_guess;
} or_schur_options;

typedef struct or_schur_report {
    int converged;
    int iterations;
    double initial_residual_norm;
    double final_residual_norm;
    int status; /* 0 ok, 1 eliminated block not PD, 2 preconditioner block not
                   PD, 3 Schur not PD, 4 iteration limit */
} or_schur_report;

/* solveDampedSchurComplement with the Cpu backend (block_schur.cpp:53-459,
 * block_schur_cpu_solver.h:14-82).  primary_step has primary_count*ps
 * entries (read as the initial guess when use_initial_guess), eliminated_step
 * eliminated_count*es entries. */
int or_schur_solve(const or_ne* eq, double damping,
                   const or_schur_options* options, double* primary_step,
                   double* eliminated_step, or_schur_report* report);

/* Damped diagonal blocks, eliminated inverses, reduced rhs and block-Jacobi
 * preconditioner inverses (block_schur.cpp:92-250). */
int or_schur_prologue(const or_ne* eq, double damping, double minimum_diagonal,
                      double* primary_diagonal_damped,
                      double* eliminated_inverse, double* reduced_rhs,
                      double* preconditioner_inverse);

/* Schur CSR pattern (block_schur_sparse_assembly.h:41-260): sizes first,
 * then the arrays. */
int or_schur_pattern_size(const or_ne* eq, int64_t* slot_count,
                          int64_t* value_count, int64_t* term_count);
int or_schur_pattern(const or_ne* eq, int64_t* slot_rows, int64_t* slot_cols,
                     int64_t* row_offsets, int64_t* col_indices,
                     int64_t* term_offsets, int64_t* term_eliminated,
                     int64_t* term_left, int64_t* term_right);
/* CPU accumulation of the Schur CSR values
 * (block_schur_sparse_assembly.h:817-903). */
int or_schur_csr_values(const or_ne* eq, double damping,
                        double minimum_diagonal, double* values);

#ifdef __cplusplus
}
#endif

#endif
