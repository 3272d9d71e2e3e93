/*
 * Oracle (test infrastructure only): CPU restatement of PlaMatrix's sparse
 * CPU path.  See oracle.h for the rules that govern this directory.
 *
 * Follows:
 *   src/sparse/sparse_ops_cpu.cpp:24-134   combineCoo / writeCombinedCoo
 *   src/sparse/sparse_ops_cpu.cpp:214-233  spmv
 *   src/sparse/sparse_ops_cpu.cpp:244-268  spmm
 *   src/sparse/iterative_solver_cpu.cpp:15-284  validateOptions, validateSystem,
 *                                          dot, multiply, jacobiInverse, solve
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static _Thread_local char g_error[512];

const char* or_last_error(void) { return g_error; }

int or_fail(int code, const char* fmt, ...)
{
    va_list args;
    va_start(args, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, args);
    va_end(args);
    return code;
}

/* ------------------------------------------------------------------------ */
/* Stable sort of a permutation by (row, col): the reference uses
 * std::stable_sort over an iota permutation (sparse_ops_cpu.cpp:58-68).  A
 * bottom-up merge sort is stable, so equal keys keep input order, which is
 * all the fold below depends on. */
static void stable_sort_by_row_col(int64_t count, const int64_t* rows,
                                   const int64_t* cols, int64_t* perm)
{
    int64_t* buffer = (int64_t*)malloc((size_t)(count > 0 ? count : 1) * sizeof(int64_t));
    int64_t* src = perm;
    int64_t* dst = buffer;
    for (int64_t width = 1; width < count; width *= 2)
    {
        for (int64_t lo = 0; lo < count; lo += 2 * width)
        {
            int64_t mid = lo + width < count ? lo + width : count;
            int64_t hi = lo + 2 * width < count ? lo + 2 * width : count;
            int64_t i = lo, j = mid, k = lo;
            while (i < mid && j < hi)
            {
                const int64_t a = src[i], b = src[j];
                /* take from the right run only when strictly smaller */
                const int right_smaller = rows[b] < rows[a]
                    || (rows[b] == rows[a] && cols[b] < cols[a]);
                dst[k++] = right_smaller ? src[j++] : src[i++];
            }
            while (i < mid) dst[k++] = src[i++];
            while (j < hi) dst[k++] = src[j++];
        }
        int64_t* t = src;
        src = dst;
        dst = t;
    }
    if (src != perm) memcpy(perm, src, (size_t)count * sizeof(int64_t));
    free(buffer);
}

#define DEFINE_COO_TO_CSR(SUFFIX, Scalar)                                                  \
    int or_coo_to_csr_##SUFFIX(int64_t rows, int64_t cols, int64_t count,                  \
                               const int64_t* row_indices, const int64_t* col_indices,    \
                               const Scalar* values, int64_t* row_offsets,                \
                               int64_t* col_out, Scalar* val_out, int64_t* nnz)           \
    {                                                                                      \
        /* combineCoo: sparse_ops_cpu.cpp:31-56 */                                         \
        if (rows < 0 || cols < 0)                                                          \
            return or_fail(OR_INVALID_ARGUMENT, "cooToCsr dimensions must be non-negative"); \
        if (count < 0)                                                                     \
            return or_fail(OR_INVALID_ARGUMENT,                                            \
                           "cooToCsr row, column, and value arrays must have equal lengths"); \
        for (int64_t position = 0; position < count; ++position)                           \
        {                                                                                  \
            const int64_t row = row_indices[position];                                     \
            const int64_t col = col_indices[position];                                     \
            if (row < 0 || row >= rows || col < 0 || col >= cols)                          \
                return or_fail(OR_OUT_OF_RANGE,                                            \
                               "cooToCsr coordinate out of range at triplet %lld: (%lld, %lld) " \
                               "for matrix %lldx%lld",                                     \
                               (long long)position, (long long)row, (long long)col,        \
                               (long long)rows, (long long)cols);                          \
        }                                                                                  \
        int64_t* perm = (int64_t*)malloc((size_t)(count > 0 ? count : 1) * sizeof(int64_t)); \
        for (int64_t i = 0; i < count; ++i) perm[i] = i;                                   \
        stable_sort_by_row_col(count, row_indices, col_indices, perm);                     \
        /* left fold in sorted (= input) order from the first value: :75-91 */             \
        int64_t combined = 0;                                                              \
        int64_t last_row = -1, last_col = -1;                                              \
        for (int64_t i = 0; i < rows + 1; ++i) row_offsets[i] = 0;                         \
        for (int64_t k = 0; k < count; ++k)                                                \
        {                                                                                  \
            const int64_t source = perm[k];                                                \
            const int64_t row = row_indices[source];                                       \
            const int64_t col = col_indices[source];                                       \
            if (combined > 0 && last_row == row && last_col == col)                        \
            {                                                                              \
                val_out[combined - 1] += values[source];                                   \
            }                                                                              \
            else                                                                           \
            {                                                                              \
                col_out[combined] = col;                                                   \
                val_out[combined] = values[source];                                        \
                ++row_offsets[row + 1];                                                    \
                ++combined;                                                                \
                last_row = row;                                                            \
                last_col = col;                                                            \
            }                                                                              \
        }                                                                                  \
        /* counting prefix: writeCombinedCoo :118-126 */                                   \
        for (int64_t row = 0; row < rows; ++row) row_offsets[row + 1] += row_offsets[row]; \
        *nnz = combined;                                                                   \
        free(perm);                                                                        \
        return OR_OK;                                                                      \
    }

DEFINE_COO_TO_CSR(f64, double)
DEFINE_COO_TO_CSR(f32, float)

/* ------------------------------------------------------------------------ */
/* spmv: sparse_ops_cpu.cpp:214-233.  `sum += v*x` in Scalar precision.  The
 * file is compiled with -ffp-contract=off so no FMA is formed, as in the
 * reference's default x86-64 build. */
#define DEFINE_SPMV(SUFFIX, Scalar)                                                        \
    int or_spmv_##SUFFIX(int64_t rows, const int64_t* row_offsets,                         \
                         const int64_t* col_indices, const Scalar* values,                 \
                         const Scalar* x, Scalar* y)                                       \
    {                                                                                      \
        if (x == y && x != NULL)                                                           \
            return or_fail(OR_INVALID_ARGUMENT, "spmv input and output data must not alias"); \
        for (int64_t row = 0; row < rows; ++row)                                           \
        {                                                                                  \
            Scalar sum = (Scalar)0;                                                        \
            for (int64_t p = row_offsets[row]; p < row_offsets[row + 1]; ++p)              \
                sum += values[p] * x[col_indices[p]];                                      \
            y[row] = sum;                                                                  \
        }                                                                                  \
        return OR_OK;                                                                      \
    }

DEFINE_SPMV(f64, double)
DEFINE_SPMV(f32, float)

/* ------------------------------------------------------------------------ */
/* spmm: sparse_ops_cpu.cpp:244-268.  Column by column, `sum += v*b` in
 * Scalar precision, B (cols x k) and Y (rows x k) column-major. */
#define DEFINE_SPMM(SUFFIX, Scalar)                                                        \
    int or_spmm_##SUFFIX(int64_t rows, int64_t cols, int64_t k, const int64_t* row_offsets, \
                         const int64_t* col_indices, const Scalar* values,                 \
                         const Scalar* b, Scalar* y)                                       \
    {                                                                                      \
        if (b == y && b != NULL)                                                           \
            return or_fail(OR_INVALID_ARGUMENT, "spmm input and output data must not alias"); \
        for (int64_t col = 0; col < k; ++col)                                              \
        {                                                                                  \
            const Scalar* bc = b + col * cols;                                             \
            Scalar* yc = y + col * rows;                                                   \
            for (int64_t row = 0; row < rows; ++row)                                       \
            {                                                                              \
                Scalar sum = (Scalar)0;                                                    \
                for (int64_t p = row_offsets[row]; p < row_offsets[row + 1]; ++p)          \
                    sum += values[p] * bc[col_indices[p]];                                 \
                yc[row] = sum;                                                             \
            }                                                                              \
        }                                                                                  \
        return OR_OK;                                                                      \
    }

DEFINE_SPMM(f64, double)
DEFINE_SPMM(f32, float)

/* ------------------------------------------------------------------------ */
/* iterative_solver_cpu.cpp:15-29 */
static int validate_options(const or_options* o)
{
    if (o->max_iterations < 0)
        return or_fail(OR_INVALID_ARGUMENT, "iterative solver maxIterations must be non-negative");
    if (!isfinite(o->relative_tolerance) || o->relative_tolerance < 0.0
        || o->relative_tolerance > 1.0 || !isfinite(o->absolute_tolerance)
        || o->absolute_tolerance < 0.0)
        return or_fail(OR_INVALID_ARGUMENT,
                       "iterative solver relative tolerance must be in [0, 1] and absolute "
                       "tolerance must be finite and non-negative");
    return OR_OK;
}

/* iterative_solver_cpu.cpp:99-107 */
static double dot(int64_t n, const double* a, const double* b)
{
    double result = 0.0;
    for (int64_t i = 0; i < n; ++i) result += a[i] * b[i];
    return result;
}

#define DEFINE_SOLVER_HELPERS(SUFFIX, Scalar, EPS)                                         \
    /* validateSystem: iterative_solver_cpu.cpp:31-97 (shape checks are the caller's) */   \
    static int validate_system_##SUFFIX(int64_t rows, int64_t cols, int64_t nnz,           \
                                        const int64_t* ro, const int64_t* ci,              \
                                        const Scalar* v, const Scalar* rhs,                \
                                        const Scalar* x)                                   \
    {                                                                                      \
        if (rows != cols)                                                                  \
            return or_fail(OR_INVALID_ARGUMENT, "iterative solver matrix must be square"); \
        if (rhs != NULL && rhs == x)                                                       \
            return or_fail(OR_INVALID_ARGUMENT,                                            \
                           "iterative solver rhs and solution must not alias");            \
        if (ro[0] != 0 || ro[rows] != nnz)                                                 \
            return or_fail(OR_INVALID_ARGUMENT,                                            \
                           "iterative solver requires CSR row offsets to start at zero "   \
                           "and end at nnz");                                              \
        for (int64_t row = 0; row < rows; ++row)                                           \
        {                                                                                  \
            if (ro[row] < 0 || ro[row + 1] < ro[row] || ro[row + 1] > nnz)                 \
                return or_fail(OR_INVALID_ARGUMENT,                                        \
                               "iterative solver invalid CSR row offsets at row %lld",     \
                               (long long)row);                                            \
        }                                                                                  \
        for (int64_t p = 0; p < nnz; ++p)                                                  \
        {                                                                                  \
            if (ci[p] < 0 || ci[p] >= cols)                                                \
                return or_fail(OR_INVALID_ARGUMENT,                                        \
                               "iterative solver CSR column index out of range at "        \
                               "position %lld", (long long)p);                             \
            if (!isfinite((double)v[p]))                                                   \
                return or_fail(OR_INVALID_ARGUMENT,                                        \
                               "iterative solver CSR value is not finite at position %lld", \
                               (long long)p);                                              \
        }                                                                                  \
        for (int64_t row = 0; row < rows; ++row)                                           \
        {                                                                                  \
            if (!isfinite((double)rhs[row]) || !isfinite((double)x[row]))                  \
                return or_fail(OR_INVALID_ARGUMENT,                                        \
                               "iterative solver rhs and solution must be finite at row %lld", \
                               (long long)row);                                            \
        }                                                                                  \
        return OR_OK;                                                                      \
    }                                                                                      \
    /* multiply: iterative_solver_cpu.cpp:109-125 (double accumulation) */                 \
    static void multiply_##SUFFIX(int64_t rows, const int64_t* ro, const int64_t* ci,      \
                                  const Scalar* v, const double* in, double* out)          \
    {                                                                                      \
        for (int64_t row = 0; row < rows; ++row)                                           \
        {                                                                                  \
            double sum = 0.0;                                                              \
            for (int64_t p = ro[row]; p < ro[row + 1]; ++p)                                \
                sum += (double)v[p] * in[ci[p]];                                           \
            out[row] = sum;                                                                \
        }                                                                                  \
    }                                                                                      \
    /* jacobiInverse: iterative_solver_cpu.cpp:127-160 */                                  \
    static int jacobi_inverse_##SUFFIX(int64_t rows, const int64_t* ro, const int64_t* ci, \
                                       const Scalar* v, double* inverse)                   \
    {                                                                                      \
        for (int64_t row = 0; row < rows; ++row)                                           \
        {                                                                                  \
            int found = 0;                                                                 \
            double diagonal = 0.0, row_scale = 0.0;                                        \
            for (int64_t p = ro[row]; p < ro[row + 1]; ++p)                                \
            {                                                                              \
                const double value = (double)v[p];                                         \
                row_scale = fmax(row_scale, fabs(value));                                  \
                if (ci[p] == row)                                                          \
                {                                                                          \
                    diagonal += value;                                                     \
                    found = 1;                                                             \
                }                                                                          \
            }                                                                              \
            const double threshold = (double)(EPS) * row_scale * 16.0;                     \
            if (!found || !isfinite(diagonal) || diagonal <= threshold                     \
                || !isfinite(1.0 / diagonal))                                              \
                return or_fail(OR_RUNTIME,                                                 \
                               "pcg Jacobi diagonal is missing, non-positive, or near "    \
                               "zero at row %lld", (long long)row);                        \
            inverse[row] = 1.0 / diagonal;                                                 \
        }                                                                                  \
        return OR_OK;                                                                      \
    }

DEFINE_SOLVER_HELPERS(f64, double, DBL_EPSILON)
DEFINE_SOLVER_HELPERS(f32, float, FLT_EPSILON)

/* The PCG loop of iterative_solver_cpu.cpp:162-284.  `apply` selects the
 * preconditioner: 0 none, 1 Jacobi (diagonal), 2 block-Jacobi (inverse blocks
 * applied in FP64, the GPU kernel's sum order
 * iterative_solver_kernels.cuh:156-166). */
#define DEFINE_SOLVE(SUFFIX, Scalar)                                                       \
    static void apply_##SUFFIX(int64_t n, int mode, const double* inverse_diagonal,        \
                               const Scalar* inverse_blocks, int64_t b,                    \
                               const double* residual, double* transformed)                \
    {                                                                                      \
        if (mode == 1)                                                                     \
        {                                                                                  \
            for (int64_t i = 0; i < n; ++i) transformed[i] = inverse_diagonal[i] * residual[i]; \
        }                                                                                  \
        else if (mode == 2)                                                                \
        {                                                                                  \
            for (int64_t row = 0; row < n; ++row)                                          \
            {                                                                              \
                const int64_t block = row / b;                                             \
                const int64_t local = row - block * b;                                     \
                const Scalar* m = inverse_blocks + block * b * b + local * b;              \
                double value = 0.0;                                                        \
                for (int64_t c = 0; c < b; ++c) value += (double)m[c] * residual[block * b + c]; \
                transformed[row] = value;                                                  \
            }                                                                              \
        }                                                                                  \
        else                                                                               \
        {                                                                                  \
            for (int64_t i = 0; i < n; ++i) transformed[i] = residual[i];                  \
        }                                                                                  \
    }                                                                                      \
    static int solve_loop_##SUFFIX(int64_t rows, const int64_t* ro, const int64_t* ci,     \
                                   const Scalar* v, const Scalar* rhs, Scalar* solution,   \
                                   const or_options* options, int mode,                    \
                                   const Scalar* inverse_blocks, int64_t b,                \
                                   or_report* report)                                      \
    {                                                                                      \
        const int64_t n = rows;                                                            \
        const size_t bytes = (size_t)(n > 0 ? n : 1) * sizeof(double);                    \
        double* x = (double*)malloc(bytes);                                                \
        double* residual = (double*)malloc(bytes);                                         \
        double* direction = (double*)malloc(bytes);                                        \
        double* transformed = (double*)malloc(bytes);                                      \
        double* matrix_direction = (double*)malloc(bytes);                                 \
        double* inverse_diagonal = NULL;                                                   \
        int status = OR_OK;                                                                \
        memset(report, 0, sizeof(*report));                                                \
        for (int64_t i = 0; i < n; ++i) x[i] = (double)solution[i];                        \
        multiply_##SUFFIX(rows, ro, ci, v, x, matrix_direction);                           \
        for (int64_t i = 0; i < n; ++i) residual[i] = (double)rhs[i] - matrix_direction[i]; \
        const double initial_squared = dot(n, residual, residual);                         \
        if (!isfinite(initial_squared) || initial_squared < 0.0)                           \
        {                                                                                  \
            status = or_fail(OR_RUNTIME, "iterative solver initial residual is not finite"); \
            goto done;                                                                     \
        }                                                                                  \
        report->initial_residual = sqrt(initial_squared);                                  \
        report->final_residual = report->initial_residual;                                 \
        const double tolerance = fmax(options->absolute_tolerance,                         \
                                      options->relative_tolerance * report->initial_residual); \
        if (report->final_residual <= tolerance)                                           \
        {                                                                                  \
            report->converged = 1;                                                         \
            goto done;                                                                     \
        }                                                                                  \
        if (mode == 1)                                                                     \
        {                                                                                  \
            inverse_diagonal = (double*)malloc(bytes);                                     \
            status = jacobi_inverse_##SUFFIX(rows, ro, ci, v, inverse_diagonal);           \
            if (status != OR_OK) goto done;                                                \
        }                                                                                  \
        apply_##SUFFIX(n, mode, inverse_diagonal, inverse_blocks, b, residual, transformed); \
        for (int64_t i = 0; i < n; ++i) direction[i] = transformed[i];                     \
        double rho = dot(n, residual, transformed);                                        \
        if (!isfinite(rho) || rho <= 0.0)                                                  \
        {                                                                                  \
            status = or_fail(OR_RUNTIME,                                                   \
                             "iterative solver breakdown: preconditioned residual is not positive"); \
            goto done;                                                                     \
        }                                                                                  \
        for (int iteration = 0; iteration < options->max_iterations; ++iteration)          \
        {                                                                                  \
            multiply_##SUFFIX(rows, ro, ci, v, direction, matrix_direction);               \
            const double denominator = dot(n, direction, matrix_direction);                \
            if (!isfinite(denominator) || denominator <= 0.0)                              \
            {                                                                              \
                status = or_fail(OR_RUNTIME,                                               \
                                 "iterative solver breakdown: matrix is not numerically SPD"); \
                goto done;                                                                 \
            }                                                                              \
            const double alpha = rho / denominator;                                        \
            if (!isfinite(alpha))                                                          \
            {                                                                              \
                status = or_fail(OR_RUNTIME, "iterative solver breakdown: step size is not finite"); \
                goto done;                                                                 \
            }                                                                              \
            for (int64_t i = 0; i < n; ++i)                                                \
            {                                                                              \
                x[i] += alpha * direction[i];                                              \
                residual[i] -= alpha * matrix_direction[i];                                \
            }                                                                              \
            report->iterations = iteration + 1;                                            \
            const double residual_squared = dot(n, residual, residual);                    \
            if (!isfinite(residual_squared) || residual_squared < 0.0)                     \
            {                                                                              \
                status = or_fail(OR_RUNTIME, "iterative solver residual is not finite");   \
                goto done;                                                                 \
            }                                                                              \
            report->final_residual = sqrt(residual_squared);                               \
            if (report->final_residual <= tolerance)                                       \
            {                                                                              \
                report->converged = 1;                                                     \
                break;                                                                     \
            }                                                                              \
            apply_##SUFFIX(n, mode, inverse_diagonal, inverse_blocks, b, residual, transformed); \
            const double next_rho = dot(n, residual, transformed);                         \
            if (!isfinite(next_rho) || next_rho <= 0.0)                                    \
            {                                                                              \
                status = or_fail(OR_RUNTIME, "iterative solver breakdown in direction update"); \
                goto done;                                                                 \
            }                                                                              \
            const double beta = next_rho / rho;                                            \
            if (!isfinite(beta))                                                           \
            {                                                                              \
                status = or_fail(OR_RUNTIME,                                               \
                                 "iterative solver breakdown: direction scale is not finite"); \
                goto done;                                                                 \
            }                                                                              \
            for (int64_t i = 0; i < n; ++i) direction[i] = transformed[i] + beta * direction[i]; \
            rho = next_rho;                                                                \
        }                                                                                  \
        for (int64_t i = 0; i < n; ++i) solution[i] = (Scalar)x[i];                        \
        if (!report->converged && options->require_convergence)                            \
            status = or_fail(OR_RUNTIME,                                                   \
                             "iterative solver did not converge in %d iterations; final "  \
                             "residual=%g", report->iterations, report->final_residual);   \
    done:                                                                                  \
        free(x);                                                                           \
        free(residual);                                                                    \
        free(direction);                                                                   \
        free(transformed);                                                                 \
        free(matrix_direction);                                                            \
        free(inverse_diagonal);                                                            \
        return status;                                                                     \
    }                                                                                      \
    int or_solve_##SUFFIX(int64_t rows, int64_t cols, int64_t nnz,                         \
                          const int64_t* row_offsets, const int64_t* col_indices,          \
                          const Scalar* values, const Scalar* rhs, Scalar* x,              \
                          const or_options* options, int preconditioned,                   \
                          or_report* report)                                               \
    {                                                                                      \
        int status = validate_options(options);                                            \
        if (status != OR_OK) return status;                                                \
        status = validate_system_##SUFFIX(rows, cols, nnz, row_offsets, col_indices,       \
                                          values, rhs, x);                                 \
        if (status != OR_OK) return status;                                                \
        const int mode = (preconditioned && options->use_jacobi) ? 1 : 0;                  \
        return solve_loop_##SUFFIX(rows, row_offsets, col_indices, values, rhs, x,         \
                                   options, mode, NULL, 1, report);                        \
    }                                                                                      \
    int or_block_pcg_##SUFFIX(int64_t rows, int64_t nnz, const int64_t* row_offsets,       \
                              const int64_t* col_indices, const Scalar* values,            \
                              const Scalar* rhs, Scalar* x,                                \
                              const Scalar* inverse_blocks, int64_t block_size,            \
                              const or_options* options, or_report* report)                \
    {                                                                                      \
        /* validateBlockPreconditioner: iterative_solver.cu:244-252 */                     \
        if (block_size <= 0 || rows % block_size != 0)                                     \
            return or_fail(OR_INVALID_ARGUMENT,                                            \
                           "block PCG requires complete row-major inverse diagonal blocks"); \
        int status = validate_options(options);                                            \
        if (status != OR_OK) return status;                                                \
        status = validate_system_##SUFFIX(rows, rows, nnz, row_offsets, col_indices,       \
                                          values, rhs, x);                                 \
        if (status != OR_OK) return status;                                                \
        return solve_loop_##SUFFIX(rows, row_offsets, col_indices, values, rhs, x,         \
                                   options, 2, inverse_blocks, block_size, report);        \
    }

DEFINE_SOLVE(f64, double)
DEFINE_SOLVE(f32, float)

/* ------------------------------------------------------------------------ */
#define DEFINE_RELRES(NAME, Scalar, Vec)                                                   \
    double NAME(int64_t rows, const int64_t* ro, const int64_t* ci, const Scalar* v,       \
                const Vec* rhs, const Vec* x)                                              \
    {                                                                                      \
        double rr = 0.0, bb = 0.0;                                                         \
        for (int64_t row = 0; row < rows; ++row)                                           \
        {                                                                                  \
            double sum = 0.0;                                                              \
            for (int64_t p = ro[row]; p < ro[row + 1]; ++p)                                \
                sum += (double)v[p] * (double)x[ci[p]];                                    \
            const double r = (double)rhs[row] - sum;                                       \
            rr += r * r;                                                                   \
            bb += (double)rhs[row] * (double)rhs[row];                                     \
        }                                                                                  \
        return bb > 0.0 ? sqrt(rr / bb) : sqrt(rr);                                        \
    }

DEFINE_RELRES(or_true_relres_f64, double, double)
DEFINE_RELRES(or_true_relres_f32, float, float)
DEFINE_RELRES(or_true_relres_f32x64, float, double)

/* robust_loss.h:27-51 */
int or_huber_f64(double s, double delta, double* cost, double* weight)
{
    if (!isfinite(s) || s < 0.0)
        return or_fail(OR_INVALID_ARGUMENT,
                       "evaluateHuberLoss: squared residual norm must be finite and non-negative");
    if (!isfinite(delta) || delta < 0.0)
        return or_fail(OR_INVALID_ARGUMENT,
                       "evaluateHuberLoss: delta must be finite and non-negative");
    if (delta == 0.0 || s <= delta * delta)
    {
        *cost = 0.5 * s;
        *weight = 1.0;
        return OR_OK;
    }
    const double norm = sqrt(s);
    *cost = delta * norm - 0.5 * delta * delta;
    *weight = delta / norm;
    return OR_OK;
}
