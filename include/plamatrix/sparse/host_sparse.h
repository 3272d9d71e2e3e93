// Host implementations behind the reference's CPU overloads: cooToCsr on
// std::vector triplets (sparse_ops.h:79-96, rule of sparse_ops_cpu.cpp:24-134),
// spmv / spmm on CSRMatrix<S, Device::CPU> (sparse_ops.h:122-146, rule of
// sparse_ops_cpu.cpp:205-268) and cg / pcg on CPU matrices
// (iterative_solver.h:94-108, rule of iterative_solver_cpu.cpp:99-304).
//
// These serve host matrices only; device matrices always take the HIP
// kernels.  Results follow the reference's CPU rules exactly:
//   - COO -> CSR orders triplets by (row, column), keeps equal coordinates in
//     input order and folds them left to right from the first value; explicit
//     zeros stay.  Here a counting pass buckets triplets by row (stable), a
//     stable sort orders each row by column, so the cost is O(T + sum r log r)
//     instead of one O(T log T) sort, with the identical result.
//   - SpMV / SpMM accumulate each row left to right in Scalar, products
//     rounded before the add (no fused multiply-add).
//   - CG / PCG iterate in double whatever Scalar is, with the stopping rule
//     ||r|| <= max(abs, rel ||r0||) on the recursive residual and the
//     reference's breakdown, diagonal and option errors.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "plamatrix/core/types.h"

namespace plamatrix
{
namespace detail
{
namespace host
{

// Products rounded to Scalar before the add, as in a build without FMA
// contraction (the reference's default x86-64 flags).
template <typename T>
inline T product(T a, T b)
{
    volatile T p = a * b;
    return p;
}

// Bucket, order and fold the triplets; values are folded by the caller so the
// index work is shared by every Scalar.
inline void orderTriplets(Index rows, Index cols, const std::vector<Index>& row_indices,
                          const std::vector<Index>& col_indices, std::size_t value_count,
                          std::vector<std::size_t>& order, std::vector<Index>& row_start)
{
    if (rows < 0 || cols < 0) throw std::invalid_argument("cooToCsr dimensions must be non-negative");
    if (row_indices.size() != col_indices.size() || row_indices.size() != value_count)
        throw std::invalid_argument("cooToCsr row, column, and value arrays must have equal lengths");
    if (row_indices.size() > static_cast<std::size_t>(std::numeric_limits<Index>::max()))
        throw std::overflow_error("cooToCsr triplet count exceeds Index range");
    const std::size_t count = row_indices.size();
    for (std::size_t k = 0; k < count; ++k)
    {
        const Index r = row_indices[k], c = col_indices[k];
        if (r < 0 || r >= rows || c < 0 || c >= cols)
            throw std::out_of_range("cooToCsr coordinate out of range at triplet " + std::to_string(k) + ": (" +
                                    std::to_string(r) + ", " + std::to_string(c) + ") for matrix " +
                                    std::to_string(rows) + "x" + std::to_string(cols));
    }
    row_start.assign(static_cast<std::size_t>(rows) + 1, 0);
    for (std::size_t k = 0; k < count; ++k) ++row_start[static_cast<std::size_t>(row_indices[k]) + 1];
    for (Index r = 0; r < rows; ++r) row_start[r + 1] += row_start[r];
    order.resize(count);
    std::vector<Index> cursor(row_start.begin(), row_start.end() - 1);
    for (std::size_t k = 0; k < count; ++k)
        order[static_cast<std::size_t>(cursor[static_cast<std::size_t>(row_indices[k])]++)] = k;
    for (Index r = 0; r < rows; ++r)
    {
        auto b = order.begin() + row_start[r], e = order.begin() + row_start[r + 1];
        if (e - b > 1)
            std::stable_sort(b, e, [&](std::size_t x, std::size_t y) { return col_indices[x] < col_indices[y]; });
    }
}

template <typename Scalar, typename Out>
void foldInto(Index rows, const std::vector<Index>& col_indices, const std::vector<Scalar>& values,
              const std::vector<std::size_t>& order, const std::vector<Index>& row_start, Out&& out)
{
    // out(combined_nnz, row_offsets, col_indices, values) sizes the output and
    // hands back its three arrays
    Index total = 0;
    for (Index r = 0; r < rows; ++r)
        for (Index k = row_start[r]; k < row_start[r + 1]; ++k)
            if (k == row_start[r] || col_indices[order[k]] != col_indices[order[k - 1]]) ++total;
    Index* ro;
    Index* ci;
    Scalar* v;
    out(total, ro, ci, v);
    ro[0] = 0;
    Index w = 0;
    for (Index r = 0; r < rows; ++r)
    {
        for (Index k = row_start[r]; k < row_start[r + 1]; ++k)
        {
            const std::size_t src = order[static_cast<std::size_t>(k)];
            if (k != row_start[r] && col_indices[src] == ci[w - 1])
            {
                v[w - 1] += values[src];  // left fold in input order
            }
            else
            {
                ci[w] = col_indices[src];
                v[w] = values[src];
                ++w;
            }
        }
        ro[r + 1] = w;
    }
}

template <typename Scalar>
void spmvRows(Index rows, const Index* ro, const Index* ci, const Scalar* v, const Scalar* x, Scalar* y)
{
    for (Index r = 0; r < rows; ++r)
    {
        Scalar acc = Scalar(0);
        for (Index k = ro[r]; k < ro[r + 1]; ++k) acc += product(v[k], x[ci[k]]);
        y[r] = acc;
    }
}

struct SolveOutcome
{
    bool converged = false;
    int iterations = 0;
    double initialResidual = 0.0;
    double finalResidual = 0.0;
};

inline void checkSolverOptions(int max_iterations, double rel, double abs_tol)
{
    if (max_iterations < 0) throw std::invalid_argument("iterative solver maxIterations must be non-negative");
    if (!std::isfinite(rel) || rel < 0.0 || rel > 1.0 || !std::isfinite(abs_tol) || abs_tol < 0.0)
        throw std::invalid_argument(
            "iterative solver relative tolerance must be in [0, 1] and absolute tolerance must be finite and "
            "non-negative");
}

// Structural and value checks of an untrusted CPU system.
template <typename Scalar>
void checkSystem(Index rows, Index cols, Index nnz, const Index* ro, const Index* ci, const Scalar* v,
                 Index rhs_rows, Index rhs_cols, Index x_rows, Index x_cols, const Scalar* b, const Scalar* x)
{
    if (rows != cols) throw std::invalid_argument("iterative solver matrix must be square");
    if (rhs_rows != rows || rhs_cols != 1) throw std::invalid_argument("iterative solver rhs must be matrix.rows() x 1");
    if (x_rows != cols || x_cols != 1)
        throw std::invalid_argument("iterative solver solution must be matrix.cols() x 1");
    if (b != nullptr && b == x) throw std::invalid_argument("iterative solver rhs and solution must not alias");
    if (ro[0] != 0 || ro[rows] != nnz)
        throw std::invalid_argument("iterative solver requires CSR row offsets to start at zero and end at nnz");
    for (Index r = 0; r < rows; ++r)
        if (ro[r] < 0 || ro[r + 1] < ro[r] || ro[r + 1] > nnz)
            throw std::invalid_argument("iterative solver invalid CSR row offsets at row " + std::to_string(r));
    for (Index k = 0; k < nnz; ++k)
    {
        if (ci[k] < 0 || ci[k] >= cols)
            throw std::invalid_argument("iterative solver CSR column index out of range at position " +
                                        std::to_string(k));
        if (!std::isfinite(static_cast<double>(v[k])))
            throw std::invalid_argument("iterative solver CSR value is not finite at position " + std::to_string(k));
    }
    for (Index r = 0; r < rows; ++r)
        if (!std::isfinite(static_cast<double>(b[r])) || !std::isfinite(static_cast<double>(x[r])))
            throw std::invalid_argument("iterative solver rhs and solution must be finite at row " + std::to_string(r));
}

// y = A u in double, rows left to right, products rounded before the add.
template <typename Scalar>
void applyDouble(Index rows, const Index* ro, const Index* ci, const Scalar* v, const double* u, double* y)
{
    for (Index r = 0; r < rows; ++r)
    {
        double acc = 0.0;
        for (Index k = ro[r]; k < ro[r + 1]; ++k) acc += product(static_cast<double>(v[k]), u[ci[k]]);
        y[r] = acc;
    }
}

inline double dotDouble(const std::vector<double>& a, const std::vector<double>& b)
{
    double s = 0.0;
    for (std::size_t i = 0; i < a.size(); ++i) s += product(a[i], b[i]);
    return s;
}

// Jacobi inverse diagonal: the diagonal entries of a row are summed in CSR
// order and must exceed 16 eps(Scalar) times the row's largest magnitude.
template <typename Scalar>
std::vector<double> inverseDiagonal(Index rows, const Index* ro, const Index* ci, const Scalar* v)
{
    std::vector<double> inv(static_cast<std::size_t>(rows));
    for (Index r = 0; r < rows; ++r)
    {
        double diag = 0.0, scale = 0.0;
        bool seen = false;
        for (Index k = ro[r]; k < ro[r + 1]; ++k)
        {
            const double a = static_cast<double>(v[k]);
            scale = std::max(scale, std::abs(a));
            if (ci[k] == r)
            {
                diag += a;
                seen = true;
            }
        }
        const double floor = static_cast<double>(std::numeric_limits<Scalar>::epsilon()) * scale * 16.0;
        const double reciprocal = 1.0 / diag;
        if (!seen || !std::isfinite(diag) || diag <= floor || !std::isfinite(reciprocal))
            throw std::runtime_error("pcg Jacobi diagonal is missing, non-positive, or near zero at row " +
                                     std::to_string(r));
        inv[static_cast<std::size_t>(r)] = reciprocal;
    }
    return inv;
}

// Conjugate gradients in double on a validated host system; x is in/out and
// written only when the iteration ends without an error.
template <typename Scalar>
SolveOutcome conjugateGradients(Index n, const Index* ro, const Index* ci, const Scalar* v, const Scalar* b, Scalar* x_io,
                                int max_iterations, double rel, double abs_tol, bool jacobi, bool require)
{
    const std::size_t size = static_cast<std::size_t>(n);
    std::vector<double> x(size), r(size), p(size), z(size), q(size);
    for (std::size_t i = 0; i < size; ++i) x[i] = static_cast<double>(x_io[i]);
    applyDouble(n, ro, ci, v, x.data(), q.data());
    for (std::size_t i = 0; i < size; ++i) r[i] = static_cast<double>(b[i]) - q[i];
    SolveOutcome out;
    const double rr0 = dotDouble(r, r);
    if (!std::isfinite(rr0) || rr0 < 0.0) throw std::runtime_error("iterative solver initial residual is not finite");
    out.initialResidual = std::sqrt(rr0);
    out.finalResidual = out.initialResidual;
    const double tol = std::max(abs_tol, rel * out.initialResidual);
    if (out.finalResidual <= tol)
    {
        out.converged = true;
        return out;
    }
    const std::vector<double> inv = jacobi ? inverseDiagonal(n, ro, ci, v) : std::vector<double>{};
    auto precondition = [&]() {
        for (std::size_t i = 0; i < size; ++i) z[i] = jacobi ? product(inv[i], r[i]) : r[i];
    };
    precondition();
    p = z;
    double rho = dotDouble(r, z);
    if (!std::isfinite(rho) || rho <= 0.0)
        throw std::runtime_error("iterative solver breakdown: preconditioned residual is not positive");
    for (int it = 0; it < max_iterations; ++it)
    {
        applyDouble(n, ro, ci, v, p.data(), q.data());
        const double pq = dotDouble(p, q);
        if (!std::isfinite(pq) || pq <= 0.0)
            throw std::runtime_error("iterative solver breakdown: matrix is not numerically SPD");
        const double alpha = rho / pq;
        if (!std::isfinite(alpha)) throw std::runtime_error("iterative solver breakdown: step size is not finite");
        for (std::size_t i = 0; i < size; ++i)
        {
            x[i] += product(alpha, p[i]);
            r[i] -= product(alpha, q[i]);
        }
        out.iterations = it + 1;
        const double rr = dotDouble(r, r);
        if (!std::isfinite(rr) || rr < 0.0) throw std::runtime_error("iterative solver residual is not finite");
        out.finalResidual = std::sqrt(rr);
        if (out.finalResidual <= tol)
        {
            out.converged = true;
            break;
        }
        precondition();
        const double rho_next = dotDouble(r, z);
        if (!std::isfinite(rho_next) || rho_next <= 0.0)
            throw std::runtime_error("iterative solver breakdown in direction update");
        const double beta = rho_next / rho;
        if (!std::isfinite(beta)) throw std::runtime_error("iterative solver breakdown: direction scale is not finite");
        for (std::size_t i = 0; i < size; ++i) p[i] = z[i] + product(beta, p[i]);
        rho = rho_next;
    }
    for (std::size_t i = 0; i < size; ++i) x_io[i] = static_cast<Scalar>(x[i]);
    if (!out.converged && require)
    {
        char buf[64];
        std::snprintf(buf, sizeof(buf), "%g", out.finalResidual);
        throw std::runtime_error("iterative solver did not converge in " + std::to_string(out.iterations) +
                                 " iterations; final residual=" + buf);
    }
    return out;
}

} // namespace host
} // namespace detail
} // namespace plamatrix
