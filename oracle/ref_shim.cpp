// Test infrastructure only: a C entry layer over PlaMatrix's own CPU build,
// compiled (by oracle/Makefile, target `ref`) together with the reference's
// CPU sources where they lie, into oracle/_ref/libplamatrix_ref.so.  The tests
// use it to check the C restatement in oracle/*.c bit for bit.  Nothing in the
// product links it.
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

#include "plamatrix/optimization/block_schur.h"
#include "plamatrix/sparse/iterative_solver.h"
#include "plamatrix/sparse/sparse_ops.h"

using namespace plamatrix;

namespace
{
thread_local std::string g_error;

template <typename F>
int guarded(F&& f)
{
    try
    {
        f();
        return 0;
    }
    catch (const std::out_of_range& e)
    {
        g_error = e.what();
        return 2;
    }
    catch (const std::invalid_argument& e)
    {
        g_error = e.what();
        return 1;
    }
    catch (const std::overflow_error& e)
    {
        g_error = e.what();
        return 3;
    }
    catch (const std::logic_error& e)
    {
        g_error = e.what();
        return 4;
    }
    catch (const std::exception& e)
    {
        g_error = e.what();
        return 5;
    }
}

template <typename S>
int cooToCsrImpl(Index rows, Index cols, Index count, const Index* r, const Index* c, const S* v,
                 Index* row_offsets, Index* col_out, S* val_out, Index* nnz)
{
    return guarded([&] {
        std::vector<Index> rv(r, r + count), cv(c, c + count);
        std::vector<S> vv(v, v + count);
        auto csr = cooToCsr(rows, cols, rv, cv, vv);
        *nnz = csr.nnz();
        std::memcpy(row_offsets, csr.rowOffsets(), sizeof(Index) * static_cast<std::size_t>(rows + 1));
        if (csr.nnz() > 0)
        {
            std::memcpy(col_out, csr.colIndices(), sizeof(Index) * static_cast<std::size_t>(csr.nnz()));
            std::memcpy(val_out, csr.values(), sizeof(S) * static_cast<std::size_t>(csr.nnz()));
        }
    });
}

template <typename S>
CSRMatrix<S, Device::CPU> makeCsr(Index rows, Index cols, Index nnz, const Index* ro, const Index* ci, const S* v)
{
    CSRMatrix<S, Device::CPU> m(rows, cols, nnz);
    std::memcpy(m.rowOffsets(), ro, sizeof(Index) * static_cast<std::size_t>(rows + 1));
    if (nnz > 0)
    {
        std::memcpy(m.colIndices(), ci, sizeof(Index) * static_cast<std::size_t>(nnz));
        std::memcpy(m.values(), v, sizeof(S) * static_cast<std::size_t>(nnz));
    }
    return m;
}

template <typename S>
int spmvImpl(Index rows, Index cols, Index nnz, const Index* ro, const Index* ci, const S* v, const S* x, S* y)
{
    return guarded([&] {
        auto m = makeCsr(rows, cols, nnz, ro, ci, v);
        DenseMatrix<S, Device::CPU> xin(cols, 1);
        std::memcpy(xin.data(), x, sizeof(S) * static_cast<std::size_t>(cols));
        auto out = spmv(m, xin);
        std::memcpy(y, out.data(), sizeof(S) * static_cast<std::size_t>(rows));
    });
}

template <typename S>
int spmmImpl(Index rows, Index cols, Index nnz, const Index* ro, const Index* ci, const S* v, Index k, const S* b,
             S* y)
{
    return guarded([&] {
        auto m = makeCsr(rows, cols, nnz, ro, ci, v);
        DenseMatrix<S, Device::CPU> bin(cols, k);
        if (cols * k > 0) std::memcpy(bin.data(), b, sizeof(S) * static_cast<std::size_t>(cols * k));
        auto out = spmm(m, bin);
        if (rows * k > 0) std::memcpy(y, out.data(), sizeof(S) * static_cast<std::size_t>(rows * k));
    });
}

struct RefOptions
{
    int maxIterations;
    double relativeTolerance;
    double absoluteTolerance;
    int useJacobi;
    int requireConvergence;
};

struct RefReport
{
    int converged;
    int iterations;
    double initialResidual;
    double finalResidual;
};

template <typename S>
int solveImpl(Index rows, Index cols, Index nnz, const Index* ro, const Index* ci, const S* v, const S* rhs, S* x,
              const RefOptions* o, int preconditioned, RefReport* report)
{
    return guarded([&] {
        auto m = makeCsr(rows, cols, nnz, ro, ci, v);
        DenseMatrix<S, Device::CPU> b(rows, 1), sol(cols, 1);
        std::memcpy(b.data(), rhs, sizeof(S) * static_cast<std::size_t>(rows));
        std::memcpy(sol.data(), x, sizeof(S) * static_cast<std::size_t>(cols));
        IterativeSolverOptions options;
        options.maxIterations = o->maxIterations;
        options.relativeTolerance = o->relativeTolerance;
        options.absoluteTolerance = o->absoluteTolerance;
        options.useJacobiPreconditioner = o->useJacobi != 0;
        options.requireConvergence = o->requireConvergence != 0;
        IterativeSolverReport r;
        try
        {
            r = preconditioned ? pcg(m, b, sol, options) : cg(m, b, sol, options);
        }
        catch (...)
        {
            std::memcpy(x, sol.data(), sizeof(S) * static_cast<std::size_t>(cols));
            throw;
        }
        std::memcpy(x, sol.data(), sizeof(S) * static_cast<std::size_t>(cols));
        report->converged = r.converged;
        report->iterations = r.iterations;
        report->initialResidual = r.initialResidual;
        report->finalResidual = r.finalResidual;
    });
}

struct RefSchurReport
{
    int converged;
    int iterations;
    double initialResidualNorm;
    double finalResidualNorm;
};
} // namespace

extern "C" {

const char* ref_last_error() { return g_error.c_str(); }

int ref_coo_to_csr_f64(Index rows, Index cols, Index count, const Index* r, const Index* c, const double* v,
                       Index* ro, Index* co, double* vo, Index* nnz)
{
    return cooToCsrImpl(rows, cols, count, r, c, v, ro, co, vo, nnz);
}

int ref_coo_to_csr_f32(Index rows, Index cols, Index count, const Index* r, const Index* c, const float* v,
                       Index* ro, Index* co, float* vo, Index* nnz)
{
    return cooToCsrImpl(rows, cols, count, r, c, v, ro, co, vo, nnz);
}

int ref_spmv_f64(Index rows, Index cols, Index nnz, const Index* ro, const Index* ci, const double* v,
                 const double* x, double* y)
{
    return spmvImpl(rows, cols, nnz, ro, ci, v, x, y);
}

int ref_spmv_f32(Index rows, Index cols, Index nnz, const Index* ro, const Index* ci, const float* v,
                 const float* x, float* y)
{
    return spmvImpl(rows, cols, nnz, ro, ci, v, x, y);
}

int ref_spmm_f64(Index rows, Index cols, Index nnz, const Index* ro, const Index* ci, const double* v, Index k,
                 const double* b, double* y)
{
    return spmmImpl(rows, cols, nnz, ro, ci, v, k, b, y);
}

int ref_spmm_f32(Index rows, Index cols, Index nnz, const Index* ro, const Index* ci, const float* v, Index k,
                 const float* b, float* y)
{
    return spmmImpl(rows, cols, nnz, ro, ci, v, k, b, y);
}

int ref_solve_f64(Index rows, Index cols, Index nnz, const Index* ro, const Index* ci, const double* v,
                  const double* rhs, double* x, const RefOptions* o, int preconditioned, RefReport* report)
{
    return solveImpl(rows, cols, nnz, ro, ci, v, rhs, x, o, preconditioned, report);
}

int ref_solve_f32(Index rows, Index cols, Index nnz, const Index* ro, const Index* ci, const float* v,
                  const float* rhs, float* x, const RefOptions* o, int preconditioned, RefReport* report)
{
    return solveImpl(rows, cols, nnz, ro, ci, v, rhs, x, o, preconditioned, report);
}

// Build BlockNormalEquations<double> from a batch of single-primary residuals
// (added in index order), optionally followed by primary-only and
// eliminated-only priors, and solve with the given backend (0 Cpu, 1 DenseCpu).
int ref_schur_solve_batch(Index pc, Index ec, Index ps, Index es, Index count, const Index* primary,
                          const Index* eliminated, const double* jp, const double* je, const double* r,
                          Index rs, const double* weights, double damping, int backend, int maxIterations,
                          double rel, double abs, double* primary_step, double* eliminated_step,
                          RefSchurReport* report)
{
    return guarded([&] {
        BlockNormalEquations<double> eq(pc, ec, ps, es);
        for (Index i = 0; i < count; ++i)
        {
            eq.addResidualBlock(primary[i], eliminated[i], jp + i * rs * ps, je + i * rs * es, r + i * rs, rs,
                                weights ? weights[i] : 1.0);
        }
        SchurComplementSolverOptions<double> options;
        options.linearBackend = backend == 1 ? SchurComplementLinearBackend::DenseCpu
                                             : SchurComplementLinearBackend::Cpu;
        options.maxIterations = maxIterations;
        options.relativeTolerance = rel;
        options.absoluteTolerance = abs;
        std::vector<double> dp, de;
        const auto rep = solveDampedSchurComplement(eq, damping, options, &dp, &de);
        report->converged = rep.converged;
        report->iterations = rep.iterations;
        report->initialResidualNorm = rep.initialResidualNorm;
        report->finalResidualNorm = rep.finalResidualNorm;
        std::memcpy(primary_step, dp.data(), sizeof(double) * dp.size());
        std::memcpy(eliminated_step, de.data(), sizeof(double) * de.size());
    });
}
}
