// CG / Jacobi-PCG / block-Jacobi PCG on the GPU (sparse/iterative_solver.h:
// 10-171 of the reference) over the C ABI.  The solve runs entirely on the
// device; the host reads the stop state once per graph-replayed chunk of
// iterations instead of once per iteration.
#pragma once

#include <cmath>
#include <stdexcept>
#include <utility>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "plamatrix/core/status.h"
#include "plamatrix/dense/dense_matrix.h"
#include "plamatrix/sparse/csr_matrix.h"
#include "plamatrix/sparse/sparse_ops.h"
#include "plamatrix_amd.h"

namespace plamatrix
{

struct IterativeSolverOptions
{
    int maxIterations = 1000;
    double relativeTolerance = 1.0e-6;
    double absoluteTolerance = 0.0;
    bool useJacobiPreconditioner = true;
    bool requireConvergence = false;
};

struct IterativeSolverReport
{
    bool converged = false;
    int iterations = 0;
    double initialResidual = 0.0;
    double finalResidual = 0.0;
};

/// Move-only reusable device storage for CG/PCG solves, bound to its first stream.
template <typename Scalar>
class IterativeSolverWorkspace
{
public:
    IterativeSolverWorkspace() { detail::check(pm_workspace_create(&_ws)); }
    ~IterativeSolverWorkspace()
    {
        if (_ws) pm_workspace_destroy(_ws);
    }
    IterativeSolverWorkspace(const IterativeSolverWorkspace&) = delete;
    IterativeSolverWorkspace& operator=(const IterativeSolverWorkspace&) = delete;
    IterativeSolverWorkspace(IterativeSolverWorkspace&& o) noexcept : _ws(std::exchange(o._ws, nullptr)) {}
    IterativeSolverWorkspace& operator=(IterativeSolverWorkspace&& o) noexcept
    {
        if (this != &o)
        {
            if (_ws) pm_workspace_destroy(_ws);
            _ws = std::exchange(o._ws, nullptr);
        }
        return *this;
    }
    /// Release storage (a no-op after a move); the bound stream must have
    /// completed its work, else std::logic_error.
    void closeAsyncAllocation()
    {
        if (_ws) detail::check(pm_workspace_close(_ws));
    }
    Index capacitySize() const noexcept { return pm_workspace_capacity_rows(_ws); }
    pm_workspace* handle() const noexcept { return _ws; }

private:
    pm_workspace* _ws = nullptr;
};

/// Device state of a fixed-iteration asynchronous solve.
struct AsyncIterativeSolverState
{
    AsyncIterativeSolverState() = default;
    ~AsyncIterativeSolverState()
    {
        if (_state) pm_fixed_state_destroy(_state);
    }
    AsyncIterativeSolverState(const AsyncIterativeSolverState&) = delete;
    AsyncIterativeSolverState& operator=(const AsyncIterativeSolverState&) = delete;
    AsyncIterativeSolverState(AsyncIterativeSolverState&& o) noexcept
        : submittedIterations(o.submittedIterations), _state(std::exchange(o._state, nullptr))
    {
    }
    AsyncIterativeSolverState& operator=(AsyncIterativeSolverState&& o) noexcept
    {
        if (this != &o)
        {
            if (_state) pm_fixed_state_destroy(_state);
            submittedIterations = o.submittedIterations;
            _state = std::exchange(o._state, nullptr);
        }
        return *this;
    }

    int submittedIterations = 0;
    pm_fixed_state* handle() const noexcept { return _state; }
    /// Release the device state once the solve's stream has completed;
    /// repeated calls are harmless.
    void closeAsyncAllocation() noexcept
    {
        if (_state) pm_fixed_state_destroy(std::exchange(_state, nullptr));
    }
    pm_fixed_state** slot() noexcept { return &_state; }

private:
    pm_fixed_state* _state = nullptr;
};

namespace detail
{
inline pm_solver_options toAbi(const IterativeSolverOptions& o)
{
    pm_solver_options out{};
    out.max_iterations = o.maxIterations;
    out.relative_tolerance = o.relativeTolerance;
    out.absolute_tolerance = o.absoluteTolerance;
    out.use_jacobi = o.useJacobiPreconditioner ? 1 : 0;
    out.require_convergence = o.requireConvergence ? 1 : 0;
    return out;
}

inline IterativeSolverReport fromAbi(const pm_solver_report& r)
{
    IterativeSolverReport out;
    out.converged = r.converged != 0;
    out.iterations = r.iterations;
    out.initialResidual = r.initial_residual;
    out.finalResidual = r.final_residual;
    return out;
}

// Shape, aliasing, stream and trust rules of the solvers
// (iterative_solver_cuda_detail.h:40-82 of the reference): stream-ordered
// operands are used on their owning stream; untrusted structure is
// revalidated by the adaptive solvers (throwing std::invalid_argument before
// any launch when it is malformed) and rejected by the fixed-iteration ones.
template <typename Scalar, typename Rhs, typename Sol>
void checkSystem(const CSRMatrix<Scalar, Device::GPU>& a, const Rhs& rhs, const Sol& x, hipStream_t stream,
                 bool require_prevalidated = false)
{
    if (a.rows() != a.cols()) throw std::invalid_argument("iterative solver requires a square matrix");
    if (rhs.rows() != a.rows() || rhs.cols() != 1 || x.rows() != a.rows() || x.cols() != 1)
        throw std::invalid_argument("iterative solver vector dimensions do not match the matrix");
    if (rhs.data() != nullptr && static_cast<const void*>(rhs.data()) == static_cast<const void*>(x.data()))
        throw std::invalid_argument("iterative solver rhs and solution must not alias");
    checkStorageStream("iterative solver matrix", a, stream);
    checkStorageStream("iterative solver rhs", rhs, stream);
    checkStorageStream("iterative solver solution", x, stream);
    if (!a.isStructureUsableOnStream(stream))
    {
        if (require_prevalidated)
            throw std::logic_error("fixed iterative solver requires prevalidated CSR structure; "
                                   "call matrix.validateStructure(stream) after mutable array access");
        a.validateStructure(stream);
    }
}
} // namespace detail

/// Adaptively solve an SPD GPU system with conjugate gradients.
namespace detail
{
template <typename Scalar>
IterativeSolverReport hostSolve(const CSRMatrix<Scalar, Device::CPU>& matrix,
                                const DenseMatrix<Scalar, Device::CPU>& rhs,
                                DenseMatrix<Scalar, Device::CPU>& solution, const IterativeSolverOptions& options,
                                bool preconditioned)
{
    host::checkSolverOptions(options.maxIterations, options.relativeTolerance, options.absoluteTolerance);
    host::checkSystem(matrix.rows(), matrix.cols(), matrix.nnz(), matrix.rowOffsets(), matrix.colIndices(),
                      matrix.values(), rhs.rows(), rhs.cols(), solution.rows(), solution.cols(), rhs.data(),
                      static_cast<const Scalar*>(solution.data()));
    const host::SolveOutcome o = host::conjugateGradients(
        matrix.rows(), matrix.rowOffsets(), matrix.colIndices(), matrix.values(), rhs.data(), solution.data(),
        options.maxIterations, options.relativeTolerance, options.absoluteTolerance,
        preconditioned && options.useJacobiPreconditioner, options.requireConvergence);
    return IterativeSolverReport{o.converged, o.iterations, o.initialResidual, o.finalResidual};
}
} // namespace detail

/// Conjugate gradients on a host system (iterative_solver.h:94-99 of the
/// reference): double-precision iteration, solution in/out.
template <typename Scalar>
IterativeSolverReport cg(const CSRMatrix<Scalar, Device::CPU>& matrix, const DenseMatrix<Scalar, Device::CPU>& rhs,
                         DenseMatrix<Scalar, Device::CPU>& solution, const IterativeSolverOptions& options = {})
{
    return detail::hostSolve(matrix, rhs, solution, options, false);
}

/// Jacobi-preconditioned conjugate gradients on a host system
/// (iterative_solver.h:101-108); useJacobiPreconditioner=false gives CG.
template <typename Scalar>
IterativeSolverReport pcg(const CSRMatrix<Scalar, Device::CPU>& matrix, const DenseMatrix<Scalar, Device::CPU>& rhs,
                          DenseMatrix<Scalar, Device::CPU>& solution, const IterativeSolverOptions& options = {})
{
    return detail::hostSolve(matrix, rhs, solution, options, true);
}

template <typename Scalar>
IterativeSolverReport cg(const CSRMatrix<Scalar, Device::GPU>& matrix, const DenseMatrix<Scalar, Device::GPU>& rhs,
                         DenseMatrix<Scalar, Device::GPU>& solution, IterativeSolverWorkspace<Scalar>& workspace,
                         const IterativeSolverOptions& options = {}, hipStream_t stream = nullptr)
{
    detail::checkSystem(matrix, rhs, solution, stream);
    const pm_csr d = matrix.descriptor(true, stream);
    const pm_solver_options o = detail::toAbi(options);
    pm_solver_report r{};
    detail::check(pm_cg(&d, rhs.data(), solution.data(), &o, nullptr, &r, workspace.handle(), stream));
    return detail::fromAbi(r);
}

/// Adaptively solve an SPD GPU system with optional Jacobi preconditioning.
template <typename Scalar>
IterativeSolverReport pcg(const CSRMatrix<Scalar, Device::GPU>& matrix, const DenseMatrix<Scalar, Device::GPU>& rhs,
                          DenseMatrix<Scalar, Device::GPU>& solution, IterativeSolverWorkspace<Scalar>& workspace,
                          const IterativeSolverOptions& options = {}, hipStream_t stream = nullptr)
{
    detail::checkSystem(matrix, rhs, solution, stream);
    const pm_csr d = matrix.descriptor(true, stream);
    const pm_solver_options o = detail::toAbi(options);
    pm_solver_report r{};
    detail::check(pm_pcg(&d, rhs.data(), solution.data(), &o, nullptr, &r, workspace.handle(), stream));
    return detail::fromAbi(r);
}

/// Adaptively solve with caller-supplied inverse diagonal blocks, stored
/// row-major as a (rows * block_size) x 1 vector.
template <typename Scalar>
IterativeSolverReport blockPcg(const CSRMatrix<Scalar, Device::GPU>& matrix,
                               const DenseMatrix<Scalar, Device::GPU>& rhs, DenseMatrix<Scalar, Device::GPU>& solution,
                               const DenseMatrix<Scalar, Device::GPU>& inverse_blocks, Index block_size,
                               IterativeSolverWorkspace<Scalar>& workspace, const IterativeSolverOptions& options = {},
                               hipStream_t stream = nullptr)
{
    detail::checkSystem(matrix, rhs, solution, stream);
    if (block_size <= 0 || inverse_blocks.size() != matrix.rows() * block_size)
        throw std::invalid_argument("blockPcg inverse blocks must hold rows * block_size values");
    detail::checkStorageStream("blockPcg inverse blocks", inverse_blocks, stream);
    const pm_csr d = matrix.descriptor(true, stream);
    const pm_solver_options o = detail::toAbi(options);
    pm_solver_report r{};
    detail::check(pm_block_pcg(&d, rhs.data(), solution.data(), inverse_blocks.data(), block_size, &o, nullptr, &r,
                               workspace.handle(), stream));
    return detail::fromAbi(r);
}

namespace detail
{
template <typename Scalar>
AsyncIterativeSolverState fixedAsync(const CSRMatrix<Scalar, Device::GPU>& matrix,
                                     const DenseMatrix<Scalar, Device::GPU>& rhs,
                                     DenseMatrix<Scalar, Device::GPU>& solution, int iterations, int preconditioned,
                                     IterativeSolverWorkspace<Scalar>& workspace, hipStream_t stream)
{
    checkSystem(matrix, rhs, solution, stream, true);
    const pm_csr d = matrix.descriptor(true, stream);
    AsyncIterativeSolverState state;
    check(pm_solve_fixed_async(&d, rhs.data(), solution.data(), iterations, preconditioned, state.slot(),
                               workspace.handle(), stream));
    state.submittedIterations = iterations;
    return state;
}
} // namespace detail

/// Submit exactly `iterations` CG steps without host convergence checks.
template <typename Scalar>
AsyncIterativeSolverState cgFixedIterationsAsync(const CSRMatrix<Scalar, Device::GPU>& matrix,
                                                 const DenseMatrix<Scalar, Device::GPU>& rhs,
                                                 DenseMatrix<Scalar, Device::GPU>& solution, int iterations,
                                                 IterativeSolverWorkspace<Scalar>& workspace,
                                                 hipStream_t stream = nullptr)
{
    return detail::fixedAsync(matrix, rhs, solution, iterations, 0, workspace, stream);
}

/// Submit exactly `iterations` Jacobi-PCG steps without host convergence checks.
template <typename Scalar>
AsyncIterativeSolverState pcgFixedIterationsAsync(const CSRMatrix<Scalar, Device::GPU>& matrix,
                                                  const DenseMatrix<Scalar, Device::GPU>& rhs,
                                                  DenseMatrix<Scalar, Device::GPU>& solution, int iterations,
                                                  IterativeSolverWorkspace<Scalar>& workspace,
                                                  hipStream_t stream = nullptr)
{
    return detail::fixedAsync(matrix, rhs, solution, iterations, 1, workspace, stream);
}

/// Read a completed fixed-iteration state and apply the convergence policy.
inline IterativeSolverReport finalizeIterativeSolverReport(const AsyncIterativeSolverState& state,
                                                           const IterativeSolverOptions& options = {})
{
    const pm_solver_options o = detail::toAbi(options);
    pm_solver_report r{};
    detail::check(pm_fixed_finalize(state.handle(), &o, &r));
    return detail::fromAbi(r);
}

// ---- extensions with no reference counterpart -----------------------------------

/// FP32 matrix, FP64 right-hand side and solution: Jacobi-PCG inner solves
/// (FP32 vectors, FP64 dots, FP64 residual refresh) inside FP64 iterative
/// refinement until ||b - A x|| / ||b|| <= target_relres in FP64.
struct RefinementReport
{
    bool converged = false;
    int outerIterations = 0;
    int innerIterations = 0;
    double initialRelres = 0.0;
    double finalRelres = 0.0;
};

inline RefinementReport pcgRefine(const CSRMatrix<float, Device::GPU>& matrix,
                                  const DenseMatrix<double, Device::GPU>& rhs,
                                  DenseMatrix<double, Device::GPU>& solution, IterativeSolverWorkspace<float>& workspace,
                                  double target_relres = 1e-8, int max_outer = 10,
                                  const IterativeSolverOptions& inner = {1000000, 1e-3, 0.0, true, false},
                                  int refresh_interval = 50, hipStream_t stream = nullptr)
{
    detail::checkSystem(matrix, rhs, solution, stream);
    const pm_csr d = matrix.descriptor(true, stream);
    const pm_solver_options o = detail::toAbi(inner);
    pm_solve_policy policy{};
    policy.use_graph = -1;
    policy.refresh_interval = refresh_interval;
    pm_refine_report r{};
    detail::check(pm_pcg_refine(&d, rhs.data(), solution.data(), &o, target_relres, max_outer, &policy, &r,
                                workspace.handle(), stream));
    RefinementReport out;
    out.converged = r.converged != 0;
    out.outerIterations = r.outer_iterations;
    out.innerIterations = r.inner_iterations;
    out.initialRelres = r.initial_relres;
    out.finalRelres = r.final_relres;
    return out;
}

/// Solve many independent systems stored as the diagonal blocks of one CSR
/// matrix (system s owns rows [row_begin[s], row_end[s]), each row_begin a
/// multiple of 64) in one set of launches with per-system device stop flags.
template <typename Scalar>
std::vector<pm_batch_report> batchedPcg(const CSRMatrix<Scalar, Device::GPU>& matrix,
                                        const std::vector<Index>& row_begin, const std::vector<Index>& row_end,
                                        const DenseMatrix<Scalar, Device::GPU>& rhs,
                                        DenseMatrix<Scalar, Device::GPU>& solution,
                                        IterativeSolverWorkspace<Scalar>& workspace,
                                        const IterativeSolverOptions& options = {}, int refresh_interval = 0,
                                        hipStream_t stream = nullptr)
{
    if (row_begin.size() != row_end.size()) throw std::invalid_argument("batchedPcg: row range arrays differ in size");
    detail::checkSystem(matrix, rhs, solution, stream);
    const pm_csr d = matrix.descriptor(true, stream);
    const pm_solver_options o = detail::toAbi(options);
    pm_solve_policy policy{};
    policy.use_graph = -1;
    policy.refresh_interval = refresh_interval;
    std::vector<pm_batch_report> reports(row_begin.size());
    detail::check(pm_batch_pcg(&d, static_cast<int64_t>(row_begin.size()), row_begin.data(), row_end.data(),
                               rhs.data(), solution.data(), &o, &policy, reports.data(), workspace.handle(), stream));
    return reports;
}

/// batchedPcg with FP64 refinement per system: FP32 matrix and inner solves,
/// FP64 right-hand sides and solutions (the initial guess), every system
/// driven to an FP64 true relative residual of target_relres or max_outer
/// outer steps.
inline std::vector<pm_batch_refine_report> batchedPcgRefine(
    const CSRMatrix<float, Device::GPU>& matrix, const std::vector<Index>& row_begin,
    const std::vector<Index>& row_end, const DenseMatrix<double, Device::GPU>& rhs,
    DenseMatrix<double, Device::GPU>& solution, IterativeSolverWorkspace<float>& workspace,
    double target_relres = 1e-6, int max_outer = 10,
    const IterativeSolverOptions& inner = {1000000, 1e-3, 0.0, true, false}, int refresh_interval = 50,
    hipStream_t stream = nullptr)
{
    if (row_begin.size() != row_end.size())
        throw std::invalid_argument("batchedPcgRefine: row range arrays differ in size");
    detail::checkSystem(matrix, rhs, solution, stream);
    const pm_csr d = matrix.descriptor(true, stream);
    const pm_solver_options o = detail::toAbi(inner);
    pm_solve_policy policy{};
    policy.use_graph = -1;
    policy.refresh_interval = refresh_interval;
    std::vector<pm_batch_refine_report> reports(row_begin.size());
    detail::check(pm_batch_pcg_refine(&d, static_cast<int64_t>(row_begin.size()), row_begin.data(), row_end.data(),
                                      rhs.data(), solution.data(), &o, target_relres, max_outer, &policy,
                                      reports.data(), workspace.handle(), stream));
    return reports;
}

} // namespace plamatrix
