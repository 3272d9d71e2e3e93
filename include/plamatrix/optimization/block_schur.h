// Block normal equations and the damped Schur-complement solve
// (optimization/block_schur.h:22-351 and robust_loss.h:27-51 of the
// reference).  BlockNormalEquations accumulates on the host with the
// reference's arithmetic and cross-block numbering, in float or double;
// solveDampedSchurComplement uploads the blocks once and runs the whole solve
// in HBM in FP64: damping, eliminated inverses, reduced right-hand side, Schur
// CSR on the cached device pattern, block-Jacobi PCG (or the dense Cholesky)
// and back-substitution (pm_schur_solve).  Float equations are widened to
// FP64 for the device solve and the steps rounded back to float.
#pragma once

#include <chrono>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "plamatrix/core/status.h"
#include "plamatrix/core/types.h"
#include "plamatrix_amd.h"

namespace plamatrix
{

template <typename Scalar>
struct RobustLossEvaluation
{
    Scalar cost = Scalar(0);
    Scalar weight = Scalar(1);
};

/// Huber loss 0.5 rho(s) of a squared residual norm s and its IRLS weight.
template <typename Scalar>
RobustLossEvaluation<Scalar> evaluateHuberLoss(Scalar squared_residual_norm, Scalar delta)
{
    static_assert(std::is_floating_point_v<Scalar>, "evaluateHuberLoss requires a floating-point scalar");
    if (!std::isfinite(squared_residual_norm) || squared_residual_norm < Scalar(0))
        throw std::invalid_argument("evaluateHuberLoss: squared residual norm must be finite and non-negative");
    if (!std::isfinite(delta) || delta < Scalar(0))
        throw std::invalid_argument("evaluateHuberLoss: delta must be finite and non-negative");
    if (delta == Scalar(0) || squared_residual_norm <= delta * delta)
        return {Scalar(0.5) * squared_residual_norm, Scalar(1)};
    const Scalar norm = std::sqrt(squared_residual_norm);
    return {delta * norm - Scalar(0.5) * delta * delta, delta / norm};
}

/// Linear solver of the reduced system.  Every served backend runs on the
/// device:
///   Cpu       block-Jacobi PCG with the reference Cpu backend's stopping rule
///             ||r|| <= abs + rel ||r0|| (block_schur_cpu_solver.h:34-35);
///   Hip       block-Jacobi PCG with the accelerated rule max(abs, rel ||r0||)
///             (the reference's Cuda / OpenCl semantics);
///   DenseCpu, DenseHip  dense FP64 Cholesky of the reduced system.
/// Cuda and OpenCl are rejected (std::invalid_argument).
enum class SchurComplementLinearBackend
{
    Cpu,
    DenseCpu,
    Cuda,
    OpenCl,
    Hip,
    DenseHip,
};

template <typename Scalar>
struct SchurComplementSolverOptions
{
    SchurComplementLinearBackend linearBackend = SchurComplementLinearBackend::Cpu;
    int deviceIndex = -1;
    int maxIterations = 100;
    Scalar relativeTolerance = Scalar(1e-8);
    Scalar absoluteTolerance = Scalar(1e-12);
    Scalar minimumDiagonal = Scalar(1e-12);
    bool useInitialGuess = false;
    /// Seed the FP64 block-PCG (Hip backend, double equations) with an FP32
    /// block-PCG pass, used when it converges (block_schur.h:43-44).
    bool useMixedPrecision = false;
    /// Kept for source compatibility (block_schur.h:45-46): the reduced
    /// values are always assembled and consumed in HBM here.
    bool schurValuesOnDevice = true;
};

template <typename Scalar>
struct SchurComplementSolverReport
{
    SchurComplementLinearBackend linearBackend = SchurComplementLinearBackend::Cpu;
    bool converged = false;
    int iterations = 0;
    Scalar initialResidualNorm = Scalar(0);
    Scalar finalResidualNorm = Scalar(0);
    /// Phase timings of the reference's report.  The phases run back to back
    /// on the device without host synchronisation, so only the whole device
    /// solve (linearSolveSeconds) is timed; the other phases read 0.
    double smallBlockInverseSeconds = 0.0;
    double schurAccumulationSeconds = 0.0;
    double csrConversionSeconds = 0.0;
    double schurAssemblySeconds = 0.0;
    double choleskyFactorizationSeconds = 0.0;
    double triangularSolveSeconds = 0.0;
    double residualCheckSeconds = 0.0;
    double linearSolveSeconds = 0.0;
    double backSubstitutionSeconds = 0.0;
    bool schurPatternReused = false;
    bool schurAssemblyOnDevice = false;
    bool mixedPrecisionUsed = false;
    std::string deviceName;
    std::string message;
};

/// Reusable device state: cached Schur pattern, numerics and solver storage.
template <typename Scalar>
class SchurComplementSolverWorkspace
{
public:
    SchurComplementSolverWorkspace()
    {
        detail::check(pm_schur_workspace_create(&_schur));
        detail::check(pm_workspace_create(&_solver));
    }
    ~SchurComplementSolverWorkspace() { release(); }
    SchurComplementSolverWorkspace(const SchurComplementSolverWorkspace&) = delete;
    SchurComplementSolverWorkspace& operator=(const SchurComplementSolverWorkspace&) = delete;

    /// Forget the cached pattern and storage.
    void clear()
    {
        release();
        detail::check(pm_schur_workspace_create(&_schur));
        detail::check(pm_workspace_create(&_solver));
    }
    std::size_t patternBuildCount() const noexcept
    {
        return static_cast<std::size_t>(pm_schur_pattern_build_count(_schur));
    }
    pm_schur_workspace* schur() const noexcept { return _schur; }
    pm_workspace* solver() const noexcept { return _solver; }

private:
    void release() noexcept
    {
        if (_schur) pm_schur_workspace_destroy(_schur);
        if (_solver) pm_workspace_destroy(_solver);
        _schur = nullptr;
        _solver = nullptr;
    }
    pm_schur_workspace* _schur = nullptr;
    pm_workspace* _solver = nullptr;
};

template <typename Scalar>
class BlockNormalEquations;

template <typename Scalar>
SchurComplementSolverReport<Scalar> solveDampedSchurComplement(const BlockNormalEquations<Scalar>& equations,
                                                               Scalar damping,
                                                               const SchurComplementSolverOptions<Scalar>& options,
                                                               SchurComplementSolverWorkspace<Scalar>& workspace,
                                                               std::vector<Scalar>* primary_step,
                                                               std::vector<Scalar>* eliminated_step);

/// Block-sparse normal equations of a bipartite least-squares problem
/// (block_normal_equations.cpp:13-411): B += w Jp^T Jp, C += w Je^T Je,
/// W += w Jp^T Je (one block per distinct (primary, eliminated) pair, numbered
/// in creation order), g += w J^T r, and primary cross blocks (row < col).
template <typename Scalar>
class BlockNormalEquations
{
    static_assert(std::is_same_v<Scalar, double> || std::is_same_v<Scalar, float>,
                  "BlockNormalEquations holds float or double blocks");

public:
    BlockNormalEquations(Index primary_block_count, Index eliminated_block_count, Index primary_block_size,
                         Index eliminated_block_size)
        : _pc(primary_block_count), _ec(eliminated_block_count), _ps(primary_block_size),
          _es(eliminated_block_size)
    {
        if (_pc < 0 || _ec < 0 || _ps <= 0 || _es <= 0)
            throw std::invalid_argument(
                "BlockNormalEquations: block counts must be non-negative and block sizes positive");
        _B.assign(static_cast<std::size_t>(_pc * _ps * _ps), Scalar(0));
        _C.assign(static_cast<std::size_t>(_ec * _es * _es), Scalar(0));
        _gp.assign(static_cast<std::size_t>(_pc * _ps), Scalar(0));
        _ge.assign(static_cast<std::size_t>(_ec * _es), Scalar(0));
        _adjacency.resize(static_cast<std::size_t>(_ec));
    }

    void addResidualBlock(Index primary_block, Index eliminated_block, const Scalar* primary_jacobian,
                          const Scalar* eliminated_jacobian, const Scalar* residual, Index residual_size,
                          Scalar weight = Scalar(1))
    {
        addResidualBlocks({primary_block}, {primary_jacobian}, eliminated_block, eliminated_jacobian, residual,
                          residual_size, weight);
    }

    void addResidualBlocks(const std::vector<Index>& primary_blocks,
                           const std::vector<const Scalar*>& primary_jacobians, Index eliminated_block,
                           const Scalar* eliminated_jacobian, const Scalar* residual, Index residual_size,
                           Scalar weight = Scalar(1))
    {
        if (eliminated_block < 0 || eliminated_block >= _ec)
            throw std::out_of_range("BlockNormalEquations::addResidualBlocks: eliminated block out of range");
        validate(eliminated_jacobian, residual_size * _es, residual, residual_size, weight,
                 "BlockNormalEquations::addResidualBlocks(eliminated)");
        addPrimaryTerms(primary_blocks, primary_jacobians, residual, residual_size, weight);
        for (std::size_t i = 0; i < primary_blocks.size(); ++i)
        {
            Scalar* w = &_crossValues[crossIndex(primary_blocks[i], eliminated_block) * static_cast<std::size_t>(_ps * _es)];
            const Scalar* jp = primary_jacobians[i];
            for (Index row = 0; row < residual_size; ++row)
                for (Index pc = 0; pc < _ps; ++pc)
                {
                    const Scalar scaled = weight * jp[row * _ps + pc];
                    for (Index ec = 0; ec < _es; ++ec) w[pc * _es + ec] += scaled * eliminated_jacobian[row * _es + ec];
                }
        }
        accumulateEliminated(eliminated_block, eliminated_jacobian, residual, residual_size, weight);
    }

    void addPrimaryResidualBlock(Index primary_block, const Scalar* primary_jacobian, const Scalar* residual,
                                 Index residual_size, Scalar weight = Scalar(1))
    {
        addPrimaryTerms({primary_block}, {primary_jacobian}, residual, residual_size, weight);
    }

    void addPrimaryResidualBlocks(const std::vector<Index>& primary_blocks,
                                  const std::vector<const Scalar*>& primary_jacobians, const Scalar* residual,
                                  Index residual_size, Scalar weight = Scalar(1))
    {
        addPrimaryTerms(primary_blocks, primary_jacobians, residual, residual_size, weight);
    }

    void addEliminatedResidualBlock(Index eliminated_block, const Scalar* eliminated_jacobian, const Scalar* residual,
                                    Index residual_size, Scalar weight = Scalar(1))
    {
        if (eliminated_block < 0 || eliminated_block >= _ec)
            throw std::out_of_range("BlockNormalEquations::addEliminatedResidualBlock: block index out of range");
        validate(eliminated_jacobian, residual_size * _es, residual, residual_size, weight,
                 "BlockNormalEquations::addEliminatedResidualBlock");
        accumulateEliminated(eliminated_block, eliminated_jacobian, residual, residual_size, weight);
    }

    /// Deterministically add another equation set with the same layout.
    void mergeFrom(const BlockNormalEquations& other)
    {
        if (_pc != other._pc || _ec != other._ec || _ps != other._ps || _es != other._es)
            throw std::invalid_argument("BlockNormalEquations::mergeFrom: block layouts must match");
        for (std::size_t i = 0; i < _B.size(); ++i) _B[i] += other._B[i];
        for (std::size_t i = 0; i < _C.size(); ++i) _C[i] += other._C[i];
        for (std::size_t i = 0; i < _gp.size(); ++i) _gp[i] += other._gp[i];
        for (std::size_t i = 0; i < _ge.size(); ++i) _ge[i] += other._ge[i];
        const std::size_t pp = static_cast<std::size_t>(_ps * _ps), pe = static_cast<std::size_t>(_ps * _es);
        for (std::size_t s = 0; s < other._pcRows.size(); ++s)
        {
            const std::size_t t = primaryCrossIndex(other._pcRows[s], other._pcCols[s]);
            for (std::size_t i = 0; i < pp; ++i) _pcValues[t * pp + i] += other._pcValues[s * pp + i];
        }
        for (std::size_t s = 0; s < other._crossPrimary.size(); ++s)
        {
            const std::size_t t = crossIndex(other._crossPrimary[s], other._crossEliminated[s]);
            for (std::size_t i = 0; i < pe; ++i) _crossValues[t * pe + i] += other._crossValues[s * pe + i];
        }
    }

    Index primaryBlockCount() const noexcept { return _pc; }
    Index eliminatedBlockCount() const noexcept { return _ec; }
    Index primaryBlockSize() const noexcept { return _ps; }
    Index eliminatedBlockSize() const noexcept { return _es; }

private:
    void validate(const Scalar* jacobian, Index jacobian_size, const Scalar* residual, Index residual_size,
                  Scalar weight, const char* op) const
    {
        if (!jacobian || !residual || residual_size <= 0)
            throw std::invalid_argument(std::string(op) + ": pointers must be non-null and residual size positive");
        if (!std::isfinite(weight) || weight < Scalar(0))
            throw std::invalid_argument(std::string(op) + ": weight must be finite and non-negative");
        for (Index i = 0; i < jacobian_size; ++i)
            if (!std::isfinite(jacobian[i])) throw std::invalid_argument(std::string(op) + ": Jacobian values must be finite");
        for (Index i = 0; i < residual_size; ++i)
            if (!std::isfinite(residual[i])) throw std::invalid_argument(std::string(op) + ": residual values must be finite");
    }

    std::size_t crossIndex(Index primary, Index eliminated)
    {
        auto& list = _adjacency[static_cast<std::size_t>(eliminated)];
        for (const std::size_t c : list)
            if (_crossPrimary[c] == primary) return c;
        const std::size_t c = _crossPrimary.size();
        _crossPrimary.push_back(primary);
        _crossEliminated.push_back(eliminated);
        _crossValues.resize(_crossValues.size() + static_cast<std::size_t>(_ps * _es), Scalar(0));
        list.push_back(c);
        return c;
    }

    std::size_t primaryCrossIndex(Index row, Index col)
    {
        if (row >= col) throw std::invalid_argument("BlockNormalEquations: primary cross blocks must be upper triangular");
        for (std::size_t i = 0; i < _pcRows.size(); ++i)
            if (_pcRows[i] == row && _pcCols[i] == col) return i;
        _pcRows.push_back(row);
        _pcCols.push_back(col);
        _pcValues.resize(_pcValues.size() + static_cast<std::size_t>(_ps * _ps), Scalar(0));
        return _pcRows.size() - 1;
    }

    void addPrimaryTerms(const std::vector<Index>& blocks, const std::vector<const Scalar*>& jacobians,
                         const Scalar* residual, Index rs, Scalar weight)
    {
        if (blocks.empty() || blocks.size() != jacobians.size())
            throw std::invalid_argument(
                "BlockNormalEquations: primary block and Jacobian lists must have equal non-zero size");
        for (std::size_t i = 0; i < blocks.size(); ++i)
        {
            if (blocks[i] < 0 || blocks[i] >= _pc)
                throw std::out_of_range("BlockNormalEquations: primary block index out of range");
            for (std::size_t j = 0; j < i; ++j)
                if (blocks[j] == blocks[i])
                    throw std::invalid_argument("BlockNormalEquations: primary block indices must be unique");
            validate(jacobians[i], rs * _ps, residual, rs, weight, "BlockNormalEquations::addPrimaryTerms");
        }
        for (std::size_t i = 0; i < blocks.size(); ++i)
        {
            Scalar* diag = &_B[static_cast<std::size_t>(blocks[i] * _ps * _ps)];
            Scalar* grad = &_gp[static_cast<std::size_t>(blocks[i] * _ps)];
            const Scalar* jac = jacobians[i];
            for (Index row = 0; row < rs; ++row)
                for (Index col = 0; col < _ps; ++col)
                {
                    const Scalar value = jac[row * _ps + col];
                    grad[col] += weight * value * residual[row];
                    for (Index k = 0; k < _ps; ++k) diag[col * _ps + k] += weight * value * jac[row * _ps + k];
                }
        }
        for (std::size_t l = 0; l < blocks.size(); ++l)
            for (std::size_t r = l + 1; r < blocks.size(); ++r)
            {
                const bool ordered = blocks[l] < blocks[r];
                const Scalar* row_jac = ordered ? jacobians[l] : jacobians[r];
                const Scalar* col_jac = ordered ? jacobians[r] : jacobians[l];
                const std::size_t t =
                    primaryCrossIndex(ordered ? blocks[l] : blocks[r], ordered ? blocks[r] : blocks[l]);
                Scalar* cross = &_pcValues[t * static_cast<std::size_t>(_ps * _ps)];
                for (Index rr = 0; rr < rs; ++rr)
                    for (Index row = 0; row < _ps; ++row)
                    {
                        const Scalar rv = row_jac[rr * _ps + row];
                        for (Index col = 0; col < _ps; ++col) cross[row * _ps + col] += weight * rv * col_jac[rr * _ps + col];
                    }
            }
    }

    void accumulateEliminated(Index e, const Scalar* jac, const Scalar* residual, Index rs, Scalar weight)
    {
        Scalar* diag = &_C[static_cast<std::size_t>(e * _es * _es)];
        Scalar* grad = &_ge[static_cast<std::size_t>(e * _es)];
        for (Index row = 0; row < rs; ++row)
            for (Index col = 0; col < _es; ++col)
            {
                const Scalar value = jac[row * _es + col];
                grad[col] += weight * value * residual[row];
                for (Index k = 0; k < _es; ++k) diag[col * _es + k] += weight * value * jac[row * _es + k];
            }
    }

    Index _pc, _ec, _ps, _es;
    std::vector<Scalar> _B, _C, _gp, _ge;
    std::vector<Index> _crossPrimary, _crossEliminated;
    std::vector<Scalar> _crossValues;
    std::vector<Index> _pcRows, _pcCols;
    std::vector<Scalar> _pcValues;
    std::vector<std::vector<std::size_t>> _adjacency;

    friend SchurComplementSolverReport<Scalar> solveDampedSchurComplement<Scalar>(
        const BlockNormalEquations<Scalar>&, Scalar, const SchurComplementSolverOptions<Scalar>&,
        SchurComplementSolverWorkspace<Scalar>&, std::vector<Scalar>*, std::vector<Scalar>*);
};

namespace detail
{
// One device copy of a host vector, freed on scope exit.
template <typename T>
struct Upload
{
    T* ptr = nullptr;
    explicit Upload(const std::vector<T>& v)
    {
        if (!v.empty())
        {
            checkHip(hipMalloc(reinterpret_cast<void**>(&ptr), sizeof(T) * v.size()), "hipMalloc");
            checkHip(hipMemcpy(ptr, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice), "hipMemcpy");
        }
    }
    explicit Upload(std::size_t n)
    {
        if (n) checkHip(hipMalloc(reinterpret_cast<void**>(&ptr), sizeof(T) * n), "hipMalloc");
    }
    ~Upload()
    {
        if (ptr) (void)hipFree(ptr);
    }
    Upload(const Upload&) = delete;
    Upload& operator=(const Upload&) = delete;
};
} // namespace detail

namespace detail
{
// FP64 device copy of a host vector of float or double.
template <typename T>
std::vector<double> widened(const std::vector<T>& v)
{
    return std::vector<double>(v.begin(), v.end());
}
} // namespace detail

template <typename Scalar>
SchurComplementSolverReport<Scalar> solveDampedSchurComplement(const BlockNormalEquations<Scalar>& eq, Scalar damping,
                                                               const SchurComplementSolverOptions<Scalar>& options,
                                                               SchurComplementSolverWorkspace<Scalar>& workspace,
                                                               std::vector<Scalar>* primary_step,
                                                               std::vector<Scalar>* eliminated_step)
{
    using Backend = SchurComplementLinearBackend;
    if (!primary_step || !eliminated_step || primary_step == eliminated_step)
        throw std::invalid_argument("solveDampedSchurComplement: output vectors must be non-null and distinct");
    const Backend backend = options.linearBackend;
    const bool dense = backend == Backend::DenseHip || backend == Backend::DenseCpu;
    if (backend != Backend::Hip && backend != Backend::Cpu && !dense)
        throw std::invalid_argument(
            "solveDampedSchurComplement: this library serves the Cpu, DenseCpu, Hip and DenseHip backends");
    if (options.deviceIndex >= 0) detail::checkHip(hipSetDevice(options.deviceIndex), "hipSetDevice");
    const std::size_t pd = static_cast<std::size_t>(eq._pc * eq._ps), ed = static_cast<std::size_t>(eq._ec * eq._es);
    const bool guess = options.useInitialGuess && primary_step->size() == pd;
    if (!guess) primary_step->assign(pd, Scalar(0));
    eliminated_step->assign(ed, Scalar(0));

    // the device solve is FP64 whatever Scalar is
    detail::Upload<double> B(detail::widened(eq._B)), C(detail::widened(eq._C)), gp(detail::widened(eq._gp)),
        ge(detail::widened(eq._ge)), W(detail::widened(eq._crossValues)), P(detail::widened(eq._pcValues));
    detail::Upload<Index> cp(eq._crossPrimary), ce(eq._crossEliminated), pr(eq._pcRows), pc(eq._pcCols);
    detail::Upload<double> dp(pd), de(ed);
    if (guess && pd)
    {
        const std::vector<double> start = detail::widened(*primary_step);
        detail::checkHip(hipMemcpy(dp.ptr, start.data(), sizeof(double) * pd, hipMemcpyHostToDevice), "hipMemcpy");
    }
    pm_normal_equations d{};
    d.primary_count = eq._pc;
    d.eliminated_count = eq._ec;
    d.primary_size = eq._ps;
    d.eliminated_size = eq._es;
    d.primary_diagonal = B.ptr;
    d.eliminated_diagonal = C.ptr;
    d.primary_gradient = gp.ptr;
    d.eliminated_gradient = ge.ptr;
    d.cross_count = static_cast<int64_t>(eq._crossPrimary.size());
    d.cross_primary = cp.ptr;
    d.cross_eliminated = ce.ptr;
    d.cross_values = W.ptr;
    d.primary_cross_count = static_cast<int64_t>(eq._pcRows.size());
    d.primary_cross_rows = pr.ptr;
    d.primary_cross_cols = pc.ptr;
    d.primary_cross_values = P.ptr;
    pm_schur_options o{};
    o.max_iterations = options.maxIterations;
    o.use_initial_guess = guess ? 1 : 0;
    o.relative_tolerance = static_cast<double>(options.relativeTolerance);
    o.absolute_tolerance = static_cast<double>(options.absoluteTolerance);
    o.minimum_diagonal = static_cast<double>(options.minimumDiagonal);
    o.linear_backend = dense ? PM_SCHUR_DENSE_CHOLESKY : PM_SCHUR_BLOCK_PCG;
    o.flags = (backend == Backend::Cpu ? PM_SCHUR_TOLERANCE_SUM : 0) |
              (backend == Backend::Hip && options.useMixedPrecision && std::is_same_v<Scalar, double>
                   ? PM_SCHUR_MIXED_PRECISION
                   : 0);
    pm_schur_report r{};
    const auto start = std::chrono::steady_clock::now();
    detail::check(pm_schur_solve(workspace.schur(), &d, static_cast<double>(damping), &o, dp.ptr, de.ptr, &r,
                                 workspace.solver(), nullptr));
    detail::checkHip(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
    const double solve_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    std::vector<double> host_p(pd), host_e(ed);
    if (pd) detail::checkHip(hipMemcpy(host_p.data(), dp.ptr, sizeof(double) * pd, hipMemcpyDeviceToHost), "hipMemcpy");
    if (ed) detail::checkHip(hipMemcpy(host_e.data(), de.ptr, sizeof(double) * ed, hipMemcpyDeviceToHost), "hipMemcpy");
    for (std::size_t i = 0; i < pd; ++i) (*primary_step)[i] = static_cast<Scalar>(host_p[i]);
    for (std::size_t i = 0; i < ed; ++i) (*eliminated_step)[i] = static_cast<Scalar>(host_e[i]);
    SchurComplementSolverReport<Scalar> report;
    report.linearBackend = backend;
    report.linearSolveSeconds = solve_seconds;
    report.converged = r.converged != 0;
    report.iterations = r.iterations;
    report.initialResidualNorm = static_cast<Scalar>(r.initial_residual_norm);
    report.finalResidualNorm = static_cast<Scalar>(r.final_residual_norm);
    report.schurPatternReused = r.pattern_reused != 0;
    report.schurAssemblyOnDevice = r.assembly_on_device != 0;
    report.mixedPrecisionUsed = r.mixed_precision_used != 0;
    report.message = r.message;
    hipDeviceProp_t prop{};
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
    {
        // the marketing name can be blank (some virtualised devices report
        // none); the architecture name always identifies the device
        report.deviceName = prop.name[0] != '\0' ? std::string(prop.name) : std::string(prop.gcnArchName);
        if (report.deviceName.empty()) report.deviceName = "HIP device " + std::to_string(dev);
    }
    return report;
}

template <typename Scalar>
SchurComplementSolverReport<Scalar> solveDampedSchurComplement(const BlockNormalEquations<Scalar>& equations,
                                                               Scalar damping,
                                                               const SchurComplementSolverOptions<Scalar>& options,
                                                               std::vector<Scalar>* primary_step,
                                                               std::vector<Scalar>* eliminated_step)
{
    SchurComplementSolverWorkspace<Scalar> workspace;
    return solveDampedSchurComplement(equations, damping, options, workspace, primary_step, eliminated_step);
}

} // namespace plamatrix
