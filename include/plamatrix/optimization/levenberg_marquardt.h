// Levenberg-Marquardt damping schedule (include/plamatrix/optimization/
// levenberg_marquardt.h:28-75): the caller evaluates each trial step and reports it as
// accepted or rejected; the damping for the next damped Schur solve shrinks by
// decreaseFactor after an acceptance and grows by increaseFactor after a
// rejection, clamped to [minimumDamping, maximumDamping].  Host-only; the
// damped solves themselves reuse one SchurComplementSolverWorkspace, so the
// Schur pattern is built once per topology across the loop.
#pragma once

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <type_traits>

namespace plamatrix
{

template <typename Scalar>
struct LevenbergMarquardtOptions
{
    Scalar initialDamping = Scalar(1e-3);
    Scalar minimumDamping = Scalar(1e-12);
    Scalar maximumDamping = Scalar(1e12);
    Scalar decreaseFactor = Scalar(0.3);
    Scalar increaseFactor = Scalar(10);
};

template <typename Scalar>
class LevenbergMarquardtStrategy
{
    static_assert(std::is_floating_point_v<Scalar>, "LevenbergMarquardtStrategy requires a floating-point scalar");

public:
    /// Throws std::invalid_argument unless 0 < minimum <= initial <= maximum
    /// (all finite), 0 < decreaseFactor < 1 and increaseFactor > 1.
    explicit LevenbergMarquardtStrategy(const LevenbergMarquardtOptions<Scalar>& options = {})
        : _opt(options), _damping(options.initialDamping)
    {
        const auto finite = [](Scalar v) { return std::isfinite(v); };
        const bool bounds = finite(_opt.minimumDamping) && finite(_opt.maximumDamping) &&
                            finite(_opt.initialDamping) && _opt.minimumDamping > Scalar(0) &&
                            _opt.minimumDamping <= _opt.maximumDamping && _opt.minimumDamping <= _opt.initialDamping &&
                            _opt.initialDamping <= _opt.maximumDamping;
        const bool factors = finite(_opt.decreaseFactor) && _opt.decreaseFactor > Scalar(0) &&
                             _opt.decreaseFactor < Scalar(1) && finite(_opt.increaseFactor) &&
                             _opt.increaseFactor > Scalar(1);
        if (!bounds || !factors) throw std::invalid_argument("LevenbergMarquardtStrategy: invalid damping options");
    }

    /// An accepted trial step: less damping next time.
    void acceptStep() noexcept
    {
        _damping = std::max(_opt.minimumDamping, _damping * _opt.decreaseFactor);
        ++_accepted;
    }

    /// A rejected (or non-finite) trial step: more damping next time.
    void rejectStep() noexcept
    {
        _damping = std::min(_opt.maximumDamping, _damping * _opt.increaseFactor);
        ++_rejected;
    }

    Scalar damping() const noexcept { return _damping; }
    int acceptedSteps() const noexcept { return _accepted; }
    int rejectedSteps() const noexcept { return _rejected; }

private:
    LevenbergMarquardtOptions<Scalar> _opt;
    Scalar _damping;
    int _accepted = 0;
    int _rejected = 0;
};

} // namespace plamatrix
