// Solver entry points shared between translation units of the library (not
// part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "plamatrix_amd.h"

namespace pm
{

// Inner relative tolerance of one FP64 refinement pass.  A pass that can end
// the refinement (relres * inner within twice the target) aims at 0.9 of what
// it must reach, target / relres: the inner solve stops on its recursive
// residual, whose FP32 gap to the true one is a fraction of a percent, and
// with inner 1e-3 and target 1e-6 two passes otherwise end a hair either side
// of the target (1.0006e-6 on a 178,929-row Laplacian, then a third pass of 746
// iterations).  Every other pass keeps `inner`.
inline double refine_pass_tolerance(double inner, double relres, double target)
{
    if (!(relres > 0.0) || !(target > 0.0) || relres * inner > 2.0 * target) return inner;
    return std::min(inner, 0.9 * target / relres);
}

// pm_block_pcg with the stopping rule chosen by the caller: tol_rule 0 is
// max(abs, rel ||r0||) (the accelerated solvers), 1 is abs + rel ||r0|| (the
// reference's CPU Schur solver, block_schur_cpu_solver.h:34-35).
int block_pcg_rule(const pm_csr* a, const void* rhs, void* x, const void* inverse_blocks, int64_t block_size,
                   const pm_solver_options* options, pm_solver_report* report, pm_workspace* ws, hipStream_t stream,
                   int tol_rule);

} // namespace pm
