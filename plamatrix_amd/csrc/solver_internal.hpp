// Solver entry points shared between translation units of the library (not
// part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "plamatrix_amd.h"

namespace pm
{

// pm_block_pcg with the stopping rule chosen by the caller: tol_rule 0 is
// max(abs, rel ||r0||) (the accelerated solvers), 1 is abs + rel ||r0|| (the
// reference's CPU Schur solver, block_schur_cpu_solver.h:34-35).
int block_pcg_rule(const pm_csr* a, const void* rhs, void* x, const void* inverse_blocks, int64_t block_size,
                   const pm_solver_options* options, pm_solver_report* report, pm_workspace* ws, hipStream_t stream,
                   int tol_rule);

} // namespace pm
