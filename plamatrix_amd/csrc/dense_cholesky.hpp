// Dense FP64 Cholesky solve of a symmetric positive definite system given as
// a scalar CSR matrix: the device counterpart of the reference's DenseCpu
// Schur backend (block_schur_dense_solver.h:126-214, 300-398).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "plamatrix_amd.h"

struct pm_workspace;

namespace pm
{

struct DenseCholeskyOutcome
{
    int not_positive_definite = 0;  // a pivot was not positive and finite
    double rhs_norm = 0.0;          // ||b||
    double residual_norm = 0.0;     // ||b - A x|| with A's CSR values
};

// Largest dimension served: the lower triangle is stored in a dense n x n
// column-major array (n^2 doubles) in the workspace.
constexpr int64_t kDenseCholeskyMaxRows = 65536;

// Solves A x = b (A: square FP64 CSR, lower triangle used; b, x: device FP64
// vectors, distinct).  Enqueues everything on `stream`, then synchronises it
// once to read the outcome.
int dense_cholesky_solve(const pm_csr* a, const double* b, double* x, pm_workspace* ws,
                         DenseCholeskyOutcome* outcome, hipStream_t stream);

} // namespace pm
