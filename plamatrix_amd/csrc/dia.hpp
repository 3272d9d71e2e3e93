// Diagonal (DIA) layout of a CSR matrix for the iterative solvers.
//
// A matrix whose entries lie on a few diagonals (stencil operators: the 3-D
// 7-point diffusion matrix has 7) streams its values diagonal by diagonal
// without per-entry column indices: 4 bytes per FP32 entry instead of 8, plus
// a 2-byte mask per row.  The layout is rebuilt from the CSR arrays at the
// start of every solve (two passes over the matrix), so in-place changes of
// the CSR between solves are always seen.  A row's present entries are
// accumulated in increasing column order with the same fused multiply-adds as
// the CSR row walk, so products are bitwise equal to the CSR kernels'.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pm_common.hpp"

struct pm_workspace;

namespace pm
{

constexpr int kMaxDia = 16;

struct DiaArgs
{
    int32_t ndiag;
    int32_t off[kMaxDia];  // column - row, increasing
};

struct DiaLayout
{
    bool active = false;
    DiaArgs args{};
    const void* values = nullptr;   // ndiag x rows, diagonal-major, A's dtype
    const uint16_t* mask = nullptr; // bit d of mask[row]: entry on diagonal d present
};

// Builds the layout into the workspace when the square CSR `a` (int32
// mirrors present) has at most kMaxDia distinct diagonals, strictly
// increasing columns in every row, and the diagonals' slots are at least
// 1/kDiaMaxFill full; otherwise out->active stays false.  Synchronises
// `stream` once to read the detected diagonals.
constexpr double kDiaMaxFill = 1.3;
int dia_build(pm_workspace* ws, const pm_csr* a, DiaLayout* out, hipStream_t stream);

} // namespace pm
