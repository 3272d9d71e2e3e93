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
#include "spmv_kernels.hpp"

struct pm_workspace;

namespace pm
{

constexpr int kMaxDia = 16;

// Entry (row, row + off[d]) has the value dv[slot[d] * rows + row + shift[d]],
// or 1 when slot[d] < 0.  A plain layout reads diagonal d from its own array
// (slot d, shift 0).  A mirrored layout (A exactly symmetric, checked bit for
// bit) reads a lower diagonal from the array of its upper mirror, shifted by
// off[d] rows, and an exactly-unit main diagonal from no array at all: a
// symmetric 7-point stencil with unit diagonal streams 3 value arrays
// instead of 7.
struct DiaArgs
{
    int32_t ndiag;
    int32_t off[kMaxDia];    // column - row, increasing
    int32_t slot[kMaxDia];   // value array of diagonal d (-1: every present entry is 1)
    int32_t shift[kMaxDia];  // row shift of that array (off[d] for a mirrored lower diagonal)
};

struct DiaLayout
{
    bool active = false;
    DiaArgs args{};
    const void* values = nullptr;   // ndiag x rows, diagonal-major, A's dtype
    const uint16_t* mask = nullptr; // bit d of mask[row]: entry on diagonal d present
    int streams = 0;                // value arrays the row sums read
};

// Builds the layout into `storage` when the square CSR `a` (read through its
// int32 mirrors when present, else its int64 arrays; fewer than 2^31 rows)
// has at most kMaxDia distinct diagonals, strictly
// increasing columns in every row, and the diagonals' slots are at least
// 1/kDiaMaxFill full; otherwise out->active stays false.  Synchronises
// `stream` once to read the detected diagonals.  With `unit_diagonal`, the
// same pass also reports whether every row's Jacobi inverse is exactly 1 (a
// diagonally scaled matrix), whether or not the layout is built.
// With `mirror`, a second pass checks the filled layout for exact symmetry
// (every lower entry present exactly where its mirror is, with the same bits)
// and for a main diagonal of exact ones, and switches the slots accordingly;
// this synchronises `stream` a second time.
constexpr double kDiaMaxFill = 1.3;
int dia_build(DeviceBuffer& storage, const pm_csr* a, DiaLayout* out, hipStream_t stream,
              bool* unit_diagonal = nullptr, bool mirror = false);

// A 64-bit signature of a layout (diagonals and storage), for plan keys.
unsigned long long dia_signature(const DiaLayout& d);

// Row sum of the diagonal layout: the row's present entries in increasing
// column order, acc = fma(value, x[col], acc) from 0 -- the CSR row walk's
// arithmetic.  Every load is issued before the mask is known (absent slots
// read an in-range value and a clamped gather that the mask discards).
// gather(col) yields the vector entry of column col.
// Value of diagonal d in `row` as the layout stores it (1 for a unit slot;
// rows before the start of a mirrored array read row 0, an absent entry).
template <typename S>
__device__ __forceinline__ S dia_value(int64_t row, int64_t rows, const DiaArgs& dia, const S* __restrict__ dv, int d)
{
    const int slot = dia.slot[d];
    if (slot < 0) return S(1);
    int64_t vr = row + dia.shift[d];
    vr = vr < 0 ? 0 : vr;
    return dv[slot * rows + vr];
}

template <typename S, int ND, typename Gather>
__device__ __forceinline__ S dia_row_gather(int64_t row, int64_t rows, const DiaArgs& dia, const S* __restrict__ dv,
                                            unsigned m, const Gather& gather)
{
    constexpr int kSlots = ND > 0 ? ND : kMaxDia;
    const int nd = ND > 0 ? ND : dia.ndiag;
    S vals[kSlots], xs[kSlots];
#pragma unroll
    for (int d = 0; d < kSlots; ++d)
    {
        if (d < nd)
        {
            int64_t col = row + dia.off[d];
            col = col < 0 ? 0 : (col >= rows ? rows - 1 : col);
            vals[d] = dia_value(row, rows, dia, dv, d);
            xs[d] = gather(col);
        }
    }
    S acc = S(0);
#pragma unroll
    for (int d = 0; d < kSlots; ++d)
        if (d < nd && ((m >> d) & 1u)) acc = fma_s(vals[d], xs[d], acc);
    return acc;
}

template <typename S, int ND>
__device__ __forceinline__ S dia_row(int64_t row, int64_t rows, const DiaArgs& dia, const S* __restrict__ dv,
                                     unsigned m, const S* __restrict__ x)
{
    return dia_row_gather<S, ND>(row, rows, dia, dv, m, [x](int64_t c) { return x[c]; });
}

} // namespace pm
