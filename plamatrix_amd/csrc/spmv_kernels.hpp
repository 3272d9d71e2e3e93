// CSR SpMV family for gfx950.
//
// A row is owned by an aligned group of L lanes (L = 1..64, a power of two,
// chosen per matrix by the autotuner or from the mean row length).  Lane j of
// the group streams entries start+j, start+j+L, ... so a wavefront reads
// 64/L consecutive rows' values and column indices as one contiguous span,
// then the group reduces with a fixed butterfly.  Each workgroup owns one
// contiguous range of rows; workgroups are renumbered so that every XCD walks
// a contiguous eighth of the matrix and its L2 keeps the x entries that the
// neighbouring rows gather.
#pragma once

#include <cmath>

#include "pm_common.hpp"
#include "workspace.hpp"

namespace pm
{

template <typename S>
__device__ __forceinline__ S fma_s(S a, S b, S c);
template <>
__device__ __forceinline__ float fma_s<float>(float a, float b, float c)
{
    return __builtin_fmaf(a, b, c);
}
template <>
__device__ __forceinline__ double fma_s<double>(double a, double b, double c)
{
    return __builtin_fma(a, b, c);
}

template <typename S>
__device__ __forceinline__ S scalar_epsilon();
template <>
__device__ __forceinline__ float scalar_epsilon<float>()
{
    return 1.1920928955078125e-7f;
}
template <>
__device__ __forceinline__ double scalar_epsilon<double>()
{
    return 2.220446049250313080847e-16;
}

// Contiguous row range of this workgroup (XCD-aware), in multiples of `gpb`
// rows (row groups per block).
__device__ __forceinline__ void block_rows(int64_t rows, int64_t gpb, int64_t& begin, int64_t& end)
{
    const int64_t nb = gridDim.x;
    const int64_t lb = xcd_logical_block(blockIdx.x, nb);
    const int64_t per = ceil_div(ceil_div(rows, nb), gpb) * gpb;
    begin = lb * per;
    end = begin + per < rows ? begin + per : rows;
}

inline int grid_for_rows(int64_t rows, int lanes)
{
    const int64_t gpb = kBlock / lanes;
    int64_t g = ceil_div(rows > 0 ? rows : 1, gpb);
    return static_cast<int>(g < kMaxPartials ? g : kMaxPartials);
}

// y = A x
template <typename S, typename I, int L>
__global__ __launch_bounds__(kBlock) void k_spmv(int64_t rows, const I* __restrict__ ro, const I* __restrict__ ci,
                                                 const S* __restrict__ v, const S* __restrict__ x,
                                                 S* __restrict__ y)
{
    constexpr int64_t gpb = kBlock / L;
    const int group = threadIdx.x / L;
    const int lane = threadIdx.x % L;
    int64_t begin, end;
    block_rows(rows, gpb, begin, end);
    for (int64_t row = begin + group; row < end; row += gpb)
    {
        const int64_t s = ro[row], e = ro[row + 1];
        S acc = S(0);
        for (int64_t k = s + lane; k < e; k += L) acc = fma_s(v[k], x[ci[k]], acc);
        acc = group_sum<L>(acc);
        if (lane == 0) y[row] = acc;
    }
}

// Solver initialisation / residual refresh:
//   r = b - A x, accumulated in FP64 and rounded to S;
//   JACOBI: inv[row] = 1/sum(diagonal entries) (FP64, rejecting a missing,
//           non-finite or <= eps(S)*16*max|a_row| diagonal, as jacobiInverse
//           in iterative_solver_cpu.cpp:127-160 does); bad rows recorded;
//   ZERO_P: p0[row] = 0 (first direction update starts from z).
template <typename S, typename I, int L, bool JACOBI, bool ZERO_P>
__global__ __launch_bounds__(kBlock) void k_residual(int64_t rows, const I* __restrict__ ro,
                                                     const I* __restrict__ ci, const S* __restrict__ v,
                                                     const S* __restrict__ b, const S* __restrict__ x,
                                                     S* __restrict__ r, S* __restrict__ inv, S* __restrict__ p0,
                                                     SolverState* st)
{
    if (st->done != kRunning) return;
    constexpr int64_t gpb = kBlock / L;
    const int group = threadIdx.x / L;
    const int lane = threadIdx.x % L;
    int64_t begin, end;
    block_rows(rows, gpb, begin, end);
    for (int64_t row = begin + group; row < end; row += gpb)
    {
        const int64_t s = ro[row], e = ro[row + 1];
        double acc = 0.0, diag = 0.0, scale = 0.0;
        int found = 0;
        for (int64_t k = s + lane; k < e; k += L)
        {
            const I c = ci[k];
            const double a = static_cast<double>(v[k]);
            acc = __builtin_fma(a, static_cast<double>(x[c]), acc);
            if (JACOBI)
            {
                scale = fmax(scale, fabs(a));
                if (c == row)
                {
                    diag += a;
                    found = 1;
                }
            }
        }
        acc = group_sum<L>(acc);
        if (JACOBI)
        {
            diag = group_sum<L>(diag);
#pragma unroll
            for (int off = L / 2; off > 0; off >>= 1)
            {
                scale = fmax(scale, __shfl_xor(scale, off, kWave));
                found |= __shfl_xor(found, off, kWave);
            }
        }
        if (lane == 0)
        {
            r[row] = static_cast<S>(static_cast<double>(b[row]) - acc);
            if (ZERO_P) p0[row] = S(0);
            if (JACOBI)
            {
                const double threshold = static_cast<double>(scalar_epsilon<S>()) * scale * 16.0;
                const double reciprocal = 1.0 / diag;
                const bool ok = found && isfinite(diag) && diag > threshold && isfinite(reciprocal);
                inv[row] = ok ? static_cast<S>(reciprocal) : static_cast<S>(NAN);
                if (!ok) atomicMin(&st->bad_row, static_cast<unsigned long long>(row));
            }
        }
    }
}

// First PCG kernel of an iteration:
//   p_new = z + beta * p_old (computed on the fly for every gathered entry and
//   written once by the row owner; p_old / p_new ping-pong between
//   iterations), q = A p_new, den = p_new . q (FP64); the last workgroup turns
//   den into alpha with the breakdown rules of alphaKernel
//   (iterative_solver_kernels.cuh:169-202).
template <typename S, typename I, int L>
__global__ __launch_bounds__(kBlock) void k_direction_spmv(int64_t rows, const I* __restrict__ ro,
                                                           const I* __restrict__ ci, const S* __restrict__ v,
                                                           const S* __restrict__ z, const S* __restrict__ p_old,
                                                           S* __restrict__ p_new, S* __restrict__ q,
                                                           SolverState* st, double* partials)
{
    if (st->done != kRunning) return;
    const S beta = static_cast<S>(st->beta);
    constexpr int64_t gpb = kBlock / L;
    const int group = threadIdx.x / L;
    const int lane = threadIdx.x % L;
    int64_t begin, end;
    block_rows(rows, gpb, begin, end);
    double den = 0.0;
    for (int64_t row = begin + group; row < end; row += gpb)
    {
        const int64_t s = ro[row], e = ro[row + 1];
        S acc = S(0);
        for (int64_t k = s + lane; k < e; k += L)
        {
            const I c = ci[k];
            acc = fma_s(v[k], fma_s(beta, p_old[c], z[c]), acc);
        }
        acc = group_sum<L>(acc);
        if (lane == 0)
        {
            const S pn = fma_s(beta, p_old[row], z[row]);
            p_new[row] = pn;
            q[row] = acc;
            den = __builtin_fma(static_cast<double>(pn), static_cast<double>(acc), den);
        }
    }
    double part[1] = {den};
    double total[1];
    if (last_block_total<1>(part, partials, &st->counters[0], total) && threadIdx.x == 0)
    {
        const double rho = st->rho;
        const double d = total[0];
        double alpha = 0.0;
        bool breakdown = false;
        if (rho == 0.0 && d == 0.0)
        {
            alpha = 0.0;
        }
        else if (isfinite(rho) && isfinite(d) && rho > 0.0 && d > 0.0)
        {
            alpha = rho / d;
            // the step is applied in S: it must be finite in S as well
            if (!isfinite(alpha) || !isfinite(static_cast<S>(alpha)))
            {
                breakdown = true;
                alpha = 0.0;
            }
        }
        else
        {
            breakdown = true;
        }
        st->alpha = alpha;
        if (breakdown) st->done = kBreakdown;
    }
}

// FP64 true residual partial sums: sum (b - A x)^2 and sum b^2.
template <typename S, typename V, typename I, int L>
__global__ __launch_bounds__(kBlock) void k_relres(int64_t rows, const I* __restrict__ ro, const I* __restrict__ ci,
                                                   const S* __restrict__ v, const V* __restrict__ b,
                                                   const V* __restrict__ x, double* partials, unsigned* counter,
                                                   double* out)
{
    constexpr int64_t gpb = kBlock / L;
    const int group = threadIdx.x / L;
    const int lane = threadIdx.x % L;
    int64_t begin, end;
    block_rows(rows, gpb, begin, end);
    double rr = 0.0, bb = 0.0;
    for (int64_t row = begin + group; row < end; row += gpb)
    {
        const int64_t s = ro[row], e = ro[row + 1];
        double acc = 0.0;
        for (int64_t k = s + lane; k < e; k += L)
            acc = __builtin_fma(static_cast<double>(v[k]), static_cast<double>(x[ci[k]]), acc);
        acc = group_sum<L>(acc);
        if (lane == 0)
        {
            const double bv = static_cast<double>(b[row]);
            const double res = bv - acc;
            rr = __builtin_fma(res, res, rr);
            bb = __builtin_fma(bv, bv, bb);
        }
    }
    double part[2] = {rr, bb};
    double total[2];
    if (last_block_total<2>(part, partials, counter, total) && threadIdx.x == 0)
    {
        out[0] = total[0];
        out[1] = total[1];
    }
}

// Dispatch helper: calls F<L>() for the runtime lanes-per-row value.
template <template <int> class F, typename... Args>
int dispatch_lanes(int lanes, Args&&... args)
{
    switch (lanes)
    {
    case 1: return F<1>::run(args...);
    case 2: return F<2>::run(args...);
    case 4: return F<4>::run(args...);
    case 8: return F<8>::run(args...);
    case 16: return F<16>::run(args...);
    case 32: return F<32>::run(args...);
    case 64: return F<64>::run(args...);
    default: return fail(PM_INVALID_ARGUMENT, "lanes per row must be a power of two in [1, 64], got %d", lanes);
    }
}

// Lanes per row from the mean row length (used when no tuned value exists).
inline int heuristic_lanes(int64_t rows, int64_t nnz)
{
    if (rows <= 0) return 1;
    const double mean = static_cast<double>(nnz) / static_cast<double>(rows);
    int lanes = 1;
    while (lanes < 64 && lanes < mean) lanes *= 2;
    return lanes;
}

} // namespace pm
