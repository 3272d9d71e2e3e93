// Conjugate gradients on gfx950: CG, Jacobi-PCG, block-Jacobi PCG, the
// fixed-iteration asynchronous variant and FP32/FP64 iterative refinement.
//
// One iteration is three kernels:
//   k_direction  reduce the previous update's r.r, r.z -> stop test, beta;
//                p = z + beta p (ping-pong buffers)
//   k_spmv_den   q = A p (one gather per entry), publish p.q
//   k_update     reduce p.q -> alpha; x += alpha p, r -= alpha q,
//                z = M^-1 r, publish r.r and r.z
// (with the factored preconditioner M^-1 = G^T G the update is x, r only and
// z = G^T (G r) follows as two more sparse products that publish r.r, r.z).
// Every dot product is accumulated in FP64 in a fixed order; every workgroup
// re-reduces the previous kernel's partial sums itself, so alpha, beta and the
// stopping test are derived identically everywhere on the device without
// atomics, fences or host round trips.  The host enqueues chunks of
// iterations (replayed as HIP graphs), keeps one chunk of lookahead in flight
// and reads one small status record per chunk; once the device marks the solve
// finished the remaining kernels return immediately.
//
// The iteration follows the reference's adaptive solver
// (iterative_solver.cu:87-163, 262-349; alphaKernel / betaKernel in
// iterative_solver_kernels.cuh:169-238), with the stopping rule
// ||r|| <= max(abs, rel * ||r0||) and the Jacobi diagonal rules of
// iterative_solver_cpu.cpp:127-160.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <string>

#include "dia.hpp"
#include "plan_cache.hpp"
#include "solver_internal.hpp"
#include "spmv_kernels.hpp"
#include "tuning.hpp"

namespace pm
{

enum Pre
{
    kPreNone = 0,
    kPreJacobi = 1,
    kPreBlock = 2,
    kPreFactored = 3,  // M^-1 = G^T G with G, G^T given as CSR
};

enum Mode
{
    kModeInit = 0,
    kModeStep = 1,
    kModeRefresh = 2,
};

constexpr int kMaxBlock = 16;
constexpr int kDirBlock = kBlock;               // threads of the direction / SpMV kernel
constexpr int kDirWaves = kDirBlock / kWave;
constexpr int kDirGrid = kMaxPartials;          // 2048 x 4 waves: 8 workgroups per CU, chunks dealt round-robin
constexpr int kPartialsDen = 0;                 // p.q partials (1 value)
constexpr int kPartialsRes = 2 * kMaxPartials;  // r.r, r.z partials (2 values)
constexpr int kPartialsDoubles = 10 * kMaxPartials;  // with the single-reduction regions at 2K..6K, 6K..10K

__global__ void k_state_init(SolverState* st, int32_t max_iterations, double rel_tol, double abs_tol,
                             int32_t tol_rule)
{
    st->tol_rule = tol_rule;
    st->rho[0] = 0.0;
    st->rho[1] = 0.0;
    st->alpha = 0.0;
    st->rr = 0.0;
    st->rr0 = 0.0;
    st->tol = 0.0;
    st->rel_tol = rel_tol;
    st->abs_tol = abs_tol;
    st->iter[0] = -1;
    st->iter[1] = -1;
    st->iteration = 0;
    st->max_iterations = max_iterations;
    st->done = kRunning;
    st->bad_row = ~0ull;
#pragma unroll
    for (int i = 0; i < 8; ++i) st->counters[i] = 0u;
}

// The decision at the head of iteration t (slot par = t % 2): every
// workgroup reduces the previous update's r.r and r.z partials in the same
// order and derives the same stop code and beta; workgroup 0 records them.
template <typename S>
__device__ __forceinline__ int direction_decision(SolverState* st, const double* partials, int par, int count_res,
                                                  double& beta)
{
    double res[2];
    reduce_partials<2>(partials + kPartialsRes, count_res, res);
    const double rr = res[0], rz = res[1];
    const int completed = st->iter[1 - par] + 1;
    // the accelerated solvers' max(abs, rel ||r0||), or the reference CPU
    // Schur solver's abs + rel ||r0|| (block_schur_cpu_solver.h:34-35)
    const double tol = completed != 0   ? st->tol
                       : st->tol_rule   ? st->abs_tol + st->rel_tol * sqrt(rr)
                                        : fmax(st->abs_tol, st->rel_tol * sqrt(rr));
    int code = kRunning;
    beta = 0.0;
    if (!isfinite(rr) || rr < 0.0) code = completed == 0 ? kInitNonFinite : kNonFinite;
    else if (sqrt(rr) <= tol) code = kConverged;
    else if (completed == 0 && st->bad_row != ~0ull) code = kBadDiagonal;
    else if (completed >= st->max_iterations) code = kIterationLimit;
    else if (completed > 0)
    {
        // betaKernel, iterative_solver_kernels.cuh:204-238
        const double rho_prev = st->rho[1 - par];
        if (rho_prev == 0.0 && rz == 0.0) beta = 0.0;
        else if (isfinite(rho_prev) && isfinite(rz) && rho_prev > 0.0 && rz > 0.0)
        {
            beta = rz / rho_prev;
            if (!isfinite(beta) || !isfinite(static_cast<S>(beta))) code = kBreakdown;
        }
        else code = kBreakdown;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        st->rr = rr;
        st->iteration = completed;
        st->iter[par] = completed;
        st->rho[par] = rz;
        if (completed == 0)
        {
            st->rr0 = rr;
            st->tol = tol;
        }
        if (code != kRunning) st->done = code;
    }
    return code;
}

// First kernel of iteration t (slot par = t % 2): every workgroup reduces the
// previous update's r.r and r.z partials in the same order, decides stop /
// beta exactly alike, and forms p_new = z + beta p_old for its rows.  Block 0
// records the decision in the state.
template <typename S>
__global__ __launch_bounds__(kBlock) void k_direction(int64_t rows, const S* __restrict__ z,
                                                      const S* __restrict__ p_old, S* __restrict__ p_new,
                                                      SolverState* st, double* partials, int par, int count_res)
{
    if (st->done != kRunning) return;
    // the first step's operands are loaded before the cross-workgroup
    // reduction so their latency overlaps it
    constexpr int U = 4;
    int64_t begin, end;
    block_range(rows, U * kBlock, begin, end);
    const int64_t first = begin + threadIdx.x;
    S fz[U], fp[U];
#pragma unroll
    for (int j = 0; j < U; ++j)
    {
        const int64_t i = first + j * kBlock;
        fz[j] = i < end ? z[i] : S(0);
        fp[j] = i < end ? p_old[i] : S(0);
    }
    double beta = 0.0;
    if (direction_decision<S>(st, partials, par, count_res, beta) != kRunning) return;
    const S b = static_cast<S>(beta);
    for (int64_t base = first; base < end; base += U * kBlock)
    {
        const bool pre = base == first;
#pragma unroll
        for (int j = 0; j < U; ++j)
        {
            const int64_t i = base + j * kBlock;
            if (i < end) p_new[i] = fma_s(b, pre ? fp[j] : p_old[i], pre ? fz[j] : z[i]);
        }
    }
}

// Second kernel of iteration t: q = A p_new and the per-workgroup partial
// p_new.q.  NT streams the matrix with non-temporal loads.
template <typename S, typename I, bool NT, bool DMA>
__global__ __launch_bounds__(kDirBlock) void k_spmv_den(int64_t rows, const I* __restrict__ ro,
                                                        const I* __restrict__ ci, const S* __restrict__ v,
                                                        const S* __restrict__ p, S* __restrict__ q,
                                                        const SolverState* st, double* partials)
{
    if (st->done != kRunning) return;
    __shared__ S lv_all[kDirWaves][kStreamCap];
    __shared__ I lc_all[kDirWaves][kStreamCap];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    double den = 0.0;
    // XCD-contiguous row ranges, as in k_spmv_den_dia (the p.q partials of
    // both kernels cover the same rows per workgroup)
    const int64_t chunks = chunk_count(rows);
    const int64_t per = ceil_div(chunks, static_cast<int64_t>(gridDim.x));
    const int64_t cbegin = xcd_logical_block(blockIdx.x, gridDim.x) * per;
    const int64_t cend = cbegin + per < chunks ? cbegin + per : chunks;
    for (int64_t c = cbegin + wave; c < cend; c += kDirWaves)
    {
        const int64_t row0 = c * kChunkRows;
        int64_t s, e;
        const S acc = chunk_row_sums<S, I, SpmvSource<S>, NT, DMA>(row0, rows, ro, ci, lv_all[wave], lc_all[wave],
                                                              SpmvSource<S>{v, p}, s, e);
        const int64_t row = row0 + lane;
        if (row < rows)
        {
            q[row] = acc;
            den = __builtin_fma(static_cast<double>(p[row]), static_cast<double>(acc), den);
        }
    }
    double part[1] = {den};
    publish_partials<1>(part, partials + kPartialsDen);
}

// k_spmv_den with kLongLanes lanes per row, for matrices whose rows are long
// and few (a 12,000-row reduced Schur system with ~110 entries per row is 188
// chunks: a wavefront per chunk leaves most of the GPU idle); the lanes of a
// row sum strided entries and then combine them, the p.q partial is that of
// the workgroup's rows.
constexpr int kLongLanes = 8;
constexpr int64_t kLongRowEntries = 32;  // average entries per row from which the lanes kernel serves

template <typename S, typename I>
__global__ __launch_bounds__(kDirBlock) void k_spmv_den_lanes(int64_t rows, const I* __restrict__ ro,
                                                              const I* __restrict__ ci, const S* __restrict__ v,
                                                              const S* __restrict__ p, S* __restrict__ q,
                                                              const SolverState* st, double* partials)
{
    if (st->done != kRunning) return;
    constexpr int64_t gpb = kDirBlock / kLongLanes;
    const int group = threadIdx.x / kLongLanes;
    const int lane = threadIdx.x % kLongLanes;
    int64_t begin, end;
    block_range(rows, gpb, begin, end);
    double den = 0.0;
    for (int64_t row = begin + group; row < end; row += gpb)
    {
        const int64_t s0 = ro[row], e0 = ro[row + 1];
        S acc = S(0);
        for (int64_t k = s0 + lane; k < e0; k += kLongLanes) acc = fma_s(v[k], p[ci[k]], acc);
        acc = group_sum<kLongLanes>(acc);
        if (lane == 0)
        {
            q[row] = acc;
            den = __builtin_fma(static_cast<double>(p[row]), static_cast<double>(acc), den);
        }
    }
    double part[1] = {den};
    publish_partials<1>(part, partials + kPartialsDen);
}

// k_residual without the Jacobi inverse, kLongLanes lanes per row (the
// long-row systems of k_spmv_den_lanes): each lane accumulates its strided
// entries in FP64, then the lanes combine.
template <typename S, typename I, bool ZERO_P>
__global__ __launch_bounds__(kBlock) void k_residual_lanes(int64_t rows, const I* __restrict__ ro,
                                                           const I* __restrict__ ci, const S* __restrict__ v,
                                                           const S* __restrict__ b, const S* __restrict__ x,
                                                           S* __restrict__ r, S* __restrict__ p0, SolverState* st)
{
    if (st->done != kRunning) return;
    constexpr int64_t gpb = kBlock / kLongLanes;
    const int group = threadIdx.x / kLongLanes;
    const int lane = threadIdx.x % kLongLanes;
    int64_t begin, end;
    block_range(rows, gpb, begin, end);
    for (int64_t row = begin + group; row < end; row += gpb)
    {
        const int64_t s0 = ro[row], e0 = ro[row + 1];
        double acc = 0.0;
        for (int64_t k = s0 + lane; k < e0; k += kLongLanes)
            acc = __builtin_fma(static_cast<double>(v[k]), static_cast<double>(x[ci[k]]), acc);
        acc = group_sum<kLongLanes>(acc);
        if (lane == 0)
        {
            r[row] = static_cast<S>(static_cast<double>(b[row]) - acc);
            if (ZERO_P) p0[row] = S(0);
        }
    }
}

// Chunks a wave of a diagonal-layout SpMV keeps in flight: one chunk's loads
// (mask, value arrays, gathers) do not cover the latency of the 256 MB
// last-level cache the iteration's working set lives in.
constexpr int kDiaInFlight = 2;

// k_spmv_den on the diagonal layout: a lane per row adds the row's present
// diagonals in increasing column order with the CSR row walk's fused
// multiply-adds, so q and p.q are bitwise those of k_spmv_den.
template <typename S, int ND>
__global__ __launch_bounds__(kDirBlock) void k_spmv_den_dia(int64_t rows, DiaArgs dia, const S* __restrict__ dv,
                                                            const uint16_t* __restrict__ mask,
                                                            const S* __restrict__ p, S* __restrict__ q,
                                                            const SolverState* st, double* partials)
{
    if (st->done != kRunning) return;
    const int lane = threadIdx.x % kWave;
    double den = 0.0;
    // each XCD walks one contiguous range of rows, so the gathered p entries
    // (the same rows plus the diagonals' reach) are fetched into its L2 once
    const int64_t chunks = chunk_count(rows);
    const int64_t per = ceil_div(chunks, static_cast<int64_t>(gridDim.x));
    const int64_t cbegin = xcd_logical_block(blockIdx.x, gridDim.x) * per;
    const int64_t cend = cbegin + per < chunks ? cbegin + per : chunks;
    for (int64_t c0 = cbegin + threadIdx.x / kWave; c0 < cend; c0 += kDiaInFlight * kDirWaves)
    {
        int64_t rw[kDiaInFlight];
        bool ok[kDiaInFlight];
        S acc[kDiaInFlight], ps[kDiaInFlight];
#pragma unroll
        for (int k = 0; k < kDiaInFlight; ++k)
        {
            const int64_t c = c0 + k * kDirWaves;
            const int64_t row = c * kChunkRows + lane;
            ok[k] = c < cend && row < rows;
            rw[k] = ok[k] ? row : rows - 1;
            acc[k] = dia_row<S, ND>(rw[k], rows, dia, dv, mask[rw[k]], p);
            ps[k] = p[rw[k]];
        }
#pragma unroll
        for (int k = 0; k < kDiaInFlight; ++k)
        {
            if (!ok[k]) continue;
            q[rw[k]] = acc[k];
            den = __builtin_fma(static_cast<double>(ps[k]), static_cast<double>(acc[k]), den);
        }
    }
    double part[1] = {den};
    publish_partials<1>(part, partials + kPartialsDen);
}

// k_direction and k_spmv_den_dia in one kernel: after the decision every
// workgroup forms p_new = z + beta p_old on the fly for its gathers (the same
// fused multiply-add as k_direction), writes p_new for its own rows and
// publishes p_new.q, so the iteration is bitwise the unfused one with one
// kernel and one pass over z and p_old fewer.
template <typename S, int ND>
__global__ __launch_bounds__(kDirBlock) void k_dir_spmv_dia(int64_t rows, DiaArgs dia, const S* __restrict__ dv,
                                                            const uint16_t* __restrict__ mask,
                                                            const S* __restrict__ z, const S* __restrict__ p_old,
                                                            S* __restrict__ p_new, S* __restrict__ q,
                                                            SolverState* st, double* partials, int par, int count_res)
{
    if (st->done != kRunning) return;
    double beta = 0.0;
    if (direction_decision<S>(st, partials, par, count_res, beta) != kRunning) return;
    const S b = static_cast<S>(beta);
    const int lane = threadIdx.x % kWave;
    double den = 0.0;
    const int64_t chunks = chunk_count(rows);
    const int64_t per = ceil_div(chunks, static_cast<int64_t>(gridDim.x));
    const int64_t cbegin = xcd_logical_block(blockIdx.x, gridDim.x) * per;
    const int64_t cend = cbegin + per < chunks ? cbegin + per : chunks;
    for (int64_t c = cbegin + threadIdx.x / kWave; c < cend; c += kDirWaves)
    {
        const int64_t row = c * kChunkRows + lane;
        if (row < rows)
        {
            const S acc = dia_row_gather<S, ND>(row, rows, dia, dv, mask[row],
                                                [=](int64_t c) { return fma_s(b, p_old[c], z[c]); });
            const S pn = fma_s(b, p_old[row], z[row]);
            p_new[row] = pn;
            q[row] = acc;
            den = __builtin_fma(static_cast<double>(pn), static_cast<double>(acc), den);
        }
    }
    double part[1] = {den};
    publish_partials<1>(part, partials + kPartialsDen);
}

// The block-Jacobi update with a lane per row: a wavefront holds
// kWave / bsize whole diagonal blocks (its last kWave % bsize lanes idle), so
// every lane reads its block's r (and q) before any lane of the wavefront
// writes r back; row j of the block's inverse is read by its own lane
// (coalesced).  Each row's x, r and z take the arithmetic of k_update's block
// path; the r.r, r.z partials group the rows per workgroup.
template <typename S, int MODE>
__global__ __launch_bounds__(kBlock) void k_update_blocks(int64_t blocks, int64_t bsize, S* __restrict__ x,
                                                          S* __restrict__ r, S* __restrict__ z,
                                                          const S* __restrict__ p, const S* __restrict__ q,
                                                          const S* __restrict__ inv, SolverState* st,
                                                          double* partials, int par, int count_den)
{
    if (st->done != kRunning) return;
    S alpha = S(0);
    if (MODE == kModeStep)
    {
        double den[1];
        reduce_partials<1>(partials + kPartialsDen, count_den, den);
        double a = 0.0;
        const bool ok = step_length<S>(st->rho[par], den[0], a);
        if (blockIdx.x == 0 && threadIdx.x == 0)
        {
            st->alpha = a;
            if (!ok) st->done = kBreakdown;
        }
        if (!ok) return;
        alpha = static_cast<S>(a);
    }
    const int lane = threadIdx.x % kWave;
    const int64_t bpw = kWave / bsize;
    const int64_t local = lane / bsize;  // block of the wavefront (bpw for an idle lane)
    const int j = static_cast<int>(lane % bsize);
    const int64_t waves = static_cast<int64_t>(gridDim.x) * (kBlock / kWave);
    double rr = 0.0, rz = 0.0;
    for (int64_t w = static_cast<int64_t>(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave;
         w * bpw < blocks; w += waves)
    {
        const int64_t u = w * bpw + local;
        const bool active = local < bpw && u < blocks;
        const int64_t base = (active ? u : 0) * bsize;
        S ri[kMaxBlock];
        for (int c = 0; c < bsize; ++c)
        {
            const int64_t i = base + c;
            ri[c] = MODE == kModeStep ? fma_s(-alpha, q[i], r[i]) : r[i];
        }
        const S* m = inv + (active ? u : 0) * bsize * bsize + j * bsize;
        S zj = S(0);
        for (int c = 0; c < bsize; ++c) zj = fma_s(m[c], ri[c], zj);
        // every lane of the wavefront has read its block's r above
        if (active)
        {
            const int64_t i = base + j;
            if (MODE == kModeStep)
            {
                x[i] = fma_s(alpha, p[i], x[i]);
                r[i] = ri[j];
            }
            z[i] = zj;
            rr = __builtin_fma(static_cast<double>(ri[j]), static_cast<double>(ri[j]), rr);
            rz = __builtin_fma(static_cast<double>(ri[j]), static_cast<double>(zj), rz);
        }
    }
    double part[2] = {rr, rz};
    publish_partials<2>(part, partials + kPartialsRes);
}

// Second kernel of iteration t (STEP), or the preconditioning half of the
// initialisation (INIT) and of a residual refresh (REFRESH).  A unit is one
// row (None / Jacobi) or one diagonal block of bsize rows (Block).
template <typename S, int PRE, int MODE>
__global__ __launch_bounds__(kBlock) void k_update(int64_t units, int64_t bsize, S* __restrict__ x,
                                                   S* __restrict__ r, S* __restrict__ z, const S* __restrict__ p,
                                                   const S* __restrict__ q, const S* __restrict__ inv,
                                                   SolverState* st, double* partials, int par, int count_den)
{
    if (st->done != kRunning) return;
    // rows (None / Jacobi): the first step's operands are loaded before the
    // cross-workgroup reduction so their latency overlaps it
    constexpr int U = 4;
    int64_t ubegin = 0, uend = 0;
    S fr[U], fp[U], fq[U], fx[U], fi[U];
    if (PRE != kPreBlock)
    {
        block_range(units, U * kBlock, ubegin, uend);
#pragma unroll
        for (int j = 0; j < U; ++j)
        {
            const int64_t u = ubegin + threadIdx.x + j * kBlock;
            fr[j] = fp[j] = fq[j] = fx[j] = fi[j] = S(0);
            if (u < uend)
            {
                fr[j] = r[u];
                if (MODE == kModeStep)
                {
                    fp[j] = p[u];
                    fq[j] = q[u];
                    fx[j] = x[u];
                }
                if (PRE == kPreJacobi) fi[j] = inv[u];
            }
        }
    }
    S alpha = S(0);
    if (MODE == kModeStep)
    {
        double den[1];
        reduce_partials<1>(partials + kPartialsDen, count_den, den);
        double a = 0.0;
        const bool ok = step_length<S>(st->rho[par], den[0], a);
        if (blockIdx.x == 0 && threadIdx.x == 0)
        {
            st->alpha = a;
            if (!ok) st->done = kBreakdown;
        }
        if (!ok) return;
        alpha = static_cast<S>(a);
    }
    double rr = 0.0, rz = 0.0;
    if (PRE == kPreBlock)
    {
        int64_t begin, end;
        block_range(units, kBlock, begin, end);
        for (int64_t u = begin + threadIdx.x; u < end; u += kBlock)
        {
            S ri[kMaxBlock];
            const int64_t base = u * bsize;
            for (int j = 0; j < bsize; ++j)
            {
                const int64_t i = base + j;
                if (MODE == kModeStep)
                {
                    x[i] = fma_s(alpha, p[i], x[i]);
                    ri[j] = fma_s(-alpha, q[i], r[i]);
                    r[i] = ri[j];
                }
                else
                {
                    ri[j] = r[i];
                }
            }
            const S* m = inv + u * bsize * bsize;
            for (int j = 0; j < bsize; ++j)
            {
                S zj = S(0);
                for (int c = 0; c < bsize; ++c) zj = fma_s(m[j * bsize + c], ri[c], zj);
                z[base + j] = zj;
                rr = __builtin_fma(static_cast<double>(ri[j]), static_cast<double>(ri[j]), rr);
                rz = __builtin_fma(static_cast<double>(ri[j]), static_cast<double>(zj), rz);
            }
        }
    }
    else
    {
        // four independent coalesced units per thread per step; the first
        // step uses the operands loaded before the reduction
        const int64_t first = ubegin + threadIdx.x;
        for (int64_t base = first; base < uend; base += U * kBlock)
        {
            S rv[U], zv[U];
#pragma unroll
            for (int j = 0; j < U; ++j)
            {
                const int64_t u = base + j * kBlock;
                if (u < uend)
                {
                    const bool pre = base == first;
                    S ri = pre ? fr[j] : r[u];
                    if (MODE == kModeStep)
                    {
                        const S pu = pre ? fp[j] : p[u], qu = pre ? fq[j] : q[u], xu = pre ? fx[j] : x[u];
                        x[u] = fma_s(alpha, pu, xu);
                        ri = fma_s(-alpha, qu, ri);
                        r[u] = ri;
                    }
                    rv[j] = ri;
                    zv[j] = PRE == kPreJacobi ? (pre ? fi[j] : inv[u]) * ri : ri;
                    if (PRE == kPreJacobi) z[u] = zv[j];
                }
                else
                {
                    rv[j] = S(0);
                    zv[j] = S(0);
                }
            }
#pragma unroll
            for (int j = 0; j < U; ++j)
            {
                rr = __builtin_fma(static_cast<double>(rv[j]), static_cast<double>(rv[j]), rr);
                rz = __builtin_fma(static_cast<double>(rv[j]), static_cast<double>(zv[j]), rz);
            }
        }
    }
    double part[2] = {rr, rz};
    publish_partials<2>(part, partials + kPartialsRes);
}

// First half of a refresh iteration: alpha from p.q, x += alpha p.
template <typename S>
__global__ __launch_bounds__(kBlock) void k_axpy_x(int64_t rows, S* __restrict__ x, const S* __restrict__ p,
                                                   SolverState* st, double* partials, int par, int count_den)
{
    if (st->done != kRunning) return;
    double den[1];
    reduce_partials<1>(partials + kPartialsDen, count_den, den);
    double a = 0.0;
    const bool ok = step_length<S>(st->rho[par], den[0], a);
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        st->alpha = a;
        if (!ok) st->done = kBreakdown;
    }
    if (!ok) return;
    const S alpha = static_cast<S>(a);
    int64_t begin, end;
    block_range(rows, kBlock, begin, end);
    for (int64_t i = begin + threadIdx.x; i < end; i += kBlock) x[i] = fma_s(alpha, p[i], x[i]);
}

// alpha = rho / p.q from the published p.q partials, recorded in the state
// by workgroup 0 (with the breakdown flag when it is not usable).  Every
// workgroup reduces the same partials in the same order, so all agree.
template <typename S>
__device__ __forceinline__ bool step_alpha(SolverState* st, const double* partials, int par, int count_den, S& alpha)
{
    double den[1];
    reduce_partials<1>(partials + kPartialsDen, count_den, den);
    double a = 0.0;
    const bool ok = step_length<S>(st->rho[par], den[0], a);
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        st->alpha = a;
        if (!ok) st->done = kBreakdown;
    }
    alpha = static_cast<S>(a);
    return ok;
}

// Factored preconditioner, step half: alpha from p.q, x += alpha p,
// r -= alpha q (z follows from the two sparse products).
template <typename S>
__global__ __launch_bounds__(kBlock) void k_step_xr(int64_t rows, S* __restrict__ x, S* __restrict__ r,
                                                    const S* __restrict__ p, const S* __restrict__ q, SolverState* st,
                                                    double* partials, int par, int count_den)
{
    if (st->done != kRunning) return;
    S alpha;
    if (!step_alpha<S>(st, partials, par, count_den, alpha)) return;
    int64_t begin, end;
    block_range(rows, kBlock, begin, end);
    for (int64_t i = begin + threadIdx.x; i < end; i += kBlock)
    {
        x[i] = fma_s(alpha, p[i], x[i]);
        r[i] = fma_s(-alpha, q[i], r[i]);
    }
}

// t = G r (G lower, CSR), skipped once the solve has stopped.
template <typename S, typename I, bool DMA>
__global__ __launch_bounds__(kBlock) void k_factor_lower(int64_t rows, const I* __restrict__ ro,
                                                         const I* __restrict__ ci, const S* __restrict__ v,
                                                         const S* __restrict__ r, S* __restrict__ t,
                                                         const SolverState* st)
{
    if (st->done != kRunning) return;
    __shared__ S lv_all[kWavesPerBlock][kStreamCap];
    __shared__ I lc_all[kWavesPerBlock][kStreamCap];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    PM_FOR_CHUNKS(c, chunk_count(rows), kWavesPerBlock)
    {
        const int64_t row0 = c * kChunkRows;
        int64_t s, e;
        const S acc = chunk_row_sums<S, I, SpmvSource<S>, false, DMA>(row0, rows, ro, ci, lv_all[wave], lc_all[wave],
                                                                      SpmvSource<S>{v, r}, s, e);
        if (row0 + lane < rows) t[row0 + lane] = acc;
    }
}

// k_step_xr and k_factor_lower in one pass: alpha, x += alpha p on the
// workgroup's rows, and t = G r_new with r_new = r - alpha q formed at every
// gathered column (the same rounding as k_step_xr's r).  r itself is left for
// k_factor_upper<STEP=true> to rewrite, since other workgroups gather it here.
template <typename S, typename I, bool DMA>
__global__ __launch_bounds__(kBlock) void k_step_lower(int64_t rows, const I* __restrict__ ro,
                                                       const I* __restrict__ ci, const S* __restrict__ v,
                                                       S* __restrict__ x, const S* __restrict__ r,
                                                       const S* __restrict__ p, const S* __restrict__ q,
                                                       S* __restrict__ t, SolverState* st, double* partials, int par,
                                                       int count_den)
{
    if (st->done != kRunning) return;
    S alpha;
    if (!step_alpha<S>(st, partials, par, count_den, alpha)) return;
    __shared__ S lv_all[kWavesPerBlock][kStreamCap];
    __shared__ I lc_all[kWavesPerBlock][kStreamCap];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    PM_FOR_CHUNKS(c, chunk_count(rows), kWavesPerBlock)
    {
        const int64_t row0 = c * kChunkRows;
        int64_t s, e;
        const S acc = chunk_row_sums<S, I, StepSource<S>, false, DMA>(row0, rows, ro, ci, lv_all[wave], lc_all[wave],
                                                                      StepSource<S>{v, r, q, -alpha}, s, e);
        const int64_t row = row0 + lane;
        if (row < rows)
        {
            t[row] = acc;
            x[row] = fma_s(alpha, p[row], x[row]);
        }
    }
}

// z = G^T t (G^T as CSR) and the partial sums r.r, r.z of the next direction.
// STEP: r -= alpha q first, on this row only (alpha from the state).
template <typename S, typename I, bool DMA, bool STEP = false>
__global__ __launch_bounds__(kBlock) void k_factor_upper(int64_t rows, const I* __restrict__ ro,
                                                         const I* __restrict__ ci, const S* __restrict__ v,
                                                         const S* __restrict__ t, S* __restrict__ r,
                                                         const S* __restrict__ q, S* __restrict__ z,
                                                         const SolverState* st, double* partials)
{
    if (st->done != kRunning) return;
    const S neg_alpha = STEP ? -static_cast<S>(st->alpha) : S(0);
    __shared__ S lv_all[kWavesPerBlock][kStreamCap];
    __shared__ I lc_all[kWavesPerBlock][kStreamCap];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    double rr = 0.0, rz = 0.0;
    PM_FOR_CHUNKS(c, chunk_count(rows), kWavesPerBlock)
    {
        const int64_t row0 = c * kChunkRows;
        int64_t s, e;
        const S acc = chunk_row_sums<S, I, SpmvSource<S>, false, DMA>(row0, rows, ro, ci, lv_all[wave], lc_all[wave],
                                                                      SpmvSource<S>{v, t}, s, e);
        const int64_t row = row0 + lane;
        if (row < rows)
        {
            z[row] = acc;
            S rs = r[row];
            if constexpr (STEP)
            {
                rs = fma_s(neg_alpha, q[row], rs);
                r[row] = rs;
            }
            const double rv = static_cast<double>(rs);
            rr = __builtin_fma(rv, rv, rr);
            rz = __builtin_fma(rv, static_cast<double>(acc), rz);
        }
    }
    double part[2] = {rr, rz};
    publish_partials<2>(part, partials + kPartialsRes);
}

// k_factor_lower / k_factor_upper on the diagonal layouts of G and G^T (same
// chunks per workgroup, same row arithmetic, so the same bits).
template <typename S, int ND>
__global__ __launch_bounds__(kBlock) void k_factor_lower_dia(int64_t rows, DiaArgs dia, const S* __restrict__ dv,
                                                             const uint16_t* __restrict__ mask,
                                                             const S* __restrict__ r, S* __restrict__ t,
                                                             const SolverState* st)
{
    if (st->done != kRunning) return;
    const int lane = threadIdx.x % kWave;
    PM_FOR_CHUNKS(c, chunk_count(rows), kWavesPerBlock)
    {
        const int64_t row = c * kChunkRows + lane;
        if (row < rows) t[row] = dia_row<S, ND>(row, rows, dia, dv, mask[row], r);
    }
}

// k_step_lower on the diagonal layout of G.
template <typename S, int ND>
__global__ __launch_bounds__(kBlock) void k_step_lower_dia(int64_t rows, DiaArgs dia, const S* __restrict__ dv,
                                                           const uint16_t* __restrict__ mask, S* __restrict__ x,
                                                           const S* __restrict__ r, const S* __restrict__ p,
                                                           const S* __restrict__ q, S* __restrict__ t,
                                                           SolverState* st, double* partials, int par, int count_den)
{
    if (st->done != kRunning) return;
    S alpha;
    if (!step_alpha<S>(st, partials, par, count_den, alpha)) return;
    const S neg_alpha = -alpha;
    const int lane = threadIdx.x % kWave;
    PM_FOR_CHUNKS(c, chunk_count(rows), kWavesPerBlock)
    {
        const int64_t row = c * kChunkRows + lane;
        if (row < rows)
        {
            t[row] = dia_row_gather<S, ND>(row, rows, dia, dv, mask[row],
                                           [=](int64_t col) { return fma_s(neg_alpha, q[col], r[col]); });
            x[row] = fma_s(alpha, p[row], x[row]);
        }
    }
}

template <typename S, int ND, bool STEP = false>
__global__ __launch_bounds__(kBlock) void k_factor_upper_dia(int64_t rows, DiaArgs dia, const S* __restrict__ dv,
                                                             const uint16_t* __restrict__ mask,
                                                             const S* __restrict__ t, S* __restrict__ r,
                                                             const S* __restrict__ q, S* __restrict__ z,
                                                             const SolverState* st, double* partials)
{
    if (st->done != kRunning) return;
    const S neg_alpha = STEP ? -static_cast<S>(st->alpha) : S(0);
    const int lane = threadIdx.x % kWave;
    double rr = 0.0, rz = 0.0;
    PM_FOR_CHUNKS(c, chunk_count(rows), kWavesPerBlock)
    {
        const int64_t row = c * kChunkRows + lane;
        if (row < rows)
        {
            const S acc = dia_row<S, ND>(row, rows, dia, dv, mask[row], t);
            z[row] = acc;
            S rs = r[row];
            if constexpr (STEP)
            {
                rs = fma_s(neg_alpha, q[row], rs);
                r[row] = rs;
            }
            const double rv = static_cast<double>(rs);
            rr = __builtin_fma(rv, rv, rr);
            rz = __builtin_fma(rv, static_cast<double>(acc), rz);
        }
    }
    double part[2] = {rr, rz};
    publish_partials<2>(part, partials + kPartialsRes);
}

// ---- single-reduction iteration (Chronopoulos & Gear, 1989) ---------------------
//
// CG / Jacobi-PCG with one reduction phase per iteration.  Two kernels on CSR:
//   k_cg_step  reduce r.r, gamma = r.u, delta = w.u and u.s of the previous
//              iteration; stop test on ||r||; beta = gamma / gamma_prev
//              (betaKernel's rules), alpha = gamma / den with den = p.Ap of the
//              new direction (alphaKernel's rules applied to it; a
//              non-positive den restarts the step from u); p = u + beta p,
//              s = w + beta s, x += alpha p, r -= alpha s, u = M^-1 r
//   k_cg_spmv  w = A u and the partials r.r, (r.u,) w.u, u.s
// The recurrence s = A p replaces the SpMV of the direction, so the
// iteration streams the same bytes as the three-kernel form with one kernel
// boundary and one cross-workgroup reduction fewer.  u is r itself without a
// preconditioner (r.u is then r.r and is not published twice), M^-1 r with
// Jacobi.
//
// On the diagonal layout one kernel does both (k_cg_fused_dia): after the
// decision every workgroup forms the new r, s, u of its own rows and, for the
// product w = A u of the new u, recomputes the new u of each neighbouring row
// from that row's r, w and s with the same fused multiply-adds its owner
// uses.  r, w, s (and u) alternate between two buffers with the iteration
// parity, so the old values a neighbour reads are never overwritten in the
// launch; p and x are only read at their own rows and stay in place.  A
// residual-refresh iteration runs the two-kernel form on the same buffers.
//
// den expands p.Ap for p = u + beta p_prev, s_prev = A p_prev and A
// symmetric: den = w.u + 2 beta u.s_prev + beta^2 p_prev.s_prev, with
// p_prev.s_prev taken as the previous den (the same expansion one step
// earlier).  The shorter textbook form delta - beta gamma / alpha_prev also
// assumes u.r_prev = 0, which FP32 rounding and every residual refresh break;
// on ill-conditioned systems it then stalls convergence (relres 0.07 after
// 20,000 iterations of a refreshed FP32 solve that the three-kernel form ends
// in 1,660), while the expanded form stays on the three-kernel iteration's
// count.  The SpMV kernels sum u.s with the other partials, over the same
// chunks in the same order as the one-kernel iteration, so the two forms
// (and the CSR and diagonal layouts) stay bitwise equal.
//
// Partials: iteration t publishes r.r, r.u, w.u, u.s (kCgValues slots of
// kMaxPartials each) into region t % 2 and its decision reads region
// (t - 1) % 2.  Region 1 is the three-kernel form's r.r / r.z area, so after
// an odd iteration (every chunk of iterations ends on one) a solve can
// continue in that form.

constexpr int kCounterRestart = 0;  // SolverState::counters slot: a single-reduction restart happened
constexpr int kCgValues = 4;        // r.r, r.u, w.u, u.s

__host__ __device__ inline double* cg_region(double* partials, int par)
{
    return partials + kPartialsRes + (par == 1 ? 0 : kCgValues * kMaxPartials);
}

// The decision at the head of single-reduction iteration t (slot par), from
// the partials `in` of iteration t - 1.
template <typename S, int PRE>
__device__ __forceinline__ int cg_decision(SolverState* st, const double* in, int par, int count, double& alpha,
                                           double& beta)
{
    // count is the solver grid (at most kSolverGrid workgroups), which halves
    // the registers the loads hold next to the step's prefetched operands
    constexpr int NV = PRE == kPreJacobi ? 4 : 3;
    double red[NV];
    if constexpr (PRE == kPreJacobi)
    {
        const double* base[NV] = {in, in + kMaxPartials, in + 2 * kMaxPartials, in + 3 * kMaxPartials};
        reduce_partials_at<NV, kSolverGrid>(base, count, red);
    }
    else
    {
        const double* base[NV] = {in, in + 2 * kMaxPartials, in + 3 * kMaxPartials};
        reduce_partials_at<NV, kSolverGrid>(base, count, red);
    }
    const double rr = red[0];
    const double gamma = red[PRE == kPreJacobi ? 1 : 0];
    const double delta = red[NV - 2];
    const double us = red[NV - 1];
    const int completed = st->iter[1 - par] + 1;
    const double tol = completed != 0   ? st->tol
                       : st->tol_rule   ? st->abs_tol + st->rel_tol * sqrt(rr)
                                        : fmax(st->abs_tol, st->rel_tol * sqrt(rr));
    int code = kRunning;
    beta = 0.0;
    alpha = 0.0;
    double den = delta;
    if (!isfinite(rr) || rr < 0.0) code = completed == 0 ? kInitNonFinite : kNonFinite;
    else if (sqrt(rr) <= tol) code = kConverged;
    else if (completed == 0 && st->bad_row != ~0ull) code = kBadDiagonal;
    else if (completed >= st->max_iterations) code = kIterationLimit;
    else
    {
        if (completed > 0)
        {
            const double gamma_prev = st->rho[1 - par];
            if (gamma_prev == 0.0 && gamma == 0.0) beta = 0.0;
            else if (isfinite(gamma_prev) && isfinite(gamma) && gamma_prev > 0.0 && gamma > 0.0)
            {
                beta = gamma / gamma_prev;
                if (!isfinite(beta) || !isfinite(static_cast<S>(beta))) code = kBreakdown;
            }
            else code = kBreakdown;
            // u.s exists only after a completed step
            if (code == kRunning && beta != 0.0) den = delta + beta * (2.0 * us + beta * st->den[1 - par]);
            // a non-positive p.Ap of the new direction restarts the step from
            // u (beta = 0, p.Ap = u.Au = delta); only a non-positive delta is
            // a breakdown of A itself
            if (code == kRunning && !(den > 0.0 && isfinite(den)))
            {
                beta = 0.0;
                den = delta;
                // tells the host to continue in the three-kernel form, which
                // forms p.Ap from a fresh product every iteration
                if (blockIdx.x == 0 && threadIdx.x == 0) st->counters[kCounterRestart] = 1u;
            }
        }
        if (code == kRunning && !step_length<S>(gamma, den, alpha)) code = kBreakdown;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        st->rr = rr;
        st->iteration = completed;
        st->iter[par] = completed;
        st->rho[par] = gamma;
        st->den[par] = den;
        st->alpha = alpha;
        if (completed == 0)
        {
            st->rr0 = rr;
            st->tol = tol;
        }
        if (code != kRunning) st->done = code;
    }
    return code;
}

// Vectors of single-reduction iteration t: iteration t's r, u, w and s_{t-1}
// (`_in`) and the new r, u, s (`_out`), which are the same arrays when the
// iteration works in place.  p and x are always updated in place.  Without a
// preconditioner u is r.
template <typename S>
struct CgVectors
{
    S* x;
    S* p;
    const S* r_in;
    const S* u_in;
    const S* w_in;
    const S* s_in;
    S* r_out;
    S* u_out;
    S* w_out;
    S* s_out;
    const S* inv;
};

// The vector half of single-reduction iteration t.  STEP_R = false (a residual
// refresh follows) leaves r and u to the refresh.
template <typename S, int PRE, bool STEP_R>
__global__ __launch_bounds__(kBlock) void k_cg_step(int64_t rows, CgVectors<S> v, SolverState* st,
                                                    const double* part_in, int par, int count)
{
    if (st->done != kRunning) return;
    constexpr int U = 8;
    int64_t begin, end;
    block_range(rows, kBlock, begin, end);
    const int64_t first = begin + threadIdx.x;
    // the first step's operands are loaded before the cross-workgroup
    // reduction so their latency overlaps it
    S fr[U], fu[U], fp[U], fs[U], fw[U], fx[U], fi[U];
#pragma unroll
    for (int j = 0; j < U; ++j)
    {
        const int64_t i = first + j * kBlock;
        const bool in = i < end;
        fr[j] = in ? v.r_in[i] : S(0);
        fu[j] = PRE == kPreJacobi ? (in ? v.u_in[i] : S(0)) : fr[j];
        fp[j] = in ? v.p[i] : S(0);
        fs[j] = in ? v.s_in[i] : S(0);
        fw[j] = in ? v.w_in[i] : S(0);
        fx[j] = in ? v.x[i] : S(0);
        fi[j] = PRE == kPreJacobi && in ? v.inv[i] : S(0);
    }
    double alpha_d = 0.0, beta_d = 0.0;
    if (cg_decision<S, PRE>(st, part_in, par, count, alpha_d, beta_d) != kRunning) return;
    const S a = static_cast<S>(alpha_d), b = static_cast<S>(beta_d);
    for (int64_t base = first; base < end; base += U * kBlock)
    {
        const bool pre = base == first;
#pragma unroll
        for (int j = 0; j < U; ++j)
        {
            const int64_t i = base + j * kBlock;
            if (i >= end) continue;
            const S ri = pre ? fr[j] : v.r_in[i];
            const S ui = PRE == kPreJacobi ? (pre ? fu[j] : v.u_in[i]) : ri;
            const S pn = fma_s(b, pre ? fp[j] : v.p[i], ui);
            const S sn = fma_s(b, pre ? fs[j] : v.s_in[i], pre ? fw[j] : v.w_in[i]);
            v.p[i] = pn;
            v.s_out[i] = sn;
            v.x[i] = fma_s(a, pn, pre ? fx[j] : v.x[i]);
            if (STEP_R)
            {
                const S rn = fma_s(-a, sn, ri);
                v.r_out[i] = rn;
                if (PRE == kPreJacobi) v.u_out[i] = (pre ? fi[j] : v.inv[i]) * rn;
            }
        }
    }
}

// u = M^-1 r after the residual refresh of iteration t (Jacobi only).
template <typename S>
__global__ __launch_bounds__(kBlock) void k_cg_precondition(int64_t rows, CgVectors<S> v, const SolverState* st)
{
    if (st->done != kRunning) return;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < rows;
         i += static_cast<int64_t>(gridDim.x) * kBlock)
        v.u_out[i] = v.inv[i] * v.r_out[i];
}

// Initialisation after the residual: s = 0 (its first update is s = w), and
// with Jacobi u = M^-1 r.
template <typename S, int PRE>
__global__ __launch_bounds__(kBlock) void k_cg_init(int64_t rows, const S* __restrict__ r, const S* __restrict__ inv,
                                                    S* __restrict__ u, S* __restrict__ s, const SolverState* st)
{
    if (st->done != kRunning) return;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < rows;
         i += static_cast<int64_t>(gridDim.x) * kBlock)
    {
        s[i] = S(0);
        if (PRE == kPreJacobi) u[i] = inv[i] * r[i];
    }
}

// w = A u (CSR) of the new u and the partials r.r, (r.u,) w.u, u.s of the
// next decision.
template <typename S, typename I, int PRE>
__global__ __launch_bounds__(kDirBlock) void k_cg_spmv(int64_t rows, const I* __restrict__ ro,
                                                       const I* __restrict__ ci, const S* __restrict__ vals,
                                                       CgVectors<S> v, const SolverState* st, double* part_out)
{
    if (st->done != kRunning) return;
    __shared__ S lv_all[kDirWaves][kStreamCap];
    __shared__ I lc_all[kDirWaves][kStreamCap];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const S* r = v.r_out;
    const S* uu = PRE == kPreJacobi ? v.u_out : r;
    double rr = 0.0, ru = 0.0, wu = 0.0, us = 0.0;
    const int64_t chunks = chunk_count(rows);
    const int64_t per = ceil_div(chunks, static_cast<int64_t>(gridDim.x));
    const int64_t cbegin = xcd_logical_block(blockIdx.x, gridDim.x) * per;
    const int64_t cend = cbegin + per < chunks ? cbegin + per : chunks;
    for (int64_t c = cbegin + wave; c < cend; c += kDirWaves)
    {
        const int64_t row0 = c * kChunkRows;
        int64_t ks, ke;
        const S acc = chunk_row_sums<S, I, SpmvSource<S>, false, false>(row0, rows, ro, ci, lv_all[wave],
                                                                        lc_all[wave], SpmvSource<S>{vals, uu}, ks, ke);
        const int64_t row = row0 + lane;
        if (row < rows)
        {
            v.w_out[row] = acc;
            const double rv = static_cast<double>(r[row]);
            const double uv = static_cast<double>(uu[row]);
            rr = __builtin_fma(rv, rv, rr);
            if (PRE == kPreJacobi) ru = __builtin_fma(rv, uv, ru);
            wu = __builtin_fma(static_cast<double>(acc), uv, wu);
            us = __builtin_fma(uv, static_cast<double>(v.s_out[row]), us);
        }
    }
    double part[4] = {rr, PRE == kPreJacobi ? ru : rr, wu, us};
    publish_partials<4>(part, part_out);
}

// k_cg_spmv on the diagonal layout.
template <typename S, int ND, int PRE>
__global__ __launch_bounds__(kDirBlock) void k_cg_spmv_dia(int64_t rows, DiaArgs dia, const S* __restrict__ dv,
                                                           const uint16_t* __restrict__ mask, CgVectors<S> v,
                                                           const SolverState* st, double* part_out)
{
    if (st->done != kRunning) return;
    const int lane = threadIdx.x % kWave;
    const S* r = v.r_out;
    const S* uu = PRE == kPreJacobi ? v.u_out : r;
    double rr = 0.0, ru = 0.0, wu = 0.0, us = 0.0;
    const int64_t chunks = chunk_count(rows);
    const int64_t per = ceil_div(chunks, static_cast<int64_t>(gridDim.x));
    const int64_t cbegin = xcd_logical_block(blockIdx.x, gridDim.x) * per;
    const int64_t cend = cbegin + per < chunks ? cbegin + per : chunks;
    // kDiaInFlight consecutive chunks of the wave per pass, every load issued
    // before the first sum (rows past the end load row rows - 1 and are
    // dropped); the partials still add the chunks in the wave's chunk order
    for (int64_t c0 = cbegin + threadIdx.x / kWave; c0 < cend; c0 += kDiaInFlight * kDirWaves)
    {
        int64_t rw[kDiaInFlight];
        bool ok[kDiaInFlight];
        S acc[kDiaInFlight], rs[kDiaInFlight], uk[kDiaInFlight], sk[kDiaInFlight];
#pragma unroll
        for (int k = 0; k < kDiaInFlight; ++k)
        {
            const int64_t c = c0 + k * kDirWaves;
            const int64_t row = c * kChunkRows + lane;
            ok[k] = c < cend && row < rows;
            rw[k] = ok[k] ? row : rows - 1;
            acc[k] = dia_row<S, ND>(rw[k], rows, dia, dv, mask[rw[k]], uu);
            rs[k] = r[rw[k]];
            uk[k] = uu[rw[k]];
            sk[k] = v.s_out[rw[k]];
        }
#pragma unroll
        for (int k = 0; k < kDiaInFlight; ++k)
        {
            if (!ok[k]) continue;
            v.w_out[rw[k]] = acc[k];
            const double rv = static_cast<double>(rs[k]);
            const double uv = static_cast<double>(uk[k]);
            rr = __builtin_fma(rv, rv, rr);
            if (PRE == kPreJacobi) ru = __builtin_fma(rv, uv, ru);
            wu = __builtin_fma(static_cast<double>(acc[k]), uv, wu);
            us = __builtin_fma(uv, static_cast<double>(sk[k]), us);
        }
    }
    double part[4] = {rr, PRE == kPreJacobi ? ru : rr, wu, us};
    publish_partials<4>(part, part_out);
}

// The operands one lane of k_cg_fused_dia loads for one row: the row's
// diagonal values and, per diagonal, the old r, w, s (and inverse diagonal)
// of the neighbouring row, plus the row's own vectors.  None of the addresses
// depend on alpha or beta, so a pass's loads can be in flight during the
// decision.
template <typename S, int ND, int PRE>
struct CgFusedRow
{
    static constexpr int kSlots = ND > 0 ? ND : kMaxDia;
    S val[kSlots], nr[kSlots], nw[kSlots], ns[kSlots], ni[kSlots];
    S r, s, w, p, x, u, iv;
    unsigned m;

    __device__ __forceinline__ void load(int64_t i, int64_t rows, const DiaArgs& dia, const S* __restrict__ dv,
                                         const uint16_t* __restrict__ mask, const CgVectors<S>& v)
    {
        const int nd = ND > 0 ? ND : dia.ndiag;
        m = mask[i];
#pragma unroll
        for (int d = 0; d < kSlots; ++d)
        {
            if (d < nd)
            {
                int64_t c = i + dia.off[d];
                c = c < 0 ? 0 : (c >= rows ? rows - 1 : c);
                val[d] = dia_value(i, rows, dia, dv, d);
                nr[d] = v.r_in[c];
                nw[d] = v.w_in[c];
                ns[d] = v.s_in[c];
                ni[d] = PRE == kPreJacobi ? v.inv[c] : S(0);
            }
        }
        r = v.r_in[i];
        s = v.s_in[i];
        w = v.w_in[i];
        p = v.p[i];
        x = v.x[i];
        u = PRE == kPreJacobi ? v.u_in[i] : r;
        iv = PRE == kPreJacobi ? v.inv[i] : S(0);
    }

    // The row of w = A u of the new u: each present neighbour's new u formed
    // with its owner's fused multiply-adds, summed in increasing column order.
    __device__ __forceinline__ S product(S a, S b, const DiaArgs& dia) const
    {
        const int nd = ND > 0 ? ND : dia.ndiag;
        S acc = S(0);
#pragma unroll
        for (int d = 0; d < kSlots; ++d)
        {
            if (d < nd && ((m >> d) & 1u))
            {
                const S rn = fma_s(-a, fma_s(b, ns[d], nw[d]), nr[d]);
                acc = fma_s(val[d], PRE == kPreJacobi ? ni[d] * rn : rn, acc);
            }
        }
        return acc;
    }
};

// One whole single-reduction iteration on the diagonal layout: the decision,
// then for every own row the step (p, s, x, r, u) and the row of w = A u of
// the new u, whose neighbouring entries are recomputed from the old r, w, s
// of those rows; partials r.r, (r.u,) w.u, u.s.  Needs distinct _in and _out
// arrays for r, w, s (and u).  The first pass's loads are issued before the
// decision's cross-workgroup reduction.
template <typename S, int ND, int PRE>
__global__ __launch_bounds__(kDirBlock) void k_cg_fused_dia(int64_t rows, DiaArgs dia, const S* __restrict__ dv,
                                                            const uint16_t* __restrict__ mask, CgVectors<S> v,
                                                            SolverState* st, const double* part_in, double* part_out,
                                                            int par, int count)
{
    if (st->done != kRunning) return;
    // rows in flight per lane and pass (the generic diagonal count holds 16
    // slots of operands per row)
    constexpr int F = ND > 0 ? kDiaInFlight : 1;
    const int lane = threadIdx.x % kWave;
    const int64_t chunks = chunk_count(rows);
    const int64_t per = ceil_div(chunks, static_cast<int64_t>(gridDim.x));
    const int64_t cbegin = xcd_logical_block(blockIdx.x, gridDim.x) * per;
    const int64_t cend = cbegin + per < chunks ? cbegin + per : chunks;
    CgFusedRow<S, ND, PRE> row[F];
    int64_t rw[F];
    bool ok[F];
    // rows past the end load row rows - 1 and are dropped
    auto load_pass = [&](int64_t c0) {
#pragma unroll
        for (int k = 0; k < F; ++k)
        {
            const int64_t c = c0 + k * kDirWaves;
            const int64_t i = c * kChunkRows + lane;
            ok[k] = c < cend && i < rows;
            rw[k] = ok[k] ? i : rows - 1;
            row[k].load(rw[k], rows, dia, dv, mask, v);
        }
    };
    int64_t c0 = cbegin + threadIdx.x / kWave;
    load_pass(c0);
    double alpha_d = 0.0, beta_d = 0.0;
    if (cg_decision<S, PRE>(st, part_in, par, count, alpha_d, beta_d) != kRunning) return;
    const S a = static_cast<S>(alpha_d), b = static_cast<S>(beta_d);
    double rr = 0.0, ru = 0.0, wu = 0.0, us = 0.0;
    while (c0 < cend)
    {
#pragma unroll
        for (int k = 0; k < F; ++k)
        {
            if (!ok[k]) continue;
            const CgFusedRow<S, ND, PRE>& o = row[k];
            const int64_t i = rw[k];
            const S acc = o.product(a, b, dia);
            const S pn = fma_s(b, o.p, o.u);
            const S sn = fma_s(b, o.s, o.w);
            const S rn = fma_s(-a, sn, o.r);
            const S un = PRE == kPreJacobi ? o.iv * rn : rn;
            v.p[i] = pn;
            v.x[i] = fma_s(a, pn, o.x);
            v.s_out[i] = sn;
            v.r_out[i] = rn;
            if (PRE == kPreJacobi) v.u_out[i] = un;
            v.w_out[i] = acc;
            const double rv = static_cast<double>(rn);
            const double uv = static_cast<double>(un);
            rr = __builtin_fma(rv, rv, rr);
            if (PRE == kPreJacobi) ru = __builtin_fma(rv, uv, ru);
            wu = __builtin_fma(static_cast<double>(acc), uv, wu);
            us = __builtin_fma(uv, static_cast<double>(sn), us);
        }
        c0 += F * kDirWaves;
        if (c0 < cend) load_pass(c0);
    }
    double part[4] = {rr, PRE == kPreJacobi ? ru : rr, wu, us};
    publish_partials<4>(part, part_out);
}

// ---- single-reduction iteration with a factored preconditioner ---------------------
//
// PCG with M^-1 = G^T G (G lower, both factors and A on the diagonal layout)
// as the single-reduction recurrence of k_cg_step / k_cg_spmv with u = G^T G r,
// three kernels per iteration:
//   k_cgf_step_lower  decision; p, s, x, r of the own rows; t = G r_new, each
//                     entry's r_new formed from that row's old r, w, s with the
//                     fused multiply-adds of its owner
//   k_cgf_upper       u = G^T t
//   k_cg_spmv_dia     w = A u and the partials r.r, r.u, w.u, u.s
// r, w, s and u alternate between two buffers with the iteration parity as in
// k_cg_fused_dia.

// The operands one lane of k_cgf_step_lower loads for one row: G's row and,
// per diagonal, the old r, w, s of the column's row, plus the row's own
// vectors (none of the addresses depend on alpha or beta).
template <typename S, int ND>
struct CgfRow
{
    static constexpr int kSlots = ND > 0 ? ND : kMaxDia;
    S val[kSlots], nr[kSlots], nw[kSlots], ns[kSlots];
    S r, s, w, p, x, u;
    unsigned m;

    __device__ __forceinline__ void load(int64_t i, int64_t rows, const DiaArgs& g, const S* __restrict__ gv,
                                         const uint16_t* __restrict__ gmask, const CgVectors<S>& v)
    {
        const int nd = ND > 0 ? ND : g.ndiag;
        m = gmask[i];
#pragma unroll
        for (int d = 0; d < kSlots; ++d)
        {
            if (d < nd)
            {
                int64_t c = i + g.off[d];
                c = c < 0 ? 0 : (c >= rows ? rows - 1 : c);
                val[d] = dia_value(i, rows, g, gv, d);
                nr[d] = v.r_in[c];
                nw[d] = v.w_in[c];
                ns[d] = v.s_in[c];
            }
        }
        r = v.r_in[i];
        s = v.s_in[i];
        w = v.w_in[i];
        p = v.p[i];
        x = v.x[i];
        u = v.u_in[i];
    }

    // The row of t = G r_new, its present entries in increasing column order.
    __device__ __forceinline__ S product(S a, S b, const DiaArgs& g) const
    {
        const int nd = ND > 0 ? ND : g.ndiag;
        S acc = S(0);
#pragma unroll
        for (int d = 0; d < kSlots; ++d)
            if (d < nd && ((m >> d) & 1u)) acc = fma_s(val[d], fma_s(-a, fma_s(b, ns[d], nw[d]), nr[d]), acc);
        return acc;
    }
};

template <typename S, int ND>
__global__ __launch_bounds__(kDirBlock) void k_cgf_step_lower(int64_t rows, DiaArgs g, const S* __restrict__ gv,
                                                              const uint16_t* __restrict__ gmask, CgVectors<S> v,
                                                              S* __restrict__ t, SolverState* st,
                                                              const double* part_in, int par, int count)
{
    if (st->done != kRunning) return;
    constexpr int F = ND > 0 ? kDiaInFlight : 1;
    const int lane = threadIdx.x % kWave;
    const int64_t chunks = chunk_count(rows);
    const int64_t per = ceil_div(chunks, static_cast<int64_t>(gridDim.x));
    const int64_t cbegin = xcd_logical_block(blockIdx.x, gridDim.x) * per;
    const int64_t cend = cbegin + per < chunks ? cbegin + per : chunks;
    CgfRow<S, ND> row[F];
    int64_t rw[F];
    bool ok[F];
    auto load_pass = [&](int64_t c0) {
#pragma unroll
        for (int k = 0; k < F; ++k)
        {
            const int64_t c = c0 + k * kDirWaves;
            const int64_t i = c * kChunkRows + lane;
            ok[k] = c < cend && i < rows;
            rw[k] = ok[k] ? i : rows - 1;
            row[k].load(rw[k], rows, g, gv, gmask, v);
        }
    };
    int64_t c0 = cbegin + threadIdx.x / kWave;
    load_pass(c0);
    double alpha_d = 0.0, beta_d = 0.0;
    if (cg_decision<S, kPreJacobi>(st, part_in, par, count, alpha_d, beta_d) != kRunning) return;
    const S a = static_cast<S>(alpha_d), b = static_cast<S>(beta_d);
    while (c0 < cend)
    {
#pragma unroll
        for (int k = 0; k < F; ++k)
        {
            if (!ok[k]) continue;
            const CgfRow<S, ND>& o = row[k];
            const int64_t i = rw[k];
            const S tr = o.product(a, b, g);
            const S pn = fma_s(b, o.p, o.u);
            const S sn = fma_s(b, o.s, o.w);
            v.p[i] = pn;
            v.x[i] = fma_s(a, pn, o.x);
            v.s_out[i] = sn;
            v.r_out[i] = fma_s(-a, sn, o.r);
            t[i] = tr;
        }
        c0 += F * kDirWaves;
        if (c0 < cend) load_pass(c0);
    }
}

// u = G^T t (the transpose factor on its diagonal layout).
template <typename S, int ND>
__global__ __launch_bounds__(kDirBlock) void k_cgf_upper(int64_t rows, DiaArgs gt, const S* __restrict__ gtv,
                                                         const uint16_t* __restrict__ gtmask,
                                                         const S* __restrict__ t, S* __restrict__ u,
                                                         const SolverState* st)
{
    if (st->done != kRunning) return;
    const int lane = threadIdx.x % kWave;
    const int64_t chunks = chunk_count(rows);
    const int64_t per = ceil_div(chunks, static_cast<int64_t>(gridDim.x));
    const int64_t cbegin = xcd_logical_block(blockIdx.x, gridDim.x) * per;
    const int64_t cend = cbegin + per < chunks ? cbegin + per : chunks;
    for (int64_t c0 = cbegin + threadIdx.x / kWave; c0 < cend; c0 += kDiaInFlight * kDirWaves)
    {
        int64_t rw[kDiaInFlight];
        bool ok[kDiaInFlight];
        S acc[kDiaInFlight];
#pragma unroll
        for (int k = 0; k < kDiaInFlight; ++k)
        {
            const int64_t c = c0 + k * kDirWaves;
            const int64_t i = c * kChunkRows + lane;
            ok[k] = c < cend && i < rows;
            rw[k] = ok[k] ? i : rows - 1;
            acc[k] = dia_row<S, ND>(rw[k], rows, gt, gtv, gtmask[rw[k]], t);
        }
#pragma unroll
        for (int k = 0; k < kDiaInFlight; ++k)
            if (ok[k]) u[rw[k]] = acc[k];
    }
}

// k_cgf_step_lower and k_cgf_upper as one kernel for a four-diagonal G:
// u_i = sum_d G[k_d, i] t_{k_d} over the rows k_d = i - off_d that hold an
// entry in column i, each t_k = G r_new of row k recomputed from G's row k and
// the old r, w, s of its columns.  t never reaches memory and G^T is not read;
// the sums run in the order of k_factor_lower_dia on G (t) and of
// k_cgf_upper on G^T (u, increasing column k), so the bits are those of the
// two-kernel pair.
template <typename S>
struct CgfPairRow
{
    static constexpr int D = 4;
    S gv[D][D];                      // G[k_d, k_d + off_e]
    S nr[D][D], nw[D][D], ns[D][D];  // old r, w, s of row k_d + off_e (e != d)
    unsigned km[D];                  // presence bits of row k_d (none past the last row)
    S r, s, w, p, x, u;

    __device__ __forceinline__ void load(int64_t i, int64_t rows, const DiaArgs& g, const S* __restrict__ gvals,
                                         const uint16_t* __restrict__ gmask, const CgVectors<S>& v)
    {
#pragma unroll
        for (int d = 0; d < D; ++d)
        {
            const int64_t k = i - g.off[d];
            const bool in = k >= 0 && k < rows;
            const int64_t kc = in ? k : (k < 0 ? 0 : rows - 1);
            km[d] = in ? gmask[kc] : 0u;
#pragma unroll
            for (int e = 0; e < D; ++e)
            {
                gv[d][e] = dia_value(kc, rows, g, gvals, e);
                if (e == d) continue;  // column i: the row's own r, w, s
                int64_t c = kc + g.off[e];
                c = c < 0 ? 0 : (c >= rows ? rows - 1 : c);
                nr[d][e] = v.r_in[c];
                nw[d][e] = v.w_in[c];
                ns[d][e] = v.s_in[c];
            }
        }
        r = v.r_in[i];
        s = v.s_in[i];
        w = v.w_in[i];
        p = v.p[i];
        x = v.x[i];
        u = v.u_in[i];
    }

    __device__ __forceinline__ S product(S a, S b) const
    {
        const S rn_own = fma_s(-a, fma_s(b, s, w), r);
        S acc = S(0);
#pragma unroll
        for (int d = D - 1; d >= 0; --d)
        {
            if (!((km[d] >> d) & 1u)) continue;  // no G[k_d, i]
            S t = S(0);
#pragma unroll
            for (int e = 0; e < D; ++e)
            {
                if (!((km[d] >> e) & 1u)) continue;
                const S rn = e == d ? rn_own : fma_s(-a, fma_s(b, ns[d][e], nw[d][e]), nr[d][e]);
                t = fma_s(gv[d][e], rn, t);
            }
            acc = fma_s(gv[d][d], t, acc);
        }
        return acc;
    }
};

// One row per lane and pass; the first pass's loads are issued before the
// decision.  (Two rows in flight need 170 VGPRs and ran slower: 63.4 against
// 55.9 us per 128^3 learned iteration; so did a variant that shares the
// neighbouring rows' r and t across lanes by wavefront shuffles, 62.6 us.)
template <typename S>
__global__ __launch_bounds__(kDirBlock) void k_cgf_step_pair(int64_t rows, DiaArgs g, const S* __restrict__ gv,
                                                             const uint16_t* __restrict__ gmask, CgVectors<S> v,
                                                             SolverState* st, const double* part_in, int par,
                                                             int count)
{
    if (st->done != kRunning) return;
    const int lane = threadIdx.x % kWave;
    const int64_t chunks = chunk_count(rows);
    const int64_t per = ceil_div(chunks, static_cast<int64_t>(gridDim.x));
    const int64_t cbegin = xcd_logical_block(blockIdx.x, gridDim.x) * per;
    const int64_t cend = cbegin + per < chunks ? cbegin + per : chunks;
    CgfPairRow<S> row;
    int64_t c0 = cbegin + threadIdx.x / kWave;
    // rows past the end load row rows - 1 and are dropped
    auto load_row = [&](int64_t c, int64_t& i) {
        i = c * kChunkRows + lane;
        const bool ok = c < cend && i < rows;
        row.load(ok ? i : rows - 1, rows, g, gv, gmask, v);
        return ok;
    };
    int64_t i;
    bool ok = load_row(c0, i);
    double alpha_d = 0.0, beta_d = 0.0;
    if (cg_decision<S, kPreJacobi>(st, part_in, par, count, alpha_d, beta_d) != kRunning) return;
    const S a = static_cast<S>(alpha_d), b = static_cast<S>(beta_d);
    while (c0 < cend)
    {
        if (ok)
        {
            const S un = row.product(a, b);
            const S pn = fma_s(b, row.p, row.u);
            const S sn = fma_s(b, row.s, row.w);
            v.p[i] = pn;
            v.x[i] = fma_s(a, pn, row.x);
            v.s_out[i] = sn;
            v.r_out[i] = fma_s(-a, sn, row.r);
            v.u_out[i] = un;
        }
        c0 += kDirWaves;
        if (c0 < cend) ok = load_row(c0, i);
    }
}

// The solver state into the host's fine-grained status slot, 8 bytes per lane.
__global__ void k_publish_state(const SolverState* __restrict__ st, SolverState* host)
{
    static_assert(sizeof(SolverState) % 8 == 0 && sizeof(SolverState) / 8 <= kWave, "state record size");
    const auto* src = reinterpret_cast<const unsigned long long*>(st);
    auto* dst = reinterpret_cast<unsigned long long*>(host);
    if (threadIdx.x < sizeof(SolverState) / 8) dst[threadIdx.x] = src[threadIdx.x];
}

// Squared residuals of a fixed-iteration solve (toDoubleKernel /
// toDoubleCheckedKernel, iterative_solver_kernels.cuh:240-252).
__global__ __launch_bounds__(kBlock) void k_fixed_out(const SolverState* st, const double* partials, int count_res,
                                                      int zero_iterations, double* out)
{
    double res[2];
    reduce_partials<2>(partials + kPartialsRes, count_res, res);
    if (threadIdx.x == 0)
    {
        const int d = st->done;
        const bool bad = d == kBreakdown || d == kNonFinite || d == kInitNonFinite || d == kBadDiagonal;
        out[0] = zero_iterations ? res[0] : st->rr0;
        out[1] = bad ? NAN : res[0];
    }
}

// ---- host side --------------------------------------------------------------

struct Problem
{
    const pm_csr* a;
    const void* rhs;
    void* x;
    const void* inv_blocks;  // block preconditioner (device), or null
    const pm_csr* gf;        // factored preconditioner: G (lower) and G^T, or null
    const pm_csr* gtf;
    int64_t bsize;
    int pre;                 // Pre
    int grid_rows;           // workgroups of the row kernels (publish p.q)
    int grid_units;          // workgroups of the unit kernels (publish r.r, r.z)
    bool nt;                 // non-temporal matrix loads in the iteration
    bool dma;                // 4-byte matrix operands staged into LDS by global_load_lds
    DiaLayout dia;           // diagonal layout of A for q = A p, when it pays
    DiaLayout dia_g, dia_gt; // ... of the factors G and G^T (both or neither)
    bool fuse_direction;     // on the diagonal layout: p_new formed inside the SpMV
    bool fuse_step;          // factored: the step folded into G r and G^T t
    int fuse_step_env;       // PLAMATRIX_AMD_FUSE_STEP: 0 off, 1 on, -1 unset
    bool single_reduction;   // CG / Jacobi-PCG as the two-kernel Chronopoulos-Gear iteration
    bool cg_fused;           // ... as one kernel per iteration on the diagonal layout (factored: G r and G^T t in one)
    bool cg_factored;        // factored PCG as the three-kernel single-reduction iteration (diagonal layouts)
    bool long_rows;          // classic CSR SpMV with kLongLanes lanes per row (few long rows)
    // workspace vectors
    void *r, *z, *q, *p0, *p1, *inv;
    void *r2, *z2, *q2, *s2; // second buffers of r, u, w, s for the one-kernel iteration (or null)
    SolverState* st;
    double* partials;
};

template <typename S, typename I>
struct Calls
{
    static const I* ro(const pm_csr* a)
    {
        if constexpr (sizeof(I) == 4) return a->row_offsets32;
        else return a->row_offsets;
    }
    static const I* ci(const pm_csr* a)
    {
        if constexpr (sizeof(I) == 4) return a->col_indices32;
        else return a->col_indices;
    }

    static int full_grid(int64_t rows) { return static_cast<int>(ceil_div(chunk_count(rows), kWavesPerBlock)); }

    template <int PRE, int MODE>
    static KernelCall update_p(const Problem& pb, const void* p, int par)
    {
        if constexpr (PRE == kPreBlock)
            return make_call(k_update_blocks<S, MODE>, dim3(pb.grid_units), dim3(kBlock), pb.a->rows / pb.bsize,
                             pb.bsize, static_cast<S*>(pb.x), static_cast<S*>(pb.r), static_cast<S*>(pb.z),
                             static_cast<const S*>(p), static_cast<const S*>(pb.q),
                             static_cast<const S*>(pb.inv_blocks), pb.st, pb.partials, par, pb.grid_rows);
        const int64_t units = PRE == kPreBlock ? pb.a->rows / pb.bsize : pb.a->rows;
        S* zz = PRE == kPreNone ? static_cast<S*>(pb.r) : static_cast<S*>(pb.z);
        const S* inv = PRE == kPreBlock ? static_cast<const S*>(pb.inv_blocks) : static_cast<const S*>(pb.inv);
        return make_call(k_update<S, PRE, MODE>, dim3(pb.grid_units), dim3(kBlock), units, pb.bsize,
                         static_cast<S*>(pb.x), static_cast<S*>(pb.r), zz, static_cast<const S*>(p),
                         static_cast<const S*>(pb.q), inv, pb.st, pb.partials, par, pb.grid_rows);
    }

    template <int MODE>
    static KernelCall update_any(const Problem& pb, const void* p, int par)
    {
        switch (pb.pre)
        {
        case kPreNone: return update_p<kPreNone, MODE>(pb, p, par);
        case kPreJacobi: return update_p<kPreJacobi, MODE>(pb, p, par);
        default: return update_p<kPreBlock, MODE>(pb, p, par);
        }
    }

    static void init(const Problem& pb, std::vector<KernelCall>& out)
    {
        const int grid = full_grid(pb.a->rows);
        if (pb.pre == kPreJacobi)
            out.push_back(make_call(k_residual<S, I, true, true>, dim3(grid), dim3(kBlock), pb.a->rows, ro(pb.a),
                                    ci(pb.a), static_cast<const S*>(pb.a->values), static_cast<const S*>(pb.rhs),
                                    static_cast<const S*>(pb.x), static_cast<S*>(pb.r), static_cast<S*>(pb.inv),
                                    static_cast<S*>(pb.p0), pb.st));
        else if (pb.long_rows)
            out.push_back(make_call(k_residual_lanes<S, I, true>, dim3(pb.grid_rows), dim3(kBlock), pb.a->rows,
                                    ro(pb.a), ci(pb.a), static_cast<const S*>(pb.a->values),
                                    static_cast<const S*>(pb.rhs), static_cast<const S*>(pb.x),
                                    static_cast<S*>(pb.r), static_cast<S*>(pb.p0), pb.st));
        else
            out.push_back(make_call(k_residual<S, I, false, true>, dim3(grid), dim3(kBlock), pb.a->rows, ro(pb.a),
                                    ci(pb.a), static_cast<const S*>(pb.a->values), static_cast<const S*>(pb.rhs),
                                    static_cast<const S*>(pb.x), static_cast<S*>(pb.r), static_cast<S*>(nullptr),
                                    static_cast<S*>(pb.p0), pb.st));
        if (pb.single_reduction)
        {
            const auto k = pb.pre == kPreJacobi ? k_cg_init<S, kPreJacobi> : k_cg_init<S, kPreNone>;
            out.push_back(make_call(k, dim3(pb.grid_units), dim3(kBlock), pb.a->rows, static_cast<const S*>(pb.r),
                                    static_cast<const S*>(pb.inv), static_cast<S*>(pb.z), static_cast<S*>(pb.p1),
                                    static_cast<const SolverState*>(pb.st)));
            // factored: u = G^T (G r)
            if (pb.cg_factored) cg_factor_apply(pb, static_cast<const S*>(pb.r), static_cast<S*>(pb.z), out);
            // the initialisation acts as iteration -1: w = A u into the first
            // buffers, partials into region 1
            cg_spmv(pb, cg_vectors(pb, 1), 1, out);
            return;
        }
        if (pb.pre == kPreFactored) factored_apply(pb, out);
        else out.push_back(update_any<kModeInit>(pb, pb.p0, 0));
    }

    // Vectors of single-reduction iteration t (parity par): u = z (Jacobi) or
    // r, w = q, p = p0, s = p1.  The one-kernel iteration reads them from the
    // first buffers and writes the second ones at even t, the other way round
    // at odd t; every chunk of iterations ends on an odd t, so between chunks
    // the vectors sit in the first buffers, where the three-kernel form finds
    // them.  The two-kernel iteration works in place.
    static CgVectors<S> cg_vectors(const Problem& pb, int par)
    {
        const bool jac = pb.pre != kPreNone;
        const bool alt = pb.cg_fused || pb.cg_factored;
        auto* r1 = static_cast<S*>(pb.r);
        auto* u1 = jac ? static_cast<S*>(pb.z) : r1;
        auto* w1 = static_cast<S*>(pb.q);
        auto* s1 = static_cast<S*>(pb.p1);
        auto* r2 = alt ? static_cast<S*>(pb.r2) : r1;
        auto* u2 = alt ? (jac ? static_cast<S*>(pb.z2) : r2) : u1;
        auto* w2 = alt ? static_cast<S*>(pb.q2) : w1;
        auto* s2 = alt ? static_cast<S*>(pb.s2) : s1;
        CgVectors<S> v{};
        v.x = static_cast<S*>(pb.x);
        v.p = static_cast<S*>(pb.p0);
        v.inv = static_cast<const S*>(pb.inv);
        const bool even = par == 0;
        v.r_in = even ? r1 : r2;
        v.u_in = even ? u1 : u2;
        v.w_in = even ? w1 : w2;
        v.s_in = even ? s1 : s2;
        v.r_out = even ? r2 : r1;
        v.u_out = even ? u2 : u1;
        v.w_out = even ? w2 : w1;
        v.s_out = even ? s2 : s1;
        return v;
    }

    // w = A u of the new u and the partials r.r, (r.u,) w.u, u.s (grid_units
    // workgroups, as many as the three-kernel update publishes, so the solve
    // can switch forms between chunks) into region par.
    static void cg_spmv(const Problem& pb, const CgVectors<S>& v, int par, std::vector<KernelCall>& out)
    {
        // a preconditioned u (Jacobi or factored) is published next to r
        const bool jac = pb.pre != kPreNone;
        const int nd = pb.dia.args.ndiag;
        double* part_out = cg_region(pb.partials, par);
        if (pb.dia.active)
        {
            const auto k = jac ? (nd == 7 ? k_cg_spmv_dia<S, 7, kPreJacobi> : nd == 5 ? k_cg_spmv_dia<S, 5, kPreJacobi>
                                                                                  : k_cg_spmv_dia<S, 0, kPreJacobi>)
                               : (nd == 7 ? k_cg_spmv_dia<S, 7, kPreNone> : nd == 5 ? k_cg_spmv_dia<S, 5, kPreNone>
                                                                                : k_cg_spmv_dia<S, 0, kPreNone>);
            out.push_back(make_call(k, dim3(pb.grid_units), dim3(kDirBlock), pb.a->rows, pb.dia.args,
                                    static_cast<const S*>(pb.dia.values), pb.dia.mask, v,
                                    static_cast<const SolverState*>(pb.st), part_out));
            return;
        }
        const auto k = jac ? k_cg_spmv<S, I, kPreJacobi> : k_cg_spmv<S, I, kPreNone>;
        out.push_back(make_call(k, dim3(pb.grid_units), dim3(kDirBlock), pb.a->rows, ro(pb.a), ci(pb.a),
                                static_cast<const S*>(pb.a->values), v, static_cast<const SolverState*>(pb.st),
                                part_out));
    }

    // u = G^T (G r) on the factors' diagonal layouts, t = G r in pb.inv.
    static void cg_factor_apply(const Problem& pb, const S* r, S* u, std::vector<KernelCall>& out)
    {
        const int64_t rows = pb.a->rows;
        auto* t = static_cast<S*>(pb.inv);
        const int grid = static_cast<int>(std::min<int64_t>(full_grid(rows), kSpmvGrid));
        out.push_back(make_call(pb.dia_g.args.ndiag == 4 ? k_factor_lower_dia<S, 4> : k_factor_lower_dia<S, 0>,
                                dim3(grid), dim3(kBlock), rows, pb.dia_g.args, static_cast<const S*>(pb.dia_g.values),
                                pb.dia_g.mask, r, t, static_cast<const SolverState*>(pb.st)));
        cg_factor_upper(pb, u, out);
    }

    static void cg_factor_upper(const Problem& pb, S* u, std::vector<KernelCall>& out)
    {
        out.push_back(make_call(pb.dia_gt.args.ndiag == 4 ? k_cgf_upper<S, 4> : k_cgf_upper<S, 0>,
                                dim3(pb.grid_units), dim3(kDirBlock), pb.a->rows, pb.dia_gt.args,
                                static_cast<const S*>(pb.dia_gt.values), pb.dia_gt.mask,
                                static_cast<const S*>(pb.inv), u, static_cast<const SolverState*>(pb.st)));
    }

    static void cg_iteration(const Problem& pb, int64_t t, bool refresh, std::vector<KernelCall>& out)
    {
        const int par = static_cast<int>(t % 2);
        const bool jac = pb.pre == kPreJacobi;
        const CgVectors<S> v = cg_vectors(pb, par);
        const double* part_in = cg_region(pb.partials, 1 - par);
        double* part_out = cg_region(pb.partials, par);
        if (pb.cg_factored)
        {
            if (refresh)
            {
                // p, s, x of the step; r = b - A x; u = G^T (G r)
                out.push_back(make_call(k_cg_step<S, kPreJacobi, false>, dim3(pb.grid_units), dim3(kBlock),
                                        pb.a->rows, v, pb.st, part_in, par, pb.grid_units));
                out.push_back(make_call(k_residual<S, I, false, false>, dim3(full_grid(pb.a->rows)), dim3(kBlock),
                                        pb.a->rows, ro(pb.a), ci(pb.a), static_cast<const S*>(pb.a->values),
                                        static_cast<const S*>(pb.rhs), static_cast<const S*>(pb.x), v.r_out,
                                        static_cast<S*>(nullptr), static_cast<S*>(nullptr), pb.st));
                cg_factor_apply(pb, v.r_out, v.u_out, out);
            }
            else if (pb.cg_fused)
            {
                // G r and G^T t in one kernel
                out.push_back(make_call(k_cgf_step_pair<S>, dim3(pb.grid_units), dim3(kDirBlock), pb.a->rows,
                                        pb.dia_g.args, static_cast<const S*>(pb.dia_g.values), pb.dia_g.mask, v,
                                        pb.st, part_in, par, pb.grid_units));
            }
            else
            {
                const auto k = pb.dia_g.args.ndiag == 4 ? k_cgf_step_lower<S, 4> : k_cgf_step_lower<S, 0>;
                out.push_back(make_call(k, dim3(pb.grid_units), dim3(kDirBlock), pb.a->rows, pb.dia_g.args,
                                        static_cast<const S*>(pb.dia_g.values), pb.dia_g.mask, v,
                                        static_cast<S*>(pb.inv), pb.st, part_in, par, pb.grid_units));
                cg_factor_upper(pb, v.u_out, out);
            }
            cg_spmv(pb, v, par, out);
            return;
        }
        if (pb.cg_fused && !refresh)
        {
            const int nd = pb.dia.args.ndiag;
            const auto k = jac ? (nd == 7   ? k_cg_fused_dia<S, 7, kPreJacobi>
                                  : nd == 5 ? k_cg_fused_dia<S, 5, kPreJacobi>
                                            : k_cg_fused_dia<S, 0, kPreJacobi>)
                               : (nd == 7   ? k_cg_fused_dia<S, 7, kPreNone>
                                  : nd == 5 ? k_cg_fused_dia<S, 5, kPreNone>
                                            : k_cg_fused_dia<S, 0, kPreNone>);
            out.push_back(make_call(k, dim3(pb.grid_units), dim3(kDirBlock), pb.a->rows, pb.dia.args,
                                    static_cast<const S*>(pb.dia.values), pb.dia.mask, v, pb.st, part_in, part_out,
                                    par, pb.grid_units));
            return;
        }
        const auto step = refresh ? (jac ? k_cg_step<S, kPreJacobi, false> : k_cg_step<S, kPreNone, false>)
                                  : (jac ? k_cg_step<S, kPreJacobi, true> : k_cg_step<S, kPreNone, true>);
        out.push_back(make_call(step, dim3(pb.grid_units), dim3(kBlock), pb.a->rows, v, pb.st, part_in, par,
                                pb.grid_units));
        if (refresh)
        {
            out.push_back(make_call(k_residual<S, I, false, false>, dim3(full_grid(pb.a->rows)), dim3(kBlock),
                                    pb.a->rows, ro(pb.a), ci(pb.a), static_cast<const S*>(pb.a->values),
                                    static_cast<const S*>(pb.rhs), static_cast<const S*>(pb.x), v.r_out,
                                    static_cast<S*>(nullptr), static_cast<S*>(nullptr), pb.st));
            if (jac)
                out.push_back(make_call(k_cg_precondition<S>, dim3(pb.grid_units), dim3(kBlock), pb.a->rows, v,
                                        static_cast<const SolverState*>(pb.st)));
        }
        cg_spmv(pb, v, par, out);
    }

    // z = G^T (G r) and the r.r, r.z partials (grid_rows workgroups).  With
    // p_step (the new direction of iteration parity par) the step comes first:
    // alpha, x += alpha p, r -= alpha q -- folded into the two products unless
    // fuse_step is off.
    static void factored_apply(const Problem& pb, std::vector<KernelCall>& out, const void* p_step = nullptr,
                               int par = 0)
    {
        const int64_t rows = pb.a->rows;
        const int grid = static_cast<int>(std::min<int64_t>(full_grid(rows), kSpmvGrid));
        const bool fused = p_step != nullptr && pb.fuse_step;
        if (p_step != nullptr && !fused)
            out.push_back(make_call(k_step_xr<S>, dim3(pb.grid_units), dim3(kBlock), rows, static_cast<S*>(pb.x),
                                    static_cast<S*>(pb.r), static_cast<const S*>(p_step),
                                    static_cast<const S*>(pb.q), pb.st, pb.partials, par, pb.grid_rows));
        const auto* r_in = static_cast<const S*>(pb.r);
        auto* t = static_cast<S*>(pb.inv);
        if (pb.dia_g.active)
        {
            const bool four = pb.dia_g.args.ndiag == 4;
            if (fused)
                out.push_back(make_call(four ? k_step_lower_dia<S, 4> : k_step_lower_dia<S, 0>, dim3(grid),
                                        dim3(kBlock), rows, pb.dia_g.args, static_cast<const S*>(pb.dia_g.values),
                                        pb.dia_g.mask, static_cast<S*>(pb.x), r_in, static_cast<const S*>(p_step),
                                        static_cast<const S*>(pb.q), t, pb.st, pb.partials, par, pb.grid_rows));
            else
                out.push_back(make_call(four ? k_factor_lower_dia<S, 4> : k_factor_lower_dia<S, 0>, dim3(grid),
                                        dim3(kBlock), rows, pb.dia_g.args, static_cast<const S*>(pb.dia_g.values),
                                        pb.dia_g.mask, r_in, t, static_cast<const SolverState*>(pb.st)));
            const bool four_t = pb.dia_gt.args.ndiag == 4;
            const auto upper = fused ? (four_t ? k_factor_upper_dia<S, 4, true> : k_factor_upper_dia<S, 0, true>)
                                     : (four_t ? k_factor_upper_dia<S, 4, false> : k_factor_upper_dia<S, 0, false>);
            out.push_back(make_call(upper, dim3(pb.grid_rows), dim3(kBlock), rows, pb.dia_gt.args,
                                    static_cast<const S*>(pb.dia_gt.values), pb.dia_gt.mask,
                                    static_cast<const S*>(t), static_cast<S*>(pb.r), static_cast<const S*>(pb.q),
                                    static_cast<S*>(pb.z), static_cast<const SolverState*>(pb.st), pb.partials));
            return;
        }
        if (fused)
            out.push_back(make_call(pb.dma ? k_step_lower<S, int32_t, true> : k_step_lower<S, int32_t, false>,
                                    dim3(grid), dim3(kBlock), rows, pb.gf->row_offsets32, pb.gf->col_indices32,
                                    static_cast<const S*>(pb.gf->values), static_cast<S*>(pb.x), r_in,
                                    static_cast<const S*>(p_step), static_cast<const S*>(pb.q), t, pb.st,
                                    pb.partials, par, pb.grid_rows));
        else
            out.push_back(make_call(pb.dma ? k_factor_lower<S, int32_t, true> : k_factor_lower<S, int32_t, false>,
                                    dim3(grid), dim3(kBlock), rows, pb.gf->row_offsets32, pb.gf->col_indices32,
                                    static_cast<const S*>(pb.gf->values), r_in, t,
                                    static_cast<const SolverState*>(pb.st)));
        const auto upper = fused ? (pb.dma ? k_factor_upper<S, int32_t, true, true>
                                           : k_factor_upper<S, int32_t, false, true>)
                                 : (pb.dma ? k_factor_upper<S, int32_t, true, false>
                                           : k_factor_upper<S, int32_t, false, false>);
        out.push_back(make_call(upper, dim3(pb.grid_rows), dim3(kBlock), rows, pb.gtf->row_offsets32,
                                pb.gtf->col_indices32, static_cast<const S*>(pb.gtf->values),
                                static_cast<const S*>(t), static_cast<S*>(pb.r), static_cast<const S*>(pb.q),
                                static_cast<S*>(pb.z), static_cast<const SolverState*>(pb.st), pb.partials));
    }

    // Iteration t: t % 2 picks the p buffers and the state slots.
    static void iteration(const Problem& pb, int64_t t, bool refresh, std::vector<KernelCall>& out)
    {
        if (pb.single_reduction)
        {
            cg_iteration(pb, t, refresh, out);
            return;
        }
        const int par = static_cast<int>(t % 2);
        const void* p_old = par == 0 ? pb.p0 : pb.p1;
        void* p_new = par == 0 ? pb.p1 : pb.p0;
        const S* zsrc = pb.pre == kPreNone ? static_cast<const S*>(pb.r) : static_cast<const S*>(pb.z);
        const int count_res = pb.pre == kPreFactored ? pb.grid_rows : pb.grid_units;
        const int nd = pb.dia.args.ndiag;
        if (pb.dia.active && pb.fuse_direction)
        {
            const auto fused = nd == 7 ? k_dir_spmv_dia<S, 7> : nd == 5 ? k_dir_spmv_dia<S, 5> : k_dir_spmv_dia<S, 0>;
            out.push_back(make_call(fused, dim3(pb.grid_rows), dim3(kDirBlock), pb.a->rows, pb.dia.args,
                                    static_cast<const S*>(pb.dia.values), pb.dia.mask, zsrc,
                                    static_cast<const S*>(p_old), static_cast<S*>(p_new), static_cast<S*>(pb.q),
                                    pb.st, pb.partials, par, count_res));
        }
        else
        {
            out.push_back(make_call(k_direction<S>, dim3(pb.grid_units), dim3(kBlock), pb.a->rows, zsrc,
                                    static_cast<const S*>(p_old), static_cast<S*>(p_new), pb.st, pb.partials, par,
                                    count_res));
        }
        if (pb.dia.active && pb.fuse_direction)
        {
            // q and p.q came with the direction
        }
        else if (pb.dia.active)
        {
            const auto dia_den = nd == 7 ? k_spmv_den_dia<S, 7> : nd == 5 ? k_spmv_den_dia<S, 5> : k_spmv_den_dia<S, 0>;
            out.push_back(make_call(dia_den, dim3(pb.grid_rows), dim3(kDirBlock), pb.a->rows, pb.dia.args,
                                    static_cast<const S*>(pb.dia.values), pb.dia.mask, static_cast<const S*>(p_new),
                                    static_cast<S*>(pb.q), static_cast<const SolverState*>(pb.st), pb.partials));
        }
        else if (pb.long_rows)
        {
            out.push_back(make_call(k_spmv_den_lanes<S, I>, dim3(pb.grid_rows), dim3(kDirBlock), pb.a->rows,
                                    ro(pb.a), ci(pb.a), static_cast<const S*>(pb.a->values),
                                    static_cast<const S*>(p_new), static_cast<S*>(pb.q),
                                    static_cast<const SolverState*>(pb.st), pb.partials));
        }
        else
        {
            const auto spmv_den = pb.nt ? (pb.dma ? k_spmv_den<S, I, true, true> : k_spmv_den<S, I, true, false>)
                                        : (pb.dma ? k_spmv_den<S, I, false, true> : k_spmv_den<S, I, false, false>);
            out.push_back(make_call(spmv_den, dim3(pb.grid_rows), dim3(kDirBlock), pb.a->rows, ro(pb.a), ci(pb.a),
                                    static_cast<const S*>(pb.a->values), static_cast<const S*>(p_new),
                                    static_cast<S*>(pb.q), static_cast<const SolverState*>(pb.st), pb.partials));
        }
        if (!refresh)
        {
            if (pb.pre == kPreFactored)
            {
                factored_apply(pb, out, p_new, par);
                return;
            }
            out.push_back(update_any<kModeStep>(pb, p_new, par));
            return;
        }
        out.push_back(make_call(k_axpy_x<S>, dim3(pb.grid_units), dim3(kBlock), pb.a->rows, static_cast<S*>(pb.x),
                                static_cast<const S*>(p_new), pb.st, pb.partials, par, pb.grid_rows));
        out.push_back(make_call(k_residual<S, I, false, false>, dim3(full_grid(pb.a->rows)), dim3(kBlock),
                                pb.a->rows, ro(pb.a), ci(pb.a), static_cast<const S*>(pb.a->values),
                                static_cast<const S*>(pb.rhs), static_cast<const S*>(pb.x), static_cast<S*>(pb.r),
                                static_cast<S*>(nullptr), static_cast<S*>(nullptr), pb.st));
        if (pb.pre == kPreFactored) factored_apply(pb, out);
        else out.push_back(update_any<kModeRefresh>(pb, p_new, par));
    }

    static void build(const Problem& pb, int what, int64_t first, int64_t count, int refresh_interval,
                      std::vector<KernelCall>& out)
    {
        if (what == 0)
        {
            init(pb, out);
            return;
        }
        for (int64_t t = first; t < first + count; ++t)
        {
            const bool refresh = refresh_interval > 0 && ((t + 1) % refresh_interval) == 0;
            iteration(pb, t, refresh, out);
        }
    }
};

template <typename S>
void build_calls(const Problem& pb, int what, int64_t first, int64_t count, int refresh, std::vector<KernelCall>& out)
{
    if (use_compact(pb.a)) Calls<S, int32_t>::build(pb, what, first, count, refresh, out);
    else Calls<S, int64_t>::build(pb, what, first, count, refresh, out);
}

int validate_options(const pm_solver_options* o)
{
    if (!o) return fail(PM_INVALID_ARGUMENT, "null solver options");
    if (o->max_iterations < 0 || !std::isfinite(o->relative_tolerance) || o->relative_tolerance < 0.0 ||
        o->relative_tolerance > 1.0 || !std::isfinite(o->absolute_tolerance) || o->absolute_tolerance < 0.0)
        return fail(PM_INVALID_ARGUMENT, "invalid iterative solver options");
    return PM_OK;
}

int check_system(const pm_csr* a, const void* rhs, const void* x, const char* op)
{
    PM_TRY(check_csr(a, op));
    if (a->rows != a->cols) return fail(PM_INVALID_ARGUMENT, "%s requires square A and Nx1 rhs/solution", op);
    if (a->rows > 0 && (!rhs || !x)) return fail(PM_INVALID_ARGUMENT, "%s: null rhs or solution", op);
    if (rhs && rhs == x) return fail(PM_INVALID_ARGUMENT, "%s: rhs and solution must not alias", op);
    return PM_OK;
}

// The factored step (x += alpha p, r -= alpha q) rides in the two factor
// products when G is on the diagonal layout (128^3 diffusion: 63.1 vs 64.7 us per
// iteration) or the system is small enough to be launch-bound; on large CSR
// factors the products' second gathers of r and q cost more than the kernel
// saved (2.1M rows: 29.8 + 35.0 us vs 12.4 + 19.9 + 22.9 us).
// PLAMATRIX_AMD_FUSE_STEP=0 / 1 forces either form.
constexpr int64_t kFuseStepCsrRows = 65536;

bool fuse_step_wanted(const Problem& pb)
{
    if (pb.fuse_step_env >= 0) return pb.fuse_step_env == 1;
    return pb.dia_g.active || pb.a->rows < kFuseStepCsrRows;
}

size_t elem_size(int32_t dtype) { return dtype == PM_F64 ? 8 : 4; }

// CG and Jacobi-PCG run as the two-kernel single-reduction iteration in the
// adaptive solves, factored PCG as the three-kernel one when A, G and G^T are
// all on the diagonal layout (PLAMATRIX_AMD_SINGLE_REDUCTION=0 keeps the
// classic forms, which the fused-direction option also needs).
bool single_reduction_wanted(const Problem& pb)
{
    const char* env = std::getenv("PLAMATRIX_AMD_SINGLE_REDUCTION");
    if (env != nullptr && env[0] == '0') return false;
    if (pb.fuse_direction) return false;
    if (pb.pre == kPreFactored) return pb.dia.active && pb.dia_g.active && pb.dia_gt.active && pb.r2 != nullptr;
    return pb.pre == kPreNone || pb.pre == kPreJacobi;
}

bool dia_wanted(const pm_csr* a);

// Reserve workspace storage and fill the workspace part of a Problem.
int prepare(pm_workspace* ws, Problem& pb, hipStream_t stream)
{
    PM_TRY(ws->bind(stream));
    PM_TRY(ws->ensure_events());
    const int64_t n = pb.a->rows;
    const size_t vec = ((elem_size(pb.a->dtype) * static_cast<size_t>(n > 0 ? n : 1)) + 255) & ~size_t(255);
    // a matrix the diagonal layout may take also gets the second buffers of
    // the one-kernel single-reduction iteration
    const bool second = dia_wanted(pb.a);
    PM_TRY(ws->vectors.reserve(vec * (second ? 10 : 6)));
    if (n > ws->capacity_rows) ws->capacity_rows = n;
    PM_TRY(ws->partials.reserve(sizeof(double) * kPartialsDoubles));
    PM_TRY(ws->state.reserve(4096));
    char* base = ws->vectors.as<char>();
    pb.r = base;
    pb.z = base + vec;
    pb.q = base + 2 * vec;
    pb.p0 = base + 3 * vec;
    pb.p1 = base + 4 * vec;
    pb.inv = base + 5 * vec;
    pb.r2 = second ? base + 6 * vec : nullptr;
    pb.z2 = second ? base + 7 * vec : nullptr;
    pb.q2 = second ? base + 8 * vec : nullptr;
    pb.s2 = second ? base + 9 * vec : nullptr;
    pb.cg_fused = false;
    pb.cg_factored = false;
    pb.long_rows = false;
    pb.st = ws->state.as<SolverState>();
    pb.partials = ws->partials.as<double>();
    const int64_t units = pb.pre == kPreBlock ? n / pb.bsize : n;
    pb.grid_rows = static_cast<int>(std::min<int64_t>(ceil_div(chunk_count(n), kDirWaves), kDirGrid));
    pb.grid_units = solver_grid_units(units);
    // block-Jacobi: a wavefront per kWave / bsize blocks in the update
    if (pb.pre == kPreBlock)
        pb.grid_units = static_cast<int>(std::max<int64_t>(
            1, std::min<int64_t>(ceil_div(ceil_div(units, kWave / pb.bsize), kBlock / kWave), kSolverGrid)));
    const char* nt = std::getenv("PLAMATRIX_AMD_SOLVER_NT");
    pb.nt = nt != nullptr && nt[0] == '1';
    // LDS-DMA staging speeds up the HBM-streamed SpMV but slows the solver's
    // cache-resident iteration (measured: 53.3 vs 51.1 us per iteration on the 128^3 diffusion matrix)
    const char* dma = std::getenv("PLAMATRIX_AMD_SOLVER_DMA");
    pb.dma = dma != nullptr && dma[0] == '1';
    // fusing the direction into the diagonal-layout SpMV saves a pass over z
    // and p_old but measured slower (41.7 vs 40.6 us per 128^3 diffusion iteration): off
    // unless PLAMATRIX_AMD_FUSE_DIRECTION=1
    const char* fuse = std::getenv("PLAMATRIX_AMD_FUSE_DIRECTION");
    pb.fuse_direction = fuse != nullptr && fuse[0] == '1';
    const char* fuse_step = std::getenv("PLAMATRIX_AMD_FUSE_STEP");
    pb.fuse_step_env = fuse_step == nullptr ? -1 : (fuse_step[0] == '0' ? 0 : 1);
    pb.fuse_step = fuse_step_wanted(pb);
    return PM_OK;
}

// The diagonal layout is tried for large matrices with few entries per row
// (its detection reads the matrix once and synchronises the stream once);
// PLAMATRIX_AMD_DIA=0 keeps every solve on the CSR kernels.
bool dia_wanted(const pm_csr* a)
{
    const char* env = std::getenv("PLAMATRIX_AMD_DIA");
    if (env != nullptr && env[0] == '0') return false;
    return a->rows >= 65536 && a->rows < (int64_t(1) << 31) && a->nnz <= static_cast<int64_t>(kMaxDia) * a->rows;
}

// Every operand address and size a cached plan's kernel arguments capture.
void append_matrix_key(std::string& key, const pm_csr* m)
{
    char buf[160];
    if (m == nullptr)
    {
        key += "|-";
        return;
    }
    std::snprintf(buf, sizeof(buf), "|%p|%p|%p|%p|%p|%lld|%lld|%lld|%d", (const void*)m->row_offsets,
                  (const void*)m->row_offsets32, (const void*)m->col_indices, (const void*)m->col_indices32, m->values,
                  static_cast<long long>(m->rows), static_cast<long long>(m->cols), static_cast<long long>(m->nnz),
                  m->dtype);
    key += buf;
}

std::string plan_key(const Problem& pb, const char* kind, int64_t first, int64_t count, int refresh)
{
    // workspace addresses are part of the key: a grown workspace must never
    // replay a plan that still points at its old buffers; so are all three
    // matrices' arrays, since a caller may move any one of them between solves
    char buf[512];
    const unsigned long long dia_sig =
        dia_signature(pb.dia) ^ (dia_signature(pb.dia_g) * 31ull) ^ (dia_signature(pb.dia_gt) * 131ull);
    std::snprintf(buf, sizeof(buf), "%s|%p|%p|%p|%p|%p|%p|%d|%lld|%lld|%d|%lld|%d|%d|%d|%llx|%p|%p|%d|%d", kind, pb.r,
                  static_cast<void*>(pb.st), static_cast<void*>(pb.partials), pb.rhs, pb.x, pb.inv_blocks, pb.pre,
                  static_cast<long long>(first), static_cast<long long>(count), refresh,
                  static_cast<long long>(pb.bsize),
                  (pb.nt ? 1 : 0) | (pb.dma ? 2 : 0) | (pb.fuse_direction ? 4 : 0) | (pb.fuse_step ? 8 : 0) |
                      (pb.single_reduction ? 16 : 0) | (pb.cg_fused ? 32 : 0) | (pb.cg_factored ? 64 : 0) |
                      (pb.long_rows ? 128 : 0),
                  pb.dia.active ? 1 : 0, pb.dia.args.ndiag, dia_sig, pb.dia.values, static_cast<const void*>(pb.dia.mask),
                  pb.grid_rows, pb.grid_units);
    std::string key(buf);
    append_matrix_key(key, pb.a);
    append_matrix_key(key, pb.gf);
    append_matrix_key(key, pb.gtf);
    return key;
}

template <typename S>
Plan* get_plan(pm_workspace* ws, const Problem& pb, int what, int64_t first, int64_t count, int refresh)
{
    const std::string key = plan_key(pb, what == 0 ? "init" : "chunk", first, count, refresh);
    Plan* plan = ws->plans->find(key);
    if (!plan)
    {
        auto fresh = std::make_unique<Plan>();
        build_calls<S>(pb, what, first, count, refresh, fresh->calls);
        plan = ws->plans->insert(key, std::move(fresh));
    }
    return plan;
}

struct Policy
{
    int chunk = 16;
    bool graph = true;
    int refresh = 0;
    bool grow = false;  // chunks double up to kMaxGrownChunk (no refresh interval, chunk not set)
};

// Without a refresh interval to divide, a solve's chunks grow 4, 4, 8, 16,
// 32, 64 iterations: short solves stop early (the 5-iteration Schur block-PCG
// of an LM step ran 27 iterations of early-exit kernels with two chunks of 16
// in flight), long launch-bound ones cross the host less often (65,536-row
// FP64 Poisson, 767 iterations: 7.91 ms with chunks of 16, 7.54 ms with 64).
constexpr int kFirstGrownChunk = 4;
constexpr int kMaxGrownChunk = 64;

// By default a chunk of iterations is replayed as a HIP graph below
// kDirectLaunchRows rows, where the kernels are short enough for the launch
// path to matter (65,536-row Poisson: 12.8 vs 13.5 us per iteration), and
// launched directly above, where the graph's own completion and launch cost
// more than the host's launches (2.1M-row diffusion: 21.6 vs 22.2 ms).
constexpr int64_t kDirectLaunchRows = 262144;

Policy resolve_policy(const pm_solve_policy* p, int64_t rows)
{
    Policy out;
    out.graph = rows < kDirectLaunchRows;
    if (p)
    {
        if (p->chunk_iterations > 0) out.chunk = p->chunk_iterations;
        if (p->use_graph >= 0) out.graph = p->use_graph != 0;
        out.refresh = p->refresh_interval > 0 ? p->refresh_interval : 0;
    }
    out.grow = out.refresh == 0 && !(p && p->chunk_iterations > 0);
    if (out.grow) out.chunk = kFirstGrownChunk;
    if (out.chunk % 2) out.chunk += 1;
    if (out.refresh > 0)
    {
        if (out.refresh % 2) out.refresh += 1;
        // a chunk must divide the refresh interval so chunk plans repeat
        int k = std::min(out.chunk, out.refresh);
        while (k > 2 && (out.refresh % k != 0 || k % 2 != 0)) --k;
        out.chunk = std::max(k, 2);
    }
    return out;
}

int finish_report(const SolverState& fin, const pm_solver_options* options, pm_solver_report* report)
{
    pm_solver_report rep{};
    rep.initial_residual = std::sqrt(fin.rr0);
    rep.final_residual = std::sqrt(fin.rr);
    rep.iterations = fin.iteration;
    switch (fin.done)
    {
    case kInitNonFinite: return fail(PM_RUNTIME_ERROR, "iterative solver initial residual is not finite");
    case kBadDiagonal:
        return fail(PM_RUNTIME_ERROR, "pcg Jacobi diagonal is missing, non-positive, or near zero at row %llu",
                    fin.bad_row);
    case kBreakdown:
    case kNonFinite: return fail(PM_RUNTIME_ERROR, "iterative solver breakdown: matrix is not numerically SPD");
    default: break;
    }
    rep.converged = fin.done == kConverged ? 1 : 0;
    if (report) *report = rep;
    if (!rep.converged && options->require_convergence)
        return fail(PM_RUNTIME_ERROR, "iterative solver did not converge in %d iterations; final residual=%g",
                    rep.iterations, rep.final_residual);
    return PM_OK;
}

template <typename S>
int adaptive_solve(const pm_csr* a, const void* rhs, void* x, const void* inv_blocks, int64_t bsize, int pre,
                   const pm_solver_options* options, const pm_solve_policy* policy_in, pm_solver_report* report,
                   pm_workspace* ws, hipStream_t stream, const pm_csr* gf = nullptr, const pm_csr* gtf = nullptr,
                   int tol_rule = 0)
{
    const Policy policy = resolve_policy(policy_in, a->rows);
    Problem pb{};
    pb.a = a;
    pb.rhs = rhs;
    pb.x = x;
    pb.inv_blocks = inv_blocks;
    pb.gf = gf;
    pb.gtf = gtf;
    pb.bsize = pre == kPreBlock ? bsize : 1;
    pb.pre = pre;
    PM_TRY(prepare(ws, pb, stream));
    if (dia_wanted(a))
    {
        // Jacobi on a matrix whose every Jacobi inverse is exactly 1 is the
        // unpreconditioned iteration bit for bit (z = 1 * r = r): it then runs
        // as such, without the inverse-diagonal reads and z writes
        // (PLAMATRIX_AMD_UNIT_DIAGONAL=0 keeps the Jacobi kernels)
        const char* unit_env = std::getenv("PLAMATRIX_AMD_UNIT_DIAGONAL");
        const bool check_unit = pb.pre == kPreJacobi && (unit_env == nullptr || unit_env[0] != '0');
        bool unit = false;
        // A symmetric bit for bit streams only its upper diagonals
        // (PLAMATRIX_AMD_DIA_MIRROR=0 streams every diagonal)
        const char* mirror_env = std::getenv("PLAMATRIX_AMD_DIA_MIRROR");
        const bool mirror = mirror_env == nullptr || mirror_env[0] != '0';
        PM_TRY(dia_build(ws->dia, a, &pb.dia, stream, check_unit ? &unit : nullptr, mirror));
        if (check_unit && unit) pb.pre = kPreNone;
    }
    if (pb.pre == kPreFactored && dia_wanted(gf) && dia_wanted(gtf))
    {
        PM_TRY(dia_build(ws->dia_g, gf, &pb.dia_g, stream));
        PM_TRY(dia_build(ws->dia_gt, gtf, &pb.dia_gt, stream));
        if (!pb.dia_g.active || !pb.dia_gt.active) pb.dia_g.active = pb.dia_gt.active = false;
    }
    pb.fuse_step = fuse_step_wanted(pb);
    pb.single_reduction = single_reduction_wanted(pb);
    {
        // few long rows (under kDirectLaunchRows rows, at least kLongRowEntries
        // entries per row on average): the classic CSR SpMV takes kLongLanes
        // lanes per row and a grid of its own (PLAMATRIX_AMD_LONG_ROWS=0 keeps
        // a wavefront per 64-row chunk)
        const char* env = std::getenv("PLAMATRIX_AMD_LONG_ROWS");
        pb.long_rows = !pb.dia.active && a->rows > 0 && a->rows < kDirectLaunchRows &&
                       a->nnz >= kLongRowEntries * a->rows && (env == nullptr || env[0] != '0');
        if (pb.long_rows)
            pb.grid_rows = static_cast<int>(
                std::min<int64_t>(ceil_div(a->rows, static_cast<int64_t>(kDirBlock / kLongLanes)), kDirGrid));
    }
    // the single-reduction kernels walk 64-row chunks, one per wavefront
    // where the solver grid allows (65,536 rows: 256 workgroups instead of
    // 64); the three-kernel form a restart hands over to uses the same grid
    if (pb.single_reduction) pb.grid_units = solver_grid_chunks(a->rows);
    // the factored three-kernel form a restart hands over to reads the r.r,
    // r.z partials with grid_rows workgroups
    pb.cg_factored = pb.single_reduction && pb.pre == kPreFactored;
    if (pb.cg_factored) pb.grid_rows = pb.grid_units;
    {
        // on the diagonal layout, one kernel per iteration; factored with a
        // four-diagonal G, G r and G^T t in one kernel (PLAMATRIX_AMD_CG_FUSED=0
        // keeps two, and three)
        const char* env = std::getenv("PLAMATRIX_AMD_CG_FUSED");
        bool pair = pb.dia_g.active && pb.dia_g.args.ndiag == 4 && pb.dia_gt.args.ndiag == 4;
        for (int d = 0; pair && d < 4; ++d)
            pair = pb.dia_g.args.off[d] <= 0 && pb.dia_gt.args.off[3 - d] == -pb.dia_g.args.off[d];
        pb.cg_fused = pb.single_reduction && (!pb.cg_factored || pair) && pb.dia.active && pb.r2 != nullptr &&
                      (env == nullptr || env[0] != '0');
    }
    ws->last_diagonals = pb.dia.active ? pb.dia.args.ndiag : 0;
    ws->last_value_streams = pb.dia.active ? pb.dia.streams : 0;
    ws->last_form = pb.cg_factored ? (pb.cg_fused ? 5 : 4) : pb.cg_fused ? 1 : pb.single_reduction ? 2 : 3;
    ws->last_preconditioner = pb.pre;
    if (a->rows == 0)
    {
        pm_solver_report rep{};
        rep.converged = 1;
        if (report) *report = rep;
        return PM_OK;
    }
    hipLaunchKernelGGL(k_state_init, dim3(1), dim3(1), 0, stream, pb.st, options->max_iterations,
                       options->relative_tolerance, options->absolute_tolerance, tol_rule);
    PM_HIP(hipGetLastError());
    PM_TRY(get_plan<S>(ws, pb, 0, 0, 0, 0)->run(stream, false));

    // iteration t = max_iterations must run its first kernel to record the
    // final decision, so chunks cover iterations [0, max_iterations].
    int64_t K = policy.chunk;
    const int64_t needed = static_cast<int64_t>(options->max_iterations) + 1;
    SolverState* hs = ws->host_state;
    SolverState fin{};
    std::deque<int64_t> inflight;
    int64_t launched = 0;
    int64_t chunk_id = 0;
    auto issue = [&]() -> int {
        const int64_t phase = policy.refresh > 0 ? launched % policy.refresh : 0;
        Plan* plan = get_plan<S>(ws, pb, 1, phase, K, policy.refresh);
        PM_TRY(plan->run(stream, policy.graph));
        const int slot = static_cast<int>(chunk_id % 2);
        // the state record goes straight into fine-grained host memory (a
        // one-wave kernel instead of a copy command per chunk)
        hipLaunchKernelGGL(k_publish_state, dim3(1), dim3(kWave), 0, stream, static_cast<const SolverState*>(pb.st),
                           hs + slot);
        PM_HIP(hipGetLastError());
        PM_HIP(hipEventRecord(ws->events[slot], stream));
        inflight.push_back(chunk_id);
        ++chunk_id;
        launched += K;
        // the first two chunks go out together; later ones follow a
        // completed chunk that did not stop the solve
        if (policy.grow && chunk_id >= 2 && K < kMaxGrownChunk) K *= 2;
        return PM_OK;
    };
    PM_TRY(issue());
    while (true)
    {
        if (launched < needed && inflight.size() < 2) PM_TRY(issue());
        const int64_t oldest = inflight.front();
        inflight.pop_front();
        const int slot = static_cast<int>(oldest % 2);
        PM_HIP(hipEventSynchronize(ws->events[slot]));
        fin = hs[slot];
        if (fin.done != kRunning) break;
        // a single-reduction restart: the chunks still to be issued take the
        // three-kernel form (same state, parity slots, partials and vectors)
        if (pb.single_reduction && fin.counters[kCounterRestart] != 0u)
            pb.single_reduction = pb.cg_fused = pb.cg_factored = false;
        if (inflight.empty() && launched >= needed)
            return fail(PM_RUNTIME_ERROR, "iterative solver stopped without a final state");
    }
    PM_HIP(hipEventRecord(ws->done_event, stream));
    return finish_report(fin, options, report);
}

int solve_dispatch(const pm_csr* a, const void* rhs, void* x, const void* inv_blocks, int64_t bsize, int pre,
                   const pm_solver_options* options, const pm_solve_policy* policy, pm_solver_report* report,
                   pm_workspace* ws, hipStream_t stream, const pm_csr* gf = nullptr, const pm_csr* gtf = nullptr,
                   int tol_rule = 0)
{
    PM_TRY(validate_options(options));
    PM_TRY(check_system(a, rhs, x, "iterative solver"));
    if (!ws) return fail(PM_INVALID_ARGUMENT, "iterative solver: null workspace");
    if (a->dtype == PM_F32)
        return adaptive_solve<float>(a, rhs, x, inv_blocks, bsize, pre, options, policy, report, ws, stream, gf, gtf,
                                     tol_rule);
    return adaptive_solve<double>(a, rhs, x, inv_blocks, bsize, pre, options, policy, report, ws, stream, gf, gtf,
                                  tol_rule);
}

// Block-Jacobi PCG with the stopping rule chosen by the caller (tol_rule 1:
// abs + rel ||r0||), for the Schur solve's CPU-backend semantics.
int block_pcg_rule(const pm_csr* a, const void* rhs, void* x, const void* inverse_blocks, int64_t block_size,
                   const pm_solver_options* options, pm_solver_report* report, pm_workspace* ws, hipStream_t stream,
                   int tol_rule)
{
    return solve_dispatch(a, rhs, x, inverse_blocks, block_size, kPreBlock, options, nullptr, report, ws, stream,
                          nullptr, nullptr, tol_rule);
}

// The factors of M^-1 = G^T G: square, same size and dtype as A, with 32-bit
// index mirrors.
int check_factors(const pm_csr* a, const pm_csr* g, const pm_csr* gt)
{
    for (const pm_csr* f : {g, gt})
    {
        PM_TRY(check_csr(f, "factored preconditioner"));
        if (f->rows != a->rows || f->cols != a->cols || f->dtype != a->dtype)
            return fail(PM_INVALID_ARGUMENT, "factored preconditioner: factor shape or dtype differs from the matrix");
        if (!f->row_offsets32 || !f->col_indices32)
            return fail(PM_INVALID_ARGUMENT, "factored preconditioner: factors need 32-bit index mirrors");
    }
    return PM_OK;
}

} // namespace pm

struct pm_fixed_state
{
    double* device = nullptr;  // [initial rr, final rr]
    hipEvent_t event = nullptr;
    hipStream_t stream = nullptr;
    int32_t submitted = 0;
};

using namespace pm;

namespace pm
{
template <typename S>
int fixed_solve(const pm_csr* a, const void* rhs, void* x, int32_t iterations, int pre, pm_fixed_state* state,
                pm_workspace* ws, hipStream_t stream)
{
    Problem pb{};
    pb.a = a;
    pb.rhs = rhs;
    pb.x = x;
    pb.bsize = 1;
    pb.pre = pre;
    PM_TRY(prepare(ws, pb, stream));
    if (a->rows == 0)
    {
        PM_HIP(hipMemsetAsync(state->device, 0, 2 * sizeof(double), stream));
    }
    else
    {
        // tolerance 0: every submitted iteration runs unless r becomes exactly 0
        hipLaunchKernelGGL(k_state_init, dim3(1), dim3(1), 0, stream, pb.st, iterations, 0.0, 0.0, 0);
        PM_HIP(hipGetLastError());
        std::vector<KernelCall> calls;
        build_calls<S>(pb, 0, 0, 0, 0, calls);
        build_calls<S>(pb, 1, 0, iterations, 0, calls);
        for (const KernelCall& k : calls) PM_TRY(launch(k, stream));
        hipLaunchKernelGGL(k_fixed_out, dim3(1), dim3(kBlock), 0, stream, pb.st, pb.partials, pb.grid_units,
                           iterations == 0 ? 1 : 0, state->device);
        PM_HIP(hipGetLastError());
    }
    PM_HIP(hipEventRecord(state->event, stream));
    PM_HIP(hipEventRecord(ws->done_event, stream));
    return PM_OK;
}
} // namespace pm

extern "C" {

int pm_cg(const pm_csr* a, const void* rhs, void* x, const pm_solver_options* options,
          const pm_solve_policy* policy, pm_solver_report* report, pm_workspace* ws, hipStream_t stream)
{
    return solve_dispatch(a, rhs, x, nullptr, 1, kPreNone, options, policy, report, ws, stream);
}

int pm_pcg(const pm_csr* a, const void* rhs, void* x, const pm_solver_options* options,
           const pm_solve_policy* policy, pm_solver_report* report, pm_workspace* ws, hipStream_t stream)
{
    PM_TRY(validate_options(options));
    return solve_dispatch(a, rhs, x, nullptr, 1, options->use_jacobi ? kPreJacobi : kPreNone, options, policy, report,
                          ws, stream);
}

int pm_block_pcg(const pm_csr* a, const void* rhs, void* x, const void* inverse_blocks, int64_t block_size,
                 const pm_solver_options* options, const pm_solve_policy* policy, pm_solver_report* report,
                 pm_workspace* ws, hipStream_t stream)
{
    PM_TRY(check_csr(a, "block PCG"));
    if (block_size <= 0 || block_size > kMaxBlock || a->rows % block_size != 0 || (a->rows > 0 && !inverse_blocks))
        return fail(PM_INVALID_ARGUMENT, "block PCG requires complete row-major inverse diagonal blocks "
                                         "(block size 1..%d dividing the row count)", kMaxBlock);
    if (inverse_blocks && (inverse_blocks == rhs || inverse_blocks == x))
        return fail(PM_INVALID_ARGUMENT, "block PCG inverse blocks must not alias rhs or solution");
    return solve_dispatch(a, rhs, x, inverse_blocks, block_size, kPreBlock, options, policy, report, ws, stream);
}

int pm_solve_fixed_async(const pm_csr* a, const void* rhs, void* x, int32_t iterations, int32_t preconditioned,
                         pm_fixed_state** state, pm_workspace* ws, hipStream_t stream)
{
    if (!state) return fail(PM_INVALID_ARGUMENT, "fixed solve: null state output");
    *state = nullptr;
    if (iterations < 0) return fail(PM_INVALID_ARGUMENT, "fixed iterative solver iterations must be non-negative");
    PM_TRY(check_system(a, rhs, x, "fixed iterative solver"));
    if (!ws) return fail(PM_INVALID_ARGUMENT, "fixed iterative solver: null workspace");
    auto* s = new pm_fixed_state();
    s->stream = stream;
    s->submitted = iterations;
    hipError_t err = hipMalloc(reinterpret_cast<void**>(&s->device), 2 * sizeof(double));
    if (err == hipSuccess) err = hipEventCreateWithFlags(&s->event, hipEventDisableTiming);
    if (err != hipSuccess)
    {
        pm_fixed_state_destroy(s);
        return check_hip(err, "fixed state allocation", __FILE__, __LINE__);
    }
    const int pre = preconditioned ? kPreJacobi : kPreNone;
    const int status = a->dtype == PM_F32 ? fixed_solve<float>(a, rhs, x, iterations, pre, s, ws, stream)
                                          : fixed_solve<double>(a, rhs, x, iterations, pre, s, ws, stream);
    if (status != PM_OK)
    {
        pm_fixed_state_destroy(s);
        return status;
    }
    *state = s;
    return PM_OK;
}

int pm_fixed_finalize(const pm_fixed_state* state, const pm_solver_options* options, pm_solver_report* report)
{
    PM_TRY(validate_options(options));
    if (!state || !state->event) return fail(PM_LOGIC_ERROR, "fixed iterative solver state is empty or closed");
    const hipError_t q = hipEventQuery(state->event);
    if (q == hipErrorNotReady)
        return fail(PM_LOGIC_ERROR, "finalizeIterativeSolverReport requires caller stream synchronization");
    if (q != hipSuccess) return check_hip(q, "hipEventQuery", __FILE__, __LINE__);
    double host[2] = {0.0, 0.0};
    PM_HIP(hipMemcpy(host, state->device, sizeof(host), hipMemcpyDeviceToHost));
    if (!std::isfinite(host[0]) || !std::isfinite(host[1]) || host[0] < 0.0 || host[1] < 0.0)
        return fail(PM_RUNTIME_ERROR, "fixed iterative solver breakdown: residual is not finite");
    pm_solver_report rep{};
    rep.iterations = state->submitted;
    rep.initial_residual = std::sqrt(host[0]);
    rep.final_residual = std::sqrt(host[1]);
    const double tol = std::max(options->absolute_tolerance, options->relative_tolerance * rep.initial_residual);
    rep.converged = rep.final_residual <= tol ? 1 : 0;
    if (report) *report = rep;
    if (!rep.converged && options->require_convergence)
        return fail(PM_RUNTIME_ERROR, "fixed iterative solver did not converge in %d iterations; final residual=%g",
                    rep.iterations, rep.final_residual);
    return PM_OK;
}

int pm_fixed_state_destroy(pm_fixed_state* state)
{
    if (!state) return PM_OK;
    if (state->event)
    {
        (void)hipEventSynchronize(state->event);
        (void)hipEventDestroy(state->event);
    }
    if (state->device) (void)hipFree(state->device);
    delete state;
    return PM_OK;
}

} // extern "C"

// ---- FP32 inner solve, FP64 outer refinement --------------------------------
namespace pm
{

// r = b - A x in FP64 (FP32 matrix, FP64 vectors) plus partial ||r||^2, ||b||^2.
template <typename I>
__global__ __launch_bounds__(kBlock) void k_residual64(int64_t rows, const I* __restrict__ ro,
                                                       const I* __restrict__ ci, const float* __restrict__ v,
                                                       const double* __restrict__ b, const double* __restrict__ x,
                                                       double* __restrict__ r, double* partials, unsigned* counter,
                                                       double* out)
{
    __shared__ float lv_all[kWavesPerBlock][kStreamCap];
    __shared__ I lc_all[kWavesPerBlock][kStreamCap];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    double rr = 0.0, bb = 0.0;
    PM_FOR_CHUNKS(c, chunk_count(rows), kWavesPerBlock)
    {
        const int64_t row0 = c * kChunkRows;
        int64_t s, e;
        const double acc = chunk_row_sums<double>(row0, rows, ro, ci, lv_all[wave], lc_all[wave],
                                                  Fp64Source<float, double>{v, x}, s, e);
        const int64_t row = row0 + lane;
        if (row < rows)
        {
            const double res = b[row] - acc;
            r[row] = res;
            rr = __builtin_fma(res, res, rr);
            bb = __builtin_fma(b[row], b[row], bb);
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

__global__ void k_scale_to_f32(int64_t n, const double* __restrict__ r, double scale, float* __restrict__ rf,
                               float* __restrict__ d)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    {
        rf[i] = static_cast<float>(r[i] * scale);
        d[i] = 0.0f;
    }
}

__global__ void k_accumulate(int64_t n, double* __restrict__ x, const float* __restrict__ d, double scale)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        x[i] = __builtin_fma(scale, static_cast<double>(d[i]), x[i]);
}

__global__ void k_zero_u32(unsigned* p) { *p = 0u; }

int refine(const pm_csr* a, const double* b, double* x, const pm_solver_options* inner, double target,
           int32_t max_outer, const pm_solve_policy* policy, pm_refine_report* report, pm_workspace* ws,
           hipStream_t stream, const pm_csr* gf = nullptr, const pm_csr* gtf = nullptr)
{
    const int64_t n = a->rows;
    const size_t r_bytes = (sizeof(double) * static_cast<size_t>(n) + 255) & ~size_t(255);
    const size_t f_bytes = (sizeof(float) * static_cast<size_t>(n) + 255) & ~size_t(255);
    PM_TRY(ws->refine.reserve(r_bytes + 2 * f_bytes + 4096));
    PM_TRY(ws->partials.reserve(sizeof(double) * kPartialsDoubles));
    char* base = ws->refine.as<char>();
    double* r64 = reinterpret_cast<double*>(base);
    float* rf = reinterpret_cast<float*>(base + r_bytes);
    float* d = reinterpret_cast<float*>(base + r_bytes + f_bytes);
    unsigned* counter = reinterpret_cast<unsigned*>(base + r_bytes + 2 * f_bytes);
    double* out = reinterpret_cast<double*>(base + r_bytes + 2 * f_bytes + 256);
    // the refinement residual uses the upper half of the partial buffer, away
    // from the solver's reduction slots
    double* rpartials = ws->partials.as<double>() + 4 * kMaxPartials;
    const int grid = static_cast<int>(std::min<int64_t>(ceil_div(n > 0 ? n : 1, kBlock), 8192));
    const int rgrid = solver_grid_chunks(n);

    pm_solver_options in = *inner;
    in.require_convergence = 0;
    in.absolute_tolerance = 0.0;
    pm_refine_report rep{};
    for (int outer = 0;; ++outer)
    {
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, stream, counter);
        if (use_compact(a))
            hipLaunchKernelGGL((k_residual64<int32_t>), dim3(rgrid), dim3(kBlock), 0, stream, n, a->row_offsets32,
                               a->col_indices32, static_cast<const float*>(a->values), b, x, r64, rpartials, counter,
                               out);
        else
            hipLaunchKernelGGL((k_residual64<int64_t>), dim3(rgrid), dim3(kBlock), 0, stream, n, a->row_offsets,
                               a->col_indices, static_cast<const float*>(a->values), b, x, r64, rpartials, counter,
                               out);
        PM_HIP(hipGetLastError());
        double host[2];
        PM_HIP(hipMemcpyAsync(host, out, sizeof(host), hipMemcpyDeviceToHost, stream));
        PM_HIP(hipStreamSynchronize(stream));
        if (!std::isfinite(host[0])) return fail(PM_RUNTIME_ERROR, "refinement residual is not finite");
        const double relres = host[1] > 0.0 ? std::sqrt(host[0] / host[1]) : std::sqrt(host[0]);
        if (outer == 0) rep.initial_relres = relres;
        rep.final_relres = relres;
        rep.outer_iterations = outer;
        if (relres <= target || host[0] == 0.0)
        {
            rep.converged = 1;
            break;
        }
        if (outer >= max_outer) break;
        const double norm = std::sqrt(host[0]);
        hipLaunchKernelGGL(k_scale_to_f32, dim3(grid), dim3(kBlock), 0, stream, n, r64, 1.0 / norm, rf, d);
        PM_HIP(hipGetLastError());
        pm_solver_report inner_rep{};
        in.relative_tolerance = refine_pass_tolerance(inner->relative_tolerance, relres, target);
        const int pre = gf ? kPreFactored : (in.use_jacobi ? kPreJacobi : kPreNone);
        PM_TRY(adaptive_solve<float>(a, rf, d, nullptr, 1, pre, &in, policy, &inner_rep, ws, stream, gf, gtf));
        rep.inner_iterations += inner_rep.iterations;
        hipLaunchKernelGGL(k_accumulate, dim3(grid), dim3(kBlock), 0, stream, n, x, d, norm);
        PM_HIP(hipGetLastError());
    }
    PM_HIP(hipEventRecord(ws->done_event, stream));
    if (report) *report = rep;
    return PM_OK;
}

} // namespace pm

extern "C" int pm_pcg_refine(const pm_csr* a, const double* rhs, double* x, const pm_solver_options* inner,
                             double target_relres, int32_t max_outer, const pm_solve_policy* policy,
                             pm_refine_report* report, pm_workspace* ws, hipStream_t stream)
{
    PM_TRY(pm::validate_options(inner));
    PM_TRY(pm::check_system(a, rhs, x, "refinement"));
    if (a->dtype != PM_F32) return pm::fail(PM_INVALID_ARGUMENT, "refinement expects an FP32 matrix");
    if (!ws) return pm::fail(PM_INVALID_ARGUMENT, "refinement: null workspace");
    if (!(target_relres >= 0.0) || max_outer < 0) return pm::fail(PM_INVALID_ARGUMENT, "refinement: invalid target");
    PM_TRY(ws->bind(stream));
    PM_TRY(ws->ensure_events());
    if (a->rows == 0)
    {
        if (report) *report = pm_refine_report{1, 0, 0, 0.0, 0.0};
        return PM_OK;
    }
    return pm::refine(a, rhs, x, inner, target_relres, max_outer, policy, report, ws, stream);
}

extern "C" int pm_pcg_factored(const pm_csr* a, const void* rhs, void* x, const pm_csr* g, const pm_csr* gt,
                               const pm_solver_options* options, const pm_solve_policy* policy,
                               pm_solver_report* report, pm_workspace* ws, hipStream_t stream)
{
    PM_TRY(pm::check_csr(a, "factored pcg"));
    PM_TRY(pm::check_factors(a, g, gt));
    return pm::solve_dispatch(a, rhs, x, nullptr, 1, pm::kPreFactored, options, policy, report, ws, stream, g, gt);
}

extern "C" int pm_pcg_refine_factored(const pm_csr* a, const double* rhs, double* x, const pm_csr* g,
                                      const pm_csr* gt, const pm_solver_options* inner, double target_relres,
                                      int32_t max_outer, const pm_solve_policy* policy, pm_refine_report* report,
                                      pm_workspace* ws, hipStream_t stream)
{
    PM_TRY(pm::validate_options(inner));
    PM_TRY(pm::check_system(a, rhs, x, "refinement"));
    if (a->dtype != PM_F32) return pm::fail(PM_INVALID_ARGUMENT, "refinement expects an FP32 matrix");
    PM_TRY(pm::check_factors(a, g, gt));
    if (!ws) return pm::fail(PM_INVALID_ARGUMENT, "refinement: null workspace");
    if (!(target_relres >= 0.0) || max_outer < 0) return pm::fail(PM_INVALID_ARGUMENT, "refinement: invalid target");
    PM_TRY(ws->bind(stream));
    PM_TRY(ws->ensure_events());
    if (a->rows == 0)
    {
        if (report) *report = pm_refine_report{1, 0, 0, 0.0, 0.0};
        return PM_OK;
    }
    return pm::refine(a, rhs, x, inner, target_relres, max_outer, policy, report, ws, stream, g, gt);
}
