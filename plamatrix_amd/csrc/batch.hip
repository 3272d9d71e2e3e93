// Batched Jacobi-PCG / CG over many independent SPD systems in one set of
// launches (no counterpart in the reference, whose GPU solver handles one
// system per call with a host convergence check per iteration,
// iterative_solver.cu:235-402).
//
// Layout: the systems are the diagonal blocks of one CSR matrix.  System s
// owns rows [row_begin[s], row_end[s]); every row_begin is a multiple of 64,
// so each 64-row wavefront chunk belongs to exactly one system (rows between
// one system's end and the next start are padding and are never touched).
// Every kernel walks all chunks of all systems grid-stride, so a batch of
// 512 uneven systems fills the chip the way one large system does.
//
// Per iteration, graph-replayed in chunks of iterations:
//   direction  p = z + beta_s p, q = A p, per-chunk FP64 partial p.q
//   alpha      one wavefront per system sums its chunk partials in a fixed
//              order: alpha_s = rho_s / (p.q)_s (breakdown rules of alphaKernel,
//              iterative_solver_kernels.cuh:169-202)
//   update     x += alpha p, r -= alpha q, z = M^-1 r, partial r.r and r.z
//   beta       per system: convergence test against max(abs, rel ||r0||),
//              iteration limit, beta rules of betaKernel (:204-238)
// Every refresh_interval iterations the update is replaced by x += alpha p and
// r = b - A x accumulated in FP64.  Each system carries its own state and
// stop flag on the device; a system that has stopped costs one flag read per
// chunk.  The host only polls the count of stopped systems once per chunk of
// iterations, with two chunks in flight.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "plan_cache.hpp"
#include "solver_internal.hpp"
#include "spmv_kernels.hpp"
#include "tuning.hpp"

namespace pm
{

constexpr int kBatchWaves = kWavesPerBlock;  // 4 wavefronts per 256-thread workgroup
constexpr int kBatchGrid = 2048;

struct BatchState
{
    double rho;
    double alpha;
    double beta;
    double rr;
    double rr0;
    double tol;
    int32_t iteration;
    int32_t done;
    unsigned long long bad_row;
    double den;  // single-reduction form: p.Ap of the last direction
};

// Chunks of the systems still running, as a device list with its length in
// device memory (so replayed graphs follow every rebuild): a superset of the
// running systems' chunks, since systems only stop between rebuilds.  The
// list is dealt round-robin over the workgroups: chunk costs differ by more
// than ten times between members (a reduced Schur row holds ~84 entries, a
// Laplacian row 5), and contiguous per-workgroup ranges measured 897 ms
// against 538 ms on the 512-system batch.
#define PM_FOR_LISTED(i, list_count, W)                                                               \
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * (W) + threadIdx.x / kWave,                    \
                 i##_n = static_cast<int64_t>(*(list_count));                                         \
         i < i##_n; i += static_cast<int64_t>(gridDim.x) * (W))

// keep_bad_row: the preconditioner's inverses are reused from the previous
// solve of the same matrix, so are the rows they rejected.
__global__ __launch_bounds__(kBlock) void k_batch_state_init(int64_t systems, BatchState* st,
                                                             unsigned long long* done_count, int keep_bad_row)
{
    for (int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; s < systems;
         s += static_cast<int64_t>(gridDim.x) * blockDim.x)
    {
        BatchState b{};
        b.bad_row = keep_bad_row ? st[s].bad_row : ~0ull;
        b.done = kRunning;
        st[s] = b;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *done_count = 0ull;
}

// r = b - A x (FP64 accumulation), z = M^-1 r (JACOBI: inverse diagonal,
// rejected like jacobiInverse, iterative_solver_cpu.cpp:127-160), INIT: p = 0;
// per-chunk partial r.r and r.z.  Stopped systems are skipped.
template <typename S, typename I, bool JACOBI, bool INIT>
__global__ __launch_bounds__(kBlock) void k_batch_residual(int64_t chunks, const int32_t* __restrict__ chunk_sys,
                                                           const int64_t* __restrict__ sys_end,
                                                           const I* __restrict__ ro, const I* __restrict__ ci,
                                                           const S* __restrict__ v, const S* __restrict__ b,
                                                           const S* __restrict__ x, S* __restrict__ r,
                                                           S* __restrict__ inv, S* __restrict__ z,
                                                           S* __restrict__ p, BatchState* st,
                                                           double* __restrict__ prr, double* __restrict__ prz,
                                                           const int64_t* __restrict__ list,
                                                           const unsigned long long* __restrict__ list_count)
{
    __shared__ S lv_all[kBatchWaves][kStreamCap];
    __shared__ I lc_all[kBatchWaves][kStreamCap];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    PM_FOR_LISTED(li, list_count, kBatchWaves)
    {
        const int64_t c = list[li];
        const int32_t s = chunk_sys[c];
        if (s < 0 || st[s].done != kRunning)
        {
            if (lane == 0) prr[c] = prz[c] = 0.0;
            continue;
        }
        const int64_t end = sys_end[s];
        const int64_t row0 = c * kChunkRows;
        int64_t ks, ke;
        const double acc =
            chunk_row_sums<double>(row0, end, ro, ci, lv_all[wave], lc_all[wave], Fp64Source<S, S>{v, x}, ks, ke);
        const int64_t row = row0 + lane;
        double rr = 0.0, rz = 0.0;
        if (row < end)
        {
            const S ri = static_cast<S>(static_cast<double>(b[row]) - acc);
            r[row] = ri;
            if (INIT) p[row] = S(0);
            S zi = ri;
            if (JACOBI)
            {
                zi = inv[row] * ri;
            }
            z[row] = zi;
            rr = static_cast<double>(ri) * static_cast<double>(ri);
            rz = static_cast<double>(ri) * static_cast<double>(zi);
        }
        rr = wave_sum(rr);
        rz = wave_sum(rz);
        if (lane == 0)
        {
            prr[c] = rr;
            prz[c] = rz;
        }
    }
}

// p_new = z + beta p_old for the chunks of running systems (SPLIT
// direction: the SpMV then gathers p_new once per entry).
template <typename S>
__global__ __launch_bounds__(kBlock) void k_batch_pupdate(const int32_t* __restrict__ chunk_sys,
                                                          const int64_t* __restrict__ sys_end,
                                                          const S* __restrict__ z, const S* __restrict__ p_old,
                                                          S* __restrict__ p_new, const BatchState* st,
                                                          const int64_t* __restrict__ list,
                                                          const unsigned long long* __restrict__ list_count)
{
    const int lane = threadIdx.x % kWave;
    PM_FOR_LISTED(li, list_count, kBatchWaves)
    {
        const int64_t c = list[li];
        const int32_t s = chunk_sys[c];
        if (s < 0 || st[s].done != kRunning) continue;
        const int64_t row = c * kChunkRows + lane;
        if (row < sys_end[s]) p_new[row] = fma_s(static_cast<S>(st[s].beta), p_old[row], z[row]);
    }
}

// p_new = z + beta p_old, q = A p_new, per-chunk FP64 partial p_new.q.
// SPLIT: p_new was formed by k_batch_pupdate and is gathered directly.
template <typename S, typename I, bool SPLIT, int LL = 8>
__global__ __launch_bounds__(kBlock) void k_batch_direction(int64_t chunks, const int32_t* __restrict__ chunk_sys,
                                                            const int64_t* __restrict__ sys_end,
                                                            const I* __restrict__ ro, const I* __restrict__ ci,
                                                            const S* __restrict__ v, const S* __restrict__ z,
                                                            const S* __restrict__ p_old, S* __restrict__ p_new,
                                                            S* __restrict__ q, const BatchState* st,
                                                            double* __restrict__ pden,
                                                            const int64_t* __restrict__ list,
                                                            const unsigned long long* __restrict__ list_count)
{
    __shared__ S lv_all[kBatchWaves][kStreamCap];
    __shared__ I lc_all[kBatchWaves][kStreamCap];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    PM_FOR_LISTED(li, list_count, kBatchWaves)
    {
        const int64_t c = list[li];
        const int32_t s = chunk_sys[c];
        if (s < 0 || st[s].done != kRunning)
        {
            if (lane == 0) pden[c] = 0.0;
            continue;
        }
        const int64_t end = sys_end[s];
        const int64_t row0 = c * kChunkRows;
        const S beta = static_cast<S>(st[s].beta);
        int64_t ks, ke;
        S acc;
        if (SPLIT)
            acc = chunk_row_sums<S, I, SpmvSource<S>, false, false, LL>(row0, end, ro, ci, lv_all[wave], lc_all[wave],
                                                                        SpmvSource<S>{v, p_new}, ks, ke);
        else
            acc = chunk_row_sums<S, I, DirectionSource<S>, false, false, LL>(
                row0, end, ro, ci, lv_all[wave], lc_all[wave], DirectionSource<S>{v, z, p_old, beta}, ks, ke);
        const int64_t row = row0 + lane;
        double den = 0.0;
        if (row < end)
        {
            S pn;
            if (SPLIT)
            {
                pn = p_new[row];
            }
            else
            {
                pn = fma_s(beta, p_old[row], z[row]);
                p_new[row] = pn;
            }
            q[row] = acc;
            den = static_cast<double>(pn) * static_cast<double>(acc);
        }
        den = wave_sum(den);
        if (lane == 0) pden[c] = den;
    }
}

// x += alpha p; STEP: r -= alpha q, z = M^-1 r, partial r.r and r.z.
// Systems with a block size above 1 (sys_bs, null: none) leave z and r.z to
// k_batch_block_apply, which runs next.
template <typename S, bool JACOBI, bool STEP>
__global__ __launch_bounds__(kBlock) void k_batch_update(int64_t chunks, const int32_t* __restrict__ chunk_sys,
                                                         const int64_t* __restrict__ sys_end, S* __restrict__ x,
                                                         S* __restrict__ r, S* __restrict__ z,
                                                         const S* __restrict__ p, const S* __restrict__ q,
                                                         const S* __restrict__ inv, const BatchState* st,
                                                         double* __restrict__ prr, double* __restrict__ prz,
                                                         const int64_t* __restrict__ list,
                                                         const unsigned long long* __restrict__ list_count,
                                                         const int32_t* __restrict__ sys_bs)
{
    const int lane = threadIdx.x % kWave;
    PM_FOR_LISTED(li, list_count, kBatchWaves)
    {
        const int64_t c = list[li];
        const int32_t s = chunk_sys[c];
        if (s < 0 || st[s].done != kRunning)
        {
            if (STEP && lane == 0) prr[c] = prz[c] = 0.0;
            continue;
        }
        const S alpha = static_cast<S>(st[s].alpha);
        const bool blocks = sys_bs != nullptr && sys_bs[s] > 1;
        const int64_t row = c * kChunkRows + lane;
        double rr = 0.0, rz = 0.0;
        if (row < sys_end[s])
        {
            x[row] = fma_s(alpha, p[row], x[row]);
            if (STEP)
            {
                const S ri = fma_s(-alpha, q[row], r[row]);
                r[row] = ri;
                rr = static_cast<double>(ri) * static_cast<double>(ri);
                if (!blocks)
                {
                    const S zi = JACOBI ? inv[row] * ri : ri;
                    z[row] = zi;
                    rz = static_cast<double>(ri) * static_cast<double>(zi);
                }
            }
        }
        if (STEP)
        {
            rr = wave_sum(rr);
            rz = wave_sum(rz);
            if (lane == 0)
            {
                prr[c] = rr;
                prz[c] = rz;
            }
        }
    }
}

// ---- factored preconditioner M_s^-1 = G_s^T G_s per system ------------------------
//
// G (lower) and G^T are block-diagonal over the same row ranges as A.  A
// system flagged in sys_jacobi (device, one byte per system; null: none) takes
// Jacobi instead (z = D^-1 r with the inverse diagonal of k_batch_jacobi_inverse).

// t = G r_new for the chunks of running factored systems.  STEP: alpha_s from
// the state, x += alpha_s p on the chunk's rows (every system), and
// r_new = r - alpha_s q formed at every gathered column with the rounding of
// k_batch_update (r itself is rewritten by k_batch_factor_upper, since other
// chunks gather it here); otherwise r_new = r.
template <typename S, typename I, bool STEP, int LL = 8>
__global__ __launch_bounds__(kBlock) void k_batch_factor_lower(const int32_t* __restrict__ chunk_sys,
                                                               const int64_t* __restrict__ sys_end,
                                                               const I* __restrict__ gro, const I* __restrict__ gci,
                                                               const S* __restrict__ gv, S* __restrict__ x,
                                                               const S* __restrict__ r, const S* __restrict__ p,
                                                               const S* __restrict__ q, S* __restrict__ t,
                                                               const BatchState* st,
                                                               const uint8_t* __restrict__ sys_jacobi,
                                                               const int64_t* __restrict__ list,
                                                               const unsigned long long* __restrict__ list_count)
{
    __shared__ S lv_all[kBatchWaves][kStreamCap];
    __shared__ I lc_all[kBatchWaves][kStreamCap];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    PM_FOR_LISTED(li, list_count, kBatchWaves)
    {
        const int64_t c = list[li];
        const int32_t s = chunk_sys[c];
        if (s < 0 || st[s].done != kRunning) continue;
        const int64_t end = sys_end[s];
        const int64_t row0 = c * kChunkRows;
        const int64_t row = row0 + lane;
        const S alpha = STEP ? static_cast<S>(st[s].alpha) : S(0);
        if (!(sys_jacobi && sys_jacobi[s]))
        {
            int64_t ks, ke;
            S acc;
            if (STEP)
                acc = chunk_row_sums<S, I, StepSource<S>, false, false, LL>(row0, end, gro, gci, lv_all[wave],
                                                                            lc_all[wave],
                                                                            StepSource<S>{gv, r, q, -alpha}, ks, ke);
            else
                acc = chunk_row_sums<S, I, SpmvSource<S>, false, false, LL>(row0, end, gro, gci, lv_all[wave],
                                                                            lc_all[wave], SpmvSource<S>{gv, r}, ks,
                                                                            ke);
            if (row < end) t[row] = acc;
        }
        if (STEP && row < end) x[row] = fma_s(alpha, p[row], x[row]);
    }
}

// The step of the factored iteration on its own (systems with long factor
// rows): x += alpha_s p, r -= alpha_s q, with the rounding of k_batch_update,
// so the factor products gather r once per entry instead of r and q.
template <typename S>
__global__ __launch_bounds__(kBlock) void k_batch_step_xr(const int32_t* __restrict__ chunk_sys,
                                                          const int64_t* __restrict__ sys_end, S* __restrict__ x,
                                                          S* __restrict__ r, const S* __restrict__ p,
                                                          const S* __restrict__ q, const BatchState* st,
                                                          const int64_t* __restrict__ list,
                                                          const unsigned long long* __restrict__ list_count)
{
    const int lane = threadIdx.x % kWave;
    PM_FOR_LISTED(li, list_count, kBatchWaves)
    {
        const int64_t c = list[li];
        const int32_t s = chunk_sys[c];
        if (s < 0 || st[s].done != kRunning) continue;
        const int64_t row = c * kChunkRows + lane;
        if (row >= sys_end[s]) continue;
        const S alpha = static_cast<S>(st[s].alpha);
        x[row] = fma_s(alpha, p[row], x[row]);
        r[row] = fma_s(-alpha, q[row], r[row]);
    }
}

// z = G^T t (Jacobi systems: z = D^-1 r_new) and the per-chunk partials r.r,
// r.z.  STEP: r_new = r - alpha_s q on the chunk's own rows, written back.
template <typename S, typename I, bool STEP, int LL = 8>
__global__ __launch_bounds__(kBlock) void k_batch_factor_upper(const int32_t* __restrict__ chunk_sys,
                                                               const int64_t* __restrict__ sys_end,
                                                               const I* __restrict__ uro, const I* __restrict__ uci,
                                                               const S* __restrict__ uv, const S* __restrict__ t,
                                                               S* __restrict__ r, const S* __restrict__ q,
                                                               const S* __restrict__ inv, S* __restrict__ z,
                                                               const BatchState* st,
                                                               const uint8_t* __restrict__ sys_jacobi,
                                                               double* __restrict__ prr, double* __restrict__ prz,
                                                               const int64_t* __restrict__ list,
                                                               const unsigned long long* __restrict__ list_count)
{
    __shared__ S lv_all[kBatchWaves][kStreamCap];
    __shared__ I lc_all[kBatchWaves][kStreamCap];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    PM_FOR_LISTED(li, list_count, kBatchWaves)
    {
        const int64_t c = list[li];
        const int32_t s = chunk_sys[c];
        if (s < 0 || st[s].done != kRunning)
        {
            if (lane == 0) prr[c] = prz[c] = 0.0;
            continue;
        }
        const int64_t end = sys_end[s];
        const int64_t row0 = c * kChunkRows;
        const int64_t row = row0 + lane;
        const bool jac = sys_jacobi && sys_jacobi[s];
        S acc = S(0);
        if (!jac)
        {
            int64_t ks, ke;
            acc = chunk_row_sums<S, I, SpmvSource<S>, false, false, LL>(row0, end, uro, uci, lv_all[wave],
                                                                        lc_all[wave], SpmvSource<S>{uv, t}, ks, ke);
        }
        double rr = 0.0, rz = 0.0;
        if (row < end)
        {
            S ri = r[row];
            if (STEP)
            {
                ri = fma_s(-static_cast<S>(st[s].alpha), q[row], ri);
                r[row] = ri;
            }
            const S zi = jac ? inv[row] * ri : acc;
            z[row] = zi;
            rr = static_cast<double>(ri) * static_cast<double>(ri);
            rz = static_cast<double>(ri) * static_cast<double>(zi);
        }
        rr = wave_sum(rr);
        rz = wave_sum(rz);
        if (lane == 0)
        {
            prr[c] = rr;
            prz[c] = rz;
        }
    }
}

// Rebuild the list of chunks whose system is still running: per workgroup
// an ordered compaction in LDS, one atomic per workgroup for its offset.
__global__ __launch_bounds__(kBlock) void k_batch_list_reset(unsigned long long* list_count) { *list_count = 0ull; }

__global__ __launch_bounds__(kBlock) void k_batch_list(int64_t chunks, const int32_t* __restrict__ chunk_sys,
                                                       const BatchState* __restrict__ st, int64_t* __restrict__ list,
                                                       unsigned long long* list_count)
{
    __shared__ unsigned wave_counts[kBlock / kWave];
    __shared__ unsigned long long base;
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    for (int64_t c0 = static_cast<int64_t>(blockIdx.x) * kBlock; c0 < chunks;
         c0 += static_cast<int64_t>(gridDim.x) * kBlock)
    {
        const int64_t c = c0 + threadIdx.x;
        bool on = false;
        if (c < chunks)
        {
            const int32_t s = chunk_sys[c];
            on = s >= 0 && st[s].done == kRunning;
        }
        const unsigned long long ballot = __ballot(on);
        const unsigned before = static_cast<unsigned>(__popcll(ballot & ((1ull << lane) - 1ull)));
        if (lane == 0) wave_counts[wave] = static_cast<unsigned>(__popcll(ballot));
        __syncthreads();
        unsigned offset = 0, total = 0;
        for (int w = 0; w < kBlock / kWave; ++w)
        {
            if (w < wave) offset += wave_counts[w];
            total += wave_counts[w];
        }
        if (threadIdx.x == 0) base = total ? atomicAdd(list_count, static_cast<unsigned long long>(total)) : 0ull;
        __syncthreads();
        if (on) list[base + offset + before] = c;
        __syncthreads();
    }
}

// Jacobi inverse diagonal of every system row (jacobiInverse,
// iterative_solver_cpu.cpp:127-160): eight lanes per row scan its entries
// (coalesced), the largest magnitude sets the threshold; a row with one
// diagonal entry takes it as is, several are summed in CSR order.  Rows of a
// rejected diagonal get NaN and the system records its first such row.
// only (device, one byte per system; null: every system) limits it to the
// systems flagged; block-Jacobi systems (sys_bs above 1) are skipped.
template <typename S, typename I>
__global__ __launch_bounds__(kBlock) void k_batch_jacobi_inverse(int64_t chunks, const int32_t* __restrict__ chunk_sys,
                                                                 const int64_t* __restrict__ sys_end,
                                                                 const I* __restrict__ ro, const I* __restrict__ ci,
                                                                 const S* __restrict__ v, S* __restrict__ inv,
                                                                 BatchState* st, const uint8_t* __restrict__ only,
                                                                 const int32_t* __restrict__ sys_bs)
{
    constexpr int L = 8;
    const int lane = threadIdx.x % kWave;
    const int sub = lane % L;
    PM_FOR_CHUNKS(c, chunks, kBatchWaves)
    {
        const int32_t s = chunk_sys[c];
        if (s < 0 || (only && !only[s]) || (sys_bs && sys_bs[s] > 1)) continue;
        const int64_t end = sys_end[s];
        for (int pass = 0; pass < L; ++pass)
        {
            const int64_t row = c * kChunkRows + pass * L + lane / L;
            const bool valid = row < end;
            double scale = 0.0, diag = 0.0;
            int count = 0;
            int64_t k0 = 0, k1 = 0;
            if (valid)
            {
                k0 = static_cast<int64_t>(ro[row]);
                k1 = static_cast<int64_t>(ro[row + 1]);
                for (int64_t k = k0 + sub; k < k1; k += L)
                {
                    const double a = static_cast<double>(v[k]);
                    scale = fmax(scale, fabs(a));
                    if (static_cast<int64_t>(ci[k]) == row)
                    {
                        diag = a;
                        ++count;
                    }
                }
            }
#pragma unroll
            for (int off = L / 2; off > 0; off >>= 1)
            {
                scale = fmax(scale, __shfl_xor(scale, off, L));
                diag += __shfl_xor(diag, off, L);
                count += __shfl_xor(count, off, L);
            }
            if (valid && sub == 0)
            {
                if (count > 1)
                {
                    diag = 0.0;
                    for (int64_t k = k0; k < k1; ++k)
                        if (static_cast<int64_t>(ci[k]) == row) diag += static_cast<double>(v[k]);
                }
                const double threshold = static_cast<double>(scalar_epsilon<S>()) * scale * 16.0;
                const double reciprocal = 1.0 / diag;
                const bool ok = count > 0 && isfinite(diag) && diag > threshold && isfinite(reciprocal);
                inv[row] = ok ? static_cast<S>(reciprocal) : static_cast<S>(NAN);
                if (!ok) atomicMin(&st[s].bad_row, static_cast<unsigned long long>(row));
            }
        }
    }
}

// ---- block-Jacobi systems -------------------------------------------------------------
//
// A system with block size bs = sys_bs[s] > 1 is preconditioned by the
// inverses of its bs x bs diagonal blocks (blockPcg's M^-1, iterative_solver.h
// :130-141), computed here from A: rows [begin + k bs, begin + (k + 1) bs)
// form block k.  Row i of the system keeps row (i - block start) of its
// block's inverse at binv[i * bmax ...], so the apply reads one short
// contiguous row per lane.

constexpr int kMaxBatchBlock = 8;
constexpr int kRegisterBlock = 6;  // the block size inverted with its block in registers (6x6 pose blocks)

// Rows of the diagonal blocks of the block-Jacobi systems, FP64, into
// tmp[row * bmax + j] (duplicates summed in CSR order): eight lanes per row
// scan its entries (coalesced); a row with a repeated block column is summed
// again serially in CSR order.
template <typename S, typename I>
__global__ __launch_bounds__(kBlock) void k_batch_block_gather(int64_t chunks, const int32_t* __restrict__ chunk_sys,
                                                               const int64_t* __restrict__ sys_end,
                                                               const int64_t* __restrict__ sys_chunk0,
                                                               const int32_t* __restrict__ sys_bs,
                                                               const I* __restrict__ ro, const I* __restrict__ ci,
                                                               const S* __restrict__ v, double* __restrict__ tmp,
                                                               int bmax)
{
    constexpr int L = 8;
    const int lane = threadIdx.x % kWave;
    const int sub = lane % L;
    PM_FOR_CHUNKS(c, chunks, kBatchWaves)
    {
        const int32_t s = chunk_sys[c];
        if (s < 0) continue;
        const int bs = sys_bs[s];
        if (bs <= 1) continue;
        const int64_t end = sys_end[s];
        const int64_t begin = sys_chunk0[s] * kChunkRows;
        for (int pass = 0; pass < L; ++pass)
        {
            const int64_t row = c * kChunkRows + pass * L + lane / L;
            const bool valid = row < end;
            const int64_t b0 = valid ? row - (row - begin) % bs : 0;
            double m[kMaxBatchBlock];
            int cnt[kMaxBatchBlock];
#pragma unroll
            for (int jj = 0; jj < kMaxBatchBlock; ++jj)
            {
                m[jj] = 0.0;
                cnt[jj] = 0;
            }
            int64_t k0 = 0, k1 = 0;
            if (valid)
            {
                k0 = static_cast<int64_t>(ro[row]);
                k1 = static_cast<int64_t>(ro[row + 1]);
                for (int64_t k = k0 + sub; k < k1; k += L)
                {
                    const int64_t d = static_cast<int64_t>(ci[k]) - b0;
                    const double a = static_cast<double>(v[k]);
#pragma unroll
                    for (int jj = 0; jj < kMaxBatchBlock; ++jj)
                        if (d == jj && jj < bs)
                        {
                            m[jj] += a;
                            ++cnt[jj];
                        }
                }
            }
#pragma unroll
            for (int off = L / 2; off > 0; off >>= 1)
            {
#pragma unroll
                for (int jj = 0; jj < kMaxBatchBlock; ++jj)
                {
                    m[jj] += __shfl_xor(m[jj], off, L);
                    cnt[jj] += __shfl_xor(cnt[jj], off, L);
                }
            }
            if (valid && sub == 0)
            {
                bool dup = false;
#pragma unroll
                for (int jj = 0; jj < kMaxBatchBlock; ++jj) dup |= cnt[jj] > 1;
                if (dup)
                {
#pragma unroll
                    for (int jj = 0; jj < kMaxBatchBlock; ++jj) m[jj] = 0.0;
                    for (int64_t k = k0; k < k1; ++k)
                    {
                        const int64_t d = static_cast<int64_t>(ci[k]) - b0;
#pragma unroll
                        for (int jj = 0; jj < kMaxBatchBlock; ++jj)
                            if (d == jj && jj < bs) m[jj] += static_cast<double>(v[k]);
                    }
                }
#pragma unroll
                for (int jj = 0; jj < kMaxBatchBlock; ++jj)
                    if (jj < bs) tmp[row * bmax + jj] = m[jj];
            }
        }
    }
}

// Inverse of an n x n (n <= N) block by Gauss-Jordan elimination without
// pivoting (an SPD block has positive pivots); false when a pivot is not
// finite or not above `threshold`.
template <int N>
__device__ __forceinline__ bool gauss_jordan_inverse(double (&m)[N][N], double (&w)[N][N], int n, double threshold)
{
    for (int k = 0; k < N; ++k)
    {
        if (k >= n) break;
        const double piv = m[k][k];
        if (!(isfinite(piv) && piv > threshold)) return false;
        const double rp = 1.0 / piv;
#pragma unroll
        for (int j = 0; j < N; ++j)
        {
            m[k][j] *= rp;
            w[k][j] *= rp;
        }
#pragma unroll
        for (int i = 0; i < N; ++i)
        {
            if (i == k || i >= n) continue;
            const double f = m[i][k];
#pragma unroll
            for (int j = 0; j < N; ++j)
            {
                m[i][j] = __builtin_fma(-f, m[k][j], m[i][j]);
                w[i][j] = __builtin_fma(-f, w[k][j], w[i][j]);
            }
        }
    }
    return true;
}

// One lane per block (the lane of its first row): the gathered block,
// inverted in FP64, rounded to S into binv.  A pivot that is not finite or not
// above 16 eps(S) times the block's largest magnitude rejects the block: its
// rows get NaN and the system records its first row, as a rejected Jacobi
// diagonal does.  BS > 0 handles the systems of that block size with the
// block in registers; BS = 0 every other size above 1.
template <typename S, int BS>
__global__ __launch_bounds__(kBlock) void k_batch_block_invert(int64_t chunks, const int32_t* __restrict__ chunk_sys,
                                                               const int64_t* __restrict__ sys_end,
                                                               const int64_t* __restrict__ sys_chunk0,
                                                               const int32_t* __restrict__ sys_bs,
                                                               const double* __restrict__ tmp, S* __restrict__ binv,
                                                               int bmax, BatchState* st)
{
    constexpr int N = BS > 0 ? BS : kMaxBatchBlock;
    const int lane = threadIdx.x % kWave;
    PM_FOR_CHUNKS(c, chunks, kBatchWaves)
    {
        const int32_t s = chunk_sys[c];
        if (s < 0) continue;
        const int bs = sys_bs[s];
        if (bs <= 1 || (BS > 0 ? bs != BS : bs == kRegisterBlock)) continue;
        const int64_t row = c * kChunkRows + lane;
        const int64_t begin = sys_chunk0[s] * kChunkRows;
        if (row >= sys_end[s] || (row - begin) % bs != 0) continue;
        const int n = BS > 0 ? BS : bs;
        double m[N][N], w[N][N];
        double scale = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i)
#pragma unroll
            for (int j = 0; j < N; ++j)
            {
                m[i][j] = i < n && j < n ? tmp[(row + i) * bmax + j] : 0.0;
                w[i][j] = i == j ? 1.0 : 0.0;
                scale = fmax(scale, fabs(m[i][j]));
            }
        const double threshold = static_cast<double>(scalar_epsilon<S>()) * scale * 16.0;
        bool ok = gauss_jordan_inverse<N>(m, w, n, threshold);
#pragma unroll
        for (int i = 0; i < N; ++i)
#pragma unroll
            for (int j = 0; j < N; ++j)
            {
                if (i >= n || j >= n) continue;
                const bool fin = ok && isfinite(w[i][j]);
                if (!fin) ok = false;
                binv[(row + i) * bmax + j] = fin ? static_cast<S>(w[i][j]) : static_cast<S>(NAN);
            }
        if (!ok) atomicMin(&st[s].bad_row, static_cast<unsigned long long>(row));
    }
}

// z = M^-1 r on the block systems' chunks: each row takes the dot of its
// inverse row with its block's r (the block may reach into a neighbouring
// chunk, whose r the preceding kernel has finished), and the chunk's partial
// r.z replaces the one the update left.
template <typename S>
__global__ __launch_bounds__(kBlock) void k_batch_block_apply(const int32_t* __restrict__ chunk_sys,
                                                              const int64_t* __restrict__ sys_end,
                                                              const int64_t* __restrict__ sys_chunk0,
                                                              const int32_t* __restrict__ sys_bs,
                                                              const S* __restrict__ r, S* __restrict__ z,
                                                              const S* __restrict__ binv, int bmax,
                                                              const BatchState* st, double* __restrict__ prz,
                                                              const int64_t* __restrict__ list,
                                                              const unsigned long long* __restrict__ list_count)
{
    const int lane = threadIdx.x % kWave;
    PM_FOR_LISTED(li, list_count, kBatchWaves)
    {
        const int64_t c = list[li];
        const int32_t s = chunk_sys[c];
        if (s < 0 || st[s].done != kRunning) continue;
        const int bs = sys_bs[s];
        if (bs <= 1) continue;
        const int64_t row = c * kChunkRows + lane;
        const int64_t begin = sys_chunk0[s] * kChunkRows;
        double rz = 0.0;
        if (row < sys_end[s])
        {
            const int64_t b0 = row - (row - begin) % bs;
            const S* wrow = binv + row * bmax;
            S acc = S(0);
            for (int j = 0; j < bs; ++j) acc = fma_s(wrow[j], r[b0 + j], acc);
            z[row] = acc;
            rz = static_cast<double>(r[row]) * static_cast<double>(acc);
        }
        rz = wave_sum(rz);
        if (lane == 0) prz[c] = rz;
    }
}

constexpr int kReduceInit = 0;
constexpr int kReduceAlpha = 1;
constexpr int kReduceBeta = 2;

// One wavefront per system: fixed-order sum of its chunk partials, then the
// scalar step of the recurrence for that system.
template <typename S, int MODE>
__global__ __launch_bounds__(kBlock) void k_batch_reduce(int64_t systems, const int64_t* __restrict__ sys_chunk0,
                                                         const int64_t* __restrict__ sys_chunk1, BatchState* st,
                                                         const double* __restrict__ pa,
                                                         const double* __restrict__ pb, int32_t max_iterations,
                                                         double rel_tol, double abs_tol,
                                                         unsigned long long* done_count,
                                                         const double* __restrict__ sys_rel)
{
    const int lane = threadIdx.x % kWave;
    const int64_t waves = static_cast<int64_t>(gridDim.x) * (blockDim.x / kWave);
    for (int64_t s = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave; s < systems;
         s += waves)
    {
        if (st[s].done != kRunning) continue;
        double a = 0.0, b = 0.0;
        const int64_t c1 = sys_chunk1[s];
        constexpr int U = 8;
        for (int64_t c = sys_chunk0[s] + lane; c < c1; c += U * kWave)
        {
            // all loads of U strided partials first, then the adds in order
            double va[U], vb[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
                const int64_t cc = c + u * kWave;
                va[u] = cc < c1 ? pa[cc] : 0.0;
                vb[u] = (MODE != kReduceAlpha && cc < c1) ? pb[cc] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
                if (c + u * kWave >= c1) break;
                a += va[u];
                if (MODE != kReduceAlpha) b += vb[u];
            }
        }
        a = wave_sum(a);
        if (MODE != kReduceAlpha) b = wave_sum(b);
        if (lane != 0) continue;
        BatchState w = st[s];
        int code = kRunning;
        if (MODE == kReduceInit)
        {
            const double rr = a, rz = b;
            w.rr = rr;
            w.rr0 = rr;
            w.tol = fmax(abs_tol, (sys_rel ? sys_rel[s] : rel_tol) * sqrt(rr));
            w.rho = rz;
            w.beta = 0.0;
            w.iteration = 0;
            if (!isfinite(rr) || rr < 0.0) code = kInitNonFinite;
            else if (sqrt(rr) <= w.tol) code = kConverged;
            else if (w.bad_row != ~0ull) code = kBadDiagonal;
            else if (max_iterations <= 0) code = kIterationLimit;
        }
        else if (MODE == kReduceAlpha)
        {
            double alpha = 0.0;
            if (!step_length<S>(w.rho, a, alpha)) code = kBreakdown;
            w.alpha = alpha;
        }
        else
        {
            const double rr = a, rz = b;
            w.iteration += 1;
            w.rr = rr;
            if (!isfinite(rr) || rr < 0.0) code = kNonFinite;
            else if (sqrt(rr) <= w.tol) code = kConverged;
            else if (w.iteration >= max_iterations) code = kIterationLimit;
            else
            {
                const double rho_prev = w.rho;
                if (rho_prev == 0.0 && rz == 0.0) w.beta = 0.0;
                else if (isfinite(rho_prev) && isfinite(rz) && rho_prev > 0.0 && rz > 0.0)
                {
                    w.beta = rz / rho_prev;
                    if (!isfinite(w.beta) || !isfinite(static_cast<S>(w.beta))) code = kBreakdown;
                }
                else code = kBreakdown;
                w.rho = rz;
            }
        }
        if (code != kRunning)
        {
            w.done = code;
            atomicAdd(done_count, 1ull);
        }
        st[s] = w;
    }
}

// ---- single-reduction (Chronopoulos & Gear) form of the batch --------------------
//
// Jacobi, CG and block-Jacobi batches run the recurrence of the single-system
// k_cg_step / k_cg_spmv per system, three kernels per iteration (four with
// block systems):
//   step     p = u + beta_s p, s = w + beta_s s (s tracks A p), x += alpha_s p,
//            r -= alpha_s s, u = M^-1 r (block systems: k_batch_block_apply next)
//   SpMV     w = A u (one gather per entry, where the direction kernel gathers
//            z and p_old), per-chunk partials r.r, r.u, w.u, u.s
//   reduce   per system, in a fixed order: stop test, beta = gamma / gamma_prev,
//            den = w.u + 2 beta u.s + beta^2 den_prev (p.Ap of the new
//            direction), alpha = gamma / den; a non-positive den restarts the
//            system's step from u (beta = 0, den = w.u)
// Vectors in place: u in z, w in q, p in p0, s in p1.  A refresh iteration
// forms p, s, x, then r = b - A x in FP64 and u = M^-1 r, then the SpMV.

template <typename S, bool JACOBI, bool STEP_R>
__global__ __launch_bounds__(kBlock) void k_batch_cg_step(const int32_t* __restrict__ chunk_sys,
                                                          const int64_t* __restrict__ sys_end, S* __restrict__ x,
                                                          S* __restrict__ p, S* __restrict__ r, S* __restrict__ u,
                                                          const S* __restrict__ w, S* __restrict__ sv,
                                                          const S* __restrict__ inv, const BatchState* st,
                                                          const int64_t* __restrict__ list,
                                                          const unsigned long long* __restrict__ list_count,
                                                          const int32_t* __restrict__ sys_bs)
{
    const int lane = threadIdx.x % kWave;
    PM_FOR_LISTED(li, list_count, kBatchWaves)
    {
        const int64_t c = list[li];
        const int32_t s = chunk_sys[c];
        if (s < 0 || st[s].done != kRunning) continue;
        const int64_t row = c * kChunkRows + lane;
        if (row >= sys_end[s]) continue;
        const S a = static_cast<S>(st[s].alpha), b = static_cast<S>(st[s].beta);
        const S pn = fma_s(b, p[row], u[row]);
        const S sn = fma_s(b, sv[row], w[row]);
        p[row] = pn;
        sv[row] = sn;
        x[row] = fma_s(a, pn, x[row]);
        if (STEP_R)
        {
            const S rn = fma_s(-a, sn, r[row]);
            r[row] = rn;
            if (!(sys_bs != nullptr && sys_bs[s] > 1)) u[row] = JACOBI ? inv[row] * rn : rn;
        }
    }
}

// s = 0 on the rows of running systems (its first update is s = w).
template <typename S>
__global__ __launch_bounds__(kBlock) void k_batch_cg_clear(const int32_t* __restrict__ chunk_sys,
                                                           const int64_t* __restrict__ sys_end, S* __restrict__ sv,
                                                           const BatchState* st, const int64_t* __restrict__ list,
                                                           const unsigned long long* __restrict__ list_count)
{
    const int lane = threadIdx.x % kWave;
    PM_FOR_LISTED(li, list_count, kBatchWaves)
    {
        const int64_t c = list[li];
        const int32_t s = chunk_sys[c];
        if (s < 0 || st[s].done != kRunning) continue;
        const int64_t row = c * kChunkRows + lane;
        if (row < sys_end[s]) sv[row] = S(0);
    }
}

template <typename S, typename I, int LL = 8>
__global__ __launch_bounds__(kBlock) void k_batch_cg_spmv(const int32_t* __restrict__ chunk_sys,
                                                          const int64_t* __restrict__ sys_end,
                                                          const I* __restrict__ ro, const I* __restrict__ ci,
                                                          const S* __restrict__ v, const S* __restrict__ r,
                                                          const S* __restrict__ u, S* __restrict__ w,
                                                          const S* __restrict__ sv, const BatchState* st,
                                                          double* __restrict__ prr, double* __restrict__ pru,
                                                          double* __restrict__ pwu, double* __restrict__ pus,
                                                          const int64_t* __restrict__ list,
                                                          const unsigned long long* __restrict__ list_count)
{
    __shared__ S lv_all[kBatchWaves][kStreamCap];
    __shared__ I lc_all[kBatchWaves][kStreamCap];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    PM_FOR_LISTED(li, list_count, kBatchWaves)
    {
        const int64_t c = list[li];
        const int32_t s = chunk_sys[c];
        if (s < 0 || st[s].done != kRunning)
        {
            if (lane == 0) prr[c] = pru[c] = pwu[c] = pus[c] = 0.0;
            continue;
        }
        const int64_t end = sys_end[s];
        const int64_t row0 = c * kChunkRows;
        int64_t ks, ke;
        const S acc = chunk_row_sums<S, I, SpmvSource<S>, false, false, LL>(row0, end, ro, ci, lv_all[wave],
                                                                            lc_all[wave], SpmvSource<S>{v, u}, ks, ke);
        const int64_t row = row0 + lane;
        double rr = 0.0, ru = 0.0, wu = 0.0, us = 0.0;
        if (row < end)
        {
            w[row] = acc;
            const double rv = static_cast<double>(r[row]);
            const double uv = static_cast<double>(u[row]);
            rr = rv * rv;
            ru = rv * uv;
            wu = static_cast<double>(acc) * uv;
            us = uv * static_cast<double>(sv[row]);
        }
        rr = wave_sum(rr);
        ru = wave_sum(ru);
        wu = wave_sum(wu);
        us = wave_sum(us);
        if (lane == 0)
        {
            prr[c] = rr;
            pru[c] = ru;
            pwu[c] = wu;
            pus[c] = us;
        }
    }
}

// One wavefront per system: fixed-order sums of its chunk partials, then the
// system's decision.  INIT: the decision at iteration 0 (tolerance from
// ||r0||, the first step from p = u, den = u.Au).
template <typename S, bool INIT>
__global__ __launch_bounds__(kBlock) void k_batch_cg_reduce(int64_t systems, const int64_t* __restrict__ sys_chunk0,
                                                            const int64_t* __restrict__ sys_chunk1, BatchState* st,
                                                            const double* __restrict__ prr,
                                                            const double* __restrict__ pru,
                                                            const double* __restrict__ pwu,
                                                            const double* __restrict__ pus, int32_t max_iterations,
                                                            double rel_tol, double abs_tol,
                                                            unsigned long long* done_count,
                                                            const double* __restrict__ sys_rel)
{
    const int lane = threadIdx.x % kWave;
    const int64_t waves = static_cast<int64_t>(gridDim.x) * (blockDim.x / kWave);
    for (int64_t s = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave; s < systems;
         s += waves)
    {
        if (st[s].done != kRunning) continue;
        double sum[4] = {0.0, 0.0, 0.0, 0.0};
        const double* const src[4] = {prr, pru, pwu, pus};
        const int64_t c1 = sys_chunk1[s];
        constexpr int U = 4;
        for (int64_t c = sys_chunk0[s] + lane; c < c1; c += U * kWave)
        {
            // all loads of U strided partials first, then the adds in order
            double va[4][U];
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int j = 0; j < U; ++j)
                {
                    const int64_t cc = c + j * kWave;
                    va[k][j] = cc < c1 ? src[k][cc] : 0.0;
                }
#pragma unroll
            for (int j = 0; j < U; ++j)
            {
                if (c + j * kWave >= c1) break;
#pragma unroll
                for (int k = 0; k < 4; ++k) sum[k] += va[k][j];
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) sum[k] = wave_sum(sum[k]);
        if (lane != 0) continue;
        const double rr = sum[0], gamma = sum[1], delta = sum[2], us = sum[3];
        BatchState w = st[s];
        int code = kRunning;
        double beta = 0.0, den = delta, alpha = 0.0;
        if (INIT)
        {
            w.rr = rr;
            w.rr0 = rr;
            w.tol = fmax(abs_tol, (sys_rel ? sys_rel[s] : rel_tol) * sqrt(rr));
            w.iteration = 0;
            if (!isfinite(rr) || rr < 0.0) code = kInitNonFinite;
            else if (sqrt(rr) <= w.tol) code = kConverged;
            else if (w.bad_row != ~0ull) code = kBadDiagonal;
            else if (max_iterations <= 0) code = kIterationLimit;
        }
        else
        {
            w.iteration += 1;
            w.rr = rr;
            if (!isfinite(rr) || rr < 0.0) code = kNonFinite;
            else if (sqrt(rr) <= w.tol) code = kConverged;
            else if (w.iteration >= max_iterations) code = kIterationLimit;
            else
            {
                const double rho_prev = w.rho;
                if (rho_prev == 0.0 && gamma == 0.0) beta = 0.0;
                else if (isfinite(rho_prev) && isfinite(gamma) && rho_prev > 0.0 && gamma > 0.0)
                {
                    beta = gamma / rho_prev;
                    if (!isfinite(beta) || !isfinite(static_cast<S>(beta))) code = kBreakdown;
                }
                else code = kBreakdown;
                if (code == kRunning && beta != 0.0) den = delta + beta * (2.0 * us + beta * w.den);
                if (code == kRunning && !(den > 0.0 && isfinite(den)))
                {
                    beta = 0.0;
                    den = delta;
                }
            }
        }
        if (code == kRunning && !step_length<S>(gamma, den, alpha)) code = kBreakdown;
        w.beta = beta;
        w.alpha = alpha;
        w.rho = gamma;
        w.den = den;
        if (code != kRunning)
        {
            w.done = code;
            atomicAdd(done_count, 1ull);
        }
        st[s] = w;
    }
}

// Stopped-system count into the host's fine-grained status slot.
__global__ void k_batch_publish(const unsigned long long* __restrict__ done_count, unsigned long long* host)
{
    if (threadIdx.x == 0) *host = *done_count;
}

// ---- host side -----------------------------------------------------------------------

struct BatchProblem
{
    const pm_csr* a;
    const void* rhs;
    void* x;
    bool jacobi;
    // factored preconditioner (G, G^T block-diagonal like a) with optional
    // per-system Jacobi fallback flags (device), else null
    const pm_csr* g;
    const pm_csr* gt;
    const uint8_t* sys_jacobi;
    int64_t systems;
    int64_t chunks;
    int32_t max_iterations;
    double rel_tol, abs_tol;
    int grid_chunks, grid_systems;
    void *r, *z, *q, *p0, *p1, *inv, *t;
    int32_t* chunk_sys;
    int64_t* sys_end;
    int64_t* sys_chunk0;
    int64_t* sys_chunk1;
    BatchState* st;
    double *pden, *prr, *prz, *pus;
    unsigned long long* done_count;
    int64_t* list;                    // chunks of running systems
    unsigned long long* list_count;
    bool split;                       // direction update as its own kernel (SpMV gathers p once)
    bool single_reduction;            // Jacobi / CG / block-Jacobi batches as the single-reduction form
    const double* sys_rel;            // per-system relative tolerance (device; null: rel_tol for all)
    int long_lanes;                   // lanes per row of chunks with long rows (8, 16, 32 or 64)
    bool fuse_step;                   // factored: the step folded into the factor products
    // block-Jacobi: per-system block size (device; null: none), the largest
    // one and the inverse rows (bmax values per row)
    const int32_t* sys_bs;
    int bmax;
    void* binv;
    double* block_rows;               // the gathered diagonal-block rows (FP64, bmax per row)
    // the Jacobi and block inverses of this matrix are in place from an
    // earlier solve (the refinement loop's later inner solves)
    bool inverses_ready;
};

template <typename S, typename I>
struct BatchCalls
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

    template <int MODE>
    static KernelCall reduce(const BatchProblem& b, const double* pa, const double* pb)
    {
        return make_call(k_batch_reduce<S, MODE>, dim3(b.grid_systems), dim3(kBlock), b.systems, b.sys_chunk0,
                         b.sys_chunk1, b.st, pa, pb, b.max_iterations, b.rel_tol, b.abs_tol, b.done_count, b.sys_rel);
    }

    template <bool INIT>
    static KernelCall residual(const BatchProblem& b, void* p)
    {
        const auto k = b.jacobi ? k_batch_residual<S, I, true, INIT> : k_batch_residual<S, I, false, INIT>;
        return make_call(k, dim3(b.grid_chunks), dim3(kBlock), b.chunks, b.chunk_sys, b.sys_end, ro(b.a), ci(b.a),
                         static_cast<const S*>(b.a->values), static_cast<const S*>(b.rhs),
                         static_cast<const S*>(b.x), static_cast<S*>(b.r), static_cast<S*>(b.inv),
                         static_cast<S*>(b.z), static_cast<S*>(p), b.st, b.prr, b.prz,
                         static_cast<const int64_t*>(b.list), static_cast<const unsigned long long*>(b.list_count));
    }

    // z = M^-1 r for the factored preconditioner: t = G r_new, z = G^T t (or
    // Jacobi for flagged systems), partials r.r and r.z.
    template <template <int> class K>
    static auto by_lanes(int lanes)
    {
        switch (lanes)
        {
        case 16: return K<16>::f;
        case 32: return K<32>::f;
        case 64: return K<64>::f;
        default: return K<8>::f;
        }
    }
    template <bool SPLIT>
    struct Dir
    {
        template <int LL>
        struct K
        {
            static constexpr auto f = k_batch_direction<S, I, SPLIT, LL>;
        };
    };
    template <int LL>
    struct CgSpmv
    {
        static constexpr auto f = k_batch_cg_spmv<S, I, LL>;
    };
    template <bool STEP>
    struct Lower
    {
        template <int LL>
        struct K
        {
            static constexpr auto f = k_batch_factor_lower<S, I, STEP, LL>;
        };
    };
    template <bool STEP>
    struct Upper
    {
        template <int LL>
        struct K
        {
            static constexpr auto f = k_batch_factor_upper<S, I, STEP, LL>;
        };
    };

    template <bool STEP>
    static void factored(const BatchProblem& b, const void* p, std::vector<KernelCall>& out)
    {
        out.push_back(make_call(by_lanes<Lower<STEP>::template K>(b.long_lanes), dim3(b.grid_chunks), dim3(kBlock), b.chunk_sys,
                                b.sys_end, ro(b.g), ci(b.g), static_cast<const S*>(b.g->values),
                                static_cast<S*>(b.x), static_cast<const S*>(b.r), static_cast<const S*>(p),
                                static_cast<const S*>(b.q), static_cast<S*>(b.t),
                                static_cast<const BatchState*>(b.st), b.sys_jacobi,
                                static_cast<const int64_t*>(b.list),
                                static_cast<const unsigned long long*>(b.list_count)));
        out.push_back(make_call(by_lanes<Upper<STEP>::template K>(b.long_lanes), dim3(b.grid_chunks), dim3(kBlock), b.chunk_sys,
                                b.sys_end, ro(b.gt), ci(b.gt), static_cast<const S*>(b.gt->values),
                                static_cast<const S*>(b.t), static_cast<S*>(b.r), static_cast<const S*>(b.q),
                                static_cast<const S*>(b.inv), static_cast<S*>(b.z),
                                static_cast<const BatchState*>(b.st), b.sys_jacobi, b.prr, b.prz,
                                static_cast<const int64_t*>(b.list),
                                static_cast<const unsigned long long*>(b.list_count)));
    }

    static bool single(const BatchProblem& b) { return b.single_reduction && !b.g; }

    // w = A u and the partials r.r, r.u, w.u, u.s (single-reduction form)
    static KernelCall cg_spmv(const BatchProblem& b, int lanes)
    {
        return make_call(by_lanes<CgSpmv>(lanes), dim3(b.grid_chunks), dim3(kBlock), b.chunk_sys, b.sys_end, ro(b.a),
                         ci(b.a), static_cast<const S*>(b.a->values), static_cast<const S*>(b.r),
                         static_cast<const S*>(b.z), static_cast<S*>(b.q), static_cast<const S*>(b.p1),
                         static_cast<const BatchState*>(b.st), b.prr, b.prz, b.pden, b.pus,
                         static_cast<const int64_t*>(b.list), static_cast<const unsigned long long*>(b.list_count));
    }

    template <bool INIT>
    static KernelCall cg_reduce(const BatchProblem& b)
    {
        return make_call(k_batch_cg_reduce<S, INIT>, dim3(b.grid_systems), dim3(kBlock), b.systems, b.sys_chunk0,
                         b.sys_chunk1, b.st, static_cast<const double*>(b.prr), static_cast<const double*>(b.prz),
                         static_cast<const double*>(b.pden), static_cast<const double*>(b.pus), b.max_iterations,
                         b.rel_tol, b.abs_tol, b.done_count, b.sys_rel);
    }

    template <bool STEP_R>
    static KernelCall cg_step(const BatchProblem& b)
    {
        const auto k = b.jacobi ? k_batch_cg_step<S, true, STEP_R> : k_batch_cg_step<S, false, STEP_R>;
        return make_call(k, dim3(b.grid_chunks), dim3(kBlock), b.chunk_sys, b.sys_end, static_cast<S*>(b.x),
                         static_cast<S*>(b.p0), static_cast<S*>(b.r), static_cast<S*>(b.z),
                         static_cast<const S*>(b.q), static_cast<S*>(b.p1), static_cast<const S*>(b.inv),
                         static_cast<const BatchState*>(b.st), static_cast<const int64_t*>(b.list),
                         static_cast<const unsigned long long*>(b.list_count), b.sys_bs);
    }

    static void init(const BatchProblem& b, std::vector<KernelCall>& out)
    {
        if ((b.jacobi || (b.g && b.sys_jacobi)) && !b.inverses_ready)
            out.push_back(make_call(k_batch_jacobi_inverse<S, I>, dim3(b.grid_chunks), dim3(kBlock), b.chunks,
                                    b.chunk_sys, b.sys_end, ro(b.a), ci(b.a), static_cast<const S*>(b.a->values),
                                    static_cast<S*>(b.inv), b.st, b.g ? b.sys_jacobi : nullptr, b.sys_bs));
        if (b.sys_bs && !b.inverses_ready)
        {
            out.push_back(make_call(k_batch_block_gather<S, I>, dim3(b.grid_chunks), dim3(kBlock), b.chunks,
                                    b.chunk_sys, b.sys_end, b.sys_chunk0, b.sys_bs, ro(b.a), ci(b.a),
                                    static_cast<const S*>(b.a->values), b.block_rows, b.bmax));
            for (const auto k : {k_batch_block_invert<S, kRegisterBlock>, k_batch_block_invert<S, 0>})
                out.push_back(make_call(k, dim3(b.grid_chunks), dim3(kBlock), b.chunks, b.chunk_sys, b.sys_end,
                                        b.sys_chunk0, b.sys_bs, static_cast<const double*>(b.block_rows),
                                        static_cast<S*>(b.binv), b.bmax, b.st));
        }
        list(b, out);
        out.push_back(residual<true>(b, b.p0));
        block_apply(b, out);
        if (single(b))
        {
            out.push_back(make_call(k_batch_cg_clear<S>, dim3(b.grid_chunks), dim3(kBlock), b.chunk_sys, b.sys_end,
                                    static_cast<S*>(b.p1), static_cast<const BatchState*>(b.st),
                                    static_cast<const int64_t*>(b.list),
                                    static_cast<const unsigned long long*>(b.list_count)));
            out.push_back(cg_spmv(b, b.long_lanes));
            out.push_back(cg_reduce<true>(b));
            return;
        }
        if (b.g) factored<false>(b, nullptr, out);
        out.push_back(reduce<kReduceInit>(b, b.prr, b.prz));
    }

    // z = M^-1 r and the partial r.z of the block-Jacobi systems.
    static void block_apply(const BatchProblem& b, std::vector<KernelCall>& out)
    {
        if (!b.sys_bs) return;
        out.push_back(make_call(k_batch_block_apply<S>, dim3(b.grid_chunks), dim3(kBlock), b.chunk_sys, b.sys_end,
                                b.sys_chunk0, b.sys_bs, static_cast<const S*>(b.r), static_cast<S*>(b.z),
                                static_cast<const S*>(b.binv), b.bmax, static_cast<const BatchState*>(b.st), b.prz,
                                static_cast<const int64_t*>(b.list),
                                static_cast<const unsigned long long*>(b.list_count)));
    }

    static void list(const BatchProblem& b, std::vector<KernelCall>& out)
    {
        out.push_back(make_call(k_batch_list_reset, dim3(1), dim3(1), b.list_count));
        const int grid = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(b.chunks, kBlock), 4096)));
        out.push_back(make_call(k_batch_list, dim3(grid), dim3(kBlock), b.chunks, b.chunk_sys,
                                static_cast<const BatchState*>(b.st), b.list, b.list_count));
    }

    static void iteration(const BatchProblem& b, int64_t t, bool refresh, std::vector<KernelCall>& out)
    {
        if (single(b))
        {
            if (refresh)
            {
                out.push_back(cg_step<false>(b));
                out.push_back(residual<false>(b, nullptr));
            }
            else out.push_back(cg_step<true>(b));
            block_apply(b, out);
            out.push_back(cg_spmv(b, b.long_lanes));
            out.push_back(cg_reduce<false>(b));
            return;
        }
        const bool even = t % 2 == 0;
        const void* p_old = even ? b.p0 : b.p1;
        void* p_new = even ? b.p1 : b.p0;
        if (b.split)
            out.push_back(make_call(k_batch_pupdate<S>, dim3(b.grid_chunks), dim3(kBlock), b.chunk_sys, b.sys_end,
                                    static_cast<const S*>(b.z), static_cast<const S*>(p_old), static_cast<S*>(p_new),
                                    static_cast<const BatchState*>(b.st), static_cast<const int64_t*>(b.list),
                                    static_cast<const unsigned long long*>(b.list_count)));
        out.push_back(make_call(b.split ? by_lanes<Dir<true>::template K>(b.long_lanes)
                                        : by_lanes<Dir<false>::template K>(b.long_lanes),
                                dim3(b.grid_chunks), dim3(kBlock), b.chunks, b.chunk_sys,
                                b.sys_end, ro(b.a), ci(b.a), static_cast<const S*>(b.a->values),
                                static_cast<const S*>(b.z), static_cast<const S*>(p_old), static_cast<S*>(p_new),
                                static_cast<S*>(b.q), static_cast<const BatchState*>(b.st), b.pden,
                                static_cast<const int64_t*>(b.list),
                                static_cast<const unsigned long long*>(b.list_count)));
        out.push_back(reduce<kReduceAlpha>(b, b.pden, nullptr));
        if (b.g && !refresh && b.fuse_step)
        {
            // the step rides in the two factor products
            factored<true>(b, p_new, out);
        }
        else if (b.g && !refresh)
        {
            out.push_back(make_call(k_batch_step_xr<S>, dim3(b.grid_chunks), dim3(kBlock), b.chunk_sys, b.sys_end,
                                    static_cast<S*>(b.x), static_cast<S*>(b.r), static_cast<const S*>(p_new),
                                    static_cast<const S*>(b.q), static_cast<const BatchState*>(b.st),
                                    static_cast<const int64_t*>(b.list),
                                    static_cast<const unsigned long long*>(b.list_count)));
            factored<false>(b, nullptr, out);
        }
        else if (!refresh)
        {
            const auto k = b.jacobi ? k_batch_update<S, true, true> : k_batch_update<S, false, true>;
            out.push_back(make_call(k, dim3(b.grid_chunks), dim3(kBlock), b.chunks, b.chunk_sys, b.sys_end,
                                    static_cast<S*>(b.x), static_cast<S*>(b.r), static_cast<S*>(b.z),
                                    static_cast<const S*>(p_new), static_cast<const S*>(b.q),
                                    static_cast<const S*>(b.inv), static_cast<const BatchState*>(b.st), b.prr,
                                    b.prz, static_cast<const int64_t*>(b.list),
                                    static_cast<const unsigned long long*>(b.list_count), b.sys_bs));
            block_apply(b, out);
        }
        else
        {
            out.push_back(make_call(k_batch_update<S, false, false>, dim3(b.grid_chunks), dim3(kBlock), b.chunks,
                                    b.chunk_sys, b.sys_end, static_cast<S*>(b.x), static_cast<S*>(b.r),
                                    static_cast<S*>(b.z), static_cast<const S*>(p_new), static_cast<const S*>(b.q),
                                    static_cast<const S*>(b.inv), static_cast<const BatchState*>(b.st), b.prr,
                                    b.prz, static_cast<const int64_t*>(b.list),
                                    static_cast<const unsigned long long*>(b.list_count), b.sys_bs));
            out.push_back(residual<false>(b, nullptr));
            if (b.g) factored<false>(b, nullptr, out);
            block_apply(b, out);
        }
        out.push_back(reduce<kReduceBeta>(b, b.prr, b.prz));
    }

    // Device time (ms) of `reps` direction launches of the first iteration
    // with `lanes` lanes per long row (idempotent: the first iteration writes
    // the same p, q and partials again).
    static float time_direction(const BatchProblem& b, int lanes, int reps, hipStream_t stream, hipEvent_t e0,
                                hipEvent_t e1)
    {
        const auto k = by_lanes<Dir<false>::template K>(lanes);
        const KernelCall spmv = cg_spmv(b, lanes);
        // the single-reduction form times its SpMV, which rewrites the same w
        // and partials from the same r, u, s
        auto launch = [&]() {
            if (single(b))
            {
                (void)pm::launch(spmv, stream);
                return;
            }
            hipLaunchKernelGGL(k, dim3(b.grid_chunks), dim3(kBlock), 0, stream, b.chunks, b.chunk_sys, b.sys_end,
                               ro(b.a), ci(b.a), static_cast<const S*>(b.a->values), static_cast<const S*>(b.z),
                               static_cast<const S*>(b.p0), static_cast<S*>(b.p1), static_cast<S*>(b.q),
                               static_cast<const BatchState*>(b.st), b.pden, static_cast<const int64_t*>(b.list),
                               static_cast<const unsigned long long*>(b.list_count));
        };
        launch();
        (void)hipEventRecord(e0, stream);
        for (int i = 0; i < reps; ++i) launch();
        (void)hipEventRecord(e1, stream);
        if (hipEventSynchronize(e1) != hipSuccess) return -1.f;
        float ms = -1.f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return -1.f;
        return ms;
    }

    static void build(const BatchProblem& b, int what, int64_t first, int64_t count, int refresh_interval,
                      std::vector<KernelCall>& out)
    {
        if (what == 0)
        {
            init(b, out);
            return;
        }
        if (what == 2)
        {
            list(b, out);
            return;
        }
        for (int64_t t = first; t < first + count; ++t)
            iteration(b, t, refresh_interval > 0 && ((t + 1) % refresh_interval) == 0, out);
    }
};

struct BatchPolicy
{
    int chunk = 16;
    bool graph = true;
    int refresh = 0;
};

BatchPolicy resolve_batch_policy(const pm_solve_policy* p)
{
    BatchPolicy out;
    if (p)
    {
        if (p->chunk_iterations > 0) out.chunk = p->chunk_iterations;
        if (p->use_graph >= 0) out.graph = p->use_graph != 0;
        out.refresh = p->refresh_interval > 0 ? p->refresh_interval : 0;
    }
    if (out.chunk % 2) out.chunk += 1;
    if (out.refresh > 0)
    {
        if (out.refresh % 2) out.refresh += 1;
        int k = std::min(out.chunk, out.refresh);
        while (k > 2 && (out.refresh % k != 0 || k % 2 != 0)) --k;
        out.chunk = std::max(k, 2);
    }
    return out;
}

std::string batch_key(const BatchProblem& b, int what, int64_t first, int64_t count, int refresh)
{
    char buf[512];
    std::snprintf(buf, sizeof(buf),
                  "batch|%d|%p|%p|%p|%p|%p|%p|%p|%p|%p|%p|%lld|%lld|%d|%lld|%lld|%d|%d|%d|%.17g|%.17g|%d|%p", what, b.r,
                  static_cast<void*>(b.st), static_cast<void*>(b.pden), b.a->values, (const void*)b.a->row_offsets32,
                  (const void*)b.a->row_offsets, (const void*)b.a->col_indices32, (const void*)b.a->col_indices,
                  b.rhs, b.x,
                  static_cast<long long>(b.a->rows), static_cast<long long>(b.systems), b.jacobi ? 1 : 0,
                  static_cast<long long>(first), static_cast<long long>(count), refresh, b.a->dtype,
                  b.max_iterations, b.rel_tol, b.abs_tol,
                  (b.split ? 1 : 0) | (b.fuse_step ? 2 : 0) | (b.single_reduction ? 4 : 0) | (b.long_lanes << 3),
                  static_cast<const void*>(b.sys_jacobi));
    std::string key(buf);
    std::snprintf(buf, sizeof(buf), "|blocks|%p|%d|%p|%p|%d|%p", static_cast<const void*>(b.sys_bs), b.bmax, b.binv,
                  static_cast<void*>(b.block_rows), b.inverses_ready ? 1 : 0, static_cast<const void*>(b.sys_rel));
    key += buf;
    if (b.g)
    {
        // every array of G and G^T the factored kernels capture
        for (const pm_csr* m : {b.g, b.gt})
        {
            std::snprintf(buf, sizeof(buf), "|%p|%p|%p|%p|%p|%lld|%d", m->values, (const void*)m->row_offsets32,
                          (const void*)m->col_indices32, (const void*)m->row_offsets, (const void*)m->col_indices,
                          static_cast<long long>(m->nnz), m->dtype);
            key += buf;
        }
    }
    return key;
}

template <typename S>
Plan* batch_plan(pm_workspace* ws, const BatchProblem& b, int what, int64_t first, int64_t count, int refresh)
{
    const std::string key = batch_key(b, what, first, count, refresh);
    Plan* plan = ws->plans->find(key);
    if (!plan)
    {
        auto fresh = std::make_unique<Plan>();
        if (use_compact(b.a)) BatchCalls<S, int32_t>::build(b, what, first, count, refresh, fresh->calls);
        else BatchCalls<S, int64_t>::build(b, what, first, count, refresh, fresh->calls);
        plan = ws->plans->insert(key, std::move(fresh));
    }
    return plan;
}

// Rows-per-wavefront autotuner of the batched iteration: the lanes given to
// each row of a chunk too long for the LDS windows (8, 16, 32 or 64 lanes per
// row: 8 to 1 rows per wavefront step).  Timed once per matrix (arrays, size,
// systems) and process on the batch's own direction kernel, then kept.
std::mutex g_lanes_mutex;
std::map<std::string, int> g_lanes;

std::string lanes_key(const BatchProblem& b)
{
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%p|%p|%p|%lld|%lld|%lld|%d", b.a->values, (const void*)b.a->col_indices32,
                  (const void*)b.a->col_indices, static_cast<long long>(b.a->rows),
                  static_cast<long long>(b.a->nnz), static_cast<long long>(b.systems), b.a->dtype);
    return buf;
}

template <typename S>
int tune_long_lanes(BatchProblem& b, hipStream_t stream)
{
    // chunks of fewer than kMaxWindows * kStreamCap entries never take the
    // long-row path: rows averaging under 24 entries keep the default
    if (b.a->rows == 0 || b.a->nnz < 24 * b.a->rows)
    {
        b.long_lanes = 8;
        return PM_OK;
    }
    const std::string key = lanes_key(b);
    {
        std::lock_guard<std::mutex> lock(g_lanes_mutex);
        auto it = g_lanes.find(key);
        if (it != g_lanes.end())
        {
            b.long_lanes = it->second;
            return PM_OK;
        }
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    PM_HIP(hipEventCreate(&e0));
    PM_HIP(hipEventCreate(&e1));
    int best = 8;
    float best_ms = -1.f;
    for (int lanes : {8, 16, 32, 64})
    {
        const float ms = use_compact(b.a) ? BatchCalls<S, int32_t>::time_direction(b, lanes, 3, stream, e0, e1)
                                          : BatchCalls<S, int64_t>::time_direction(b, lanes, 3, stream, e0, e1);
        if (ms >= 0.f && (best_ms < 0.f || ms < best_ms))
        {
            best_ms = ms;
            best = lanes;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    PM_HIP(hipGetLastError());
    b.long_lanes = best;
    std::lock_guard<std::mutex> lock(g_lanes_mutex);
    g_lanes[key] = best;
    return PM_OK;
}

template <typename S>
int batch_solve(BatchProblem& b, const BatchPolicy& policy, pm_batch_report* reports, pm_workspace* ws,
                hipStream_t stream)
{
    hipLaunchKernelGGL(k_batch_state_init, dim3(static_cast<int>(std::min<int64_t>(ceil_div(b.systems, kBlock), 1024))),
                       dim3(kBlock), 0, stream, b.systems, b.st, b.done_count, b.inverses_ready ? 1 : 0);
    PM_HIP(hipGetLastError());
    PM_TRY(batch_plan<S>(ws, b, 0, 0, 0, 0)->run(stream, false));
    if (b.long_lanes == 0) PM_TRY(tune_long_lanes<S>(b, stream));
    const int64_t K = policy.chunk;
    const int64_t needed = b.max_iterations;
    unsigned long long* hs = ws->host_batch;
    std::deque<int64_t> inflight;
    int64_t launched = 0, chunk_id = 0;
    auto issue = [&]() -> int {
        const int64_t phase = policy.refresh > 0 ? launched % policy.refresh : 0;
        PM_TRY(batch_plan<S>(ws, b, 1, phase, K, policy.refresh)->run(stream, policy.graph));
        const int slot = static_cast<int>(chunk_id % 2);
        // the count goes straight into fine-grained host memory (a one-wave
        // kernel instead of a copy command per chunk)
        hipLaunchKernelGGL(k_batch_publish, dim3(1), dim3(kWave), 0, stream,
                           static_cast<const unsigned long long*>(b.done_count), hs + slot);
        PM_HIP(hipGetLastError());
        PM_HIP(hipEventRecord(ws->events[slot], stream));
        inflight.push_back(chunk_id);
        ++chunk_id;
        launched += K;
        return PM_OK;
    };
    // the initialisation alone may stop every system
    PM_HIP(hipMemcpyAsync(hs, b.done_count, sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
    PM_HIP(hipStreamSynchronize(stream));
    bool finished = static_cast<int64_t>(hs[0]) >= b.systems;
    unsigned long long listed_at = hs[0];  // stopped systems when the list was last built
    if (!finished && needed > 0)
    {
        PM_TRY(issue());
        while (true)
        {
            if (launched < needed && inflight.size() < 2) PM_TRY(issue());
            const int64_t oldest = inflight.front();
            inflight.pop_front();
            const int slot = static_cast<int>(oldest % 2);
            PM_HIP(hipEventSynchronize(ws->events[slot]));
            if (static_cast<int64_t>(hs[slot]) >= b.systems) break;
            if (hs[slot] > listed_at)
            {
                // drop the stopped systems' chunks from the chunks still to be issued
                listed_at = hs[slot];
                PM_TRY(batch_plan<S>(ws, b, 2, 0, 0, 0)->run(stream, false));
            }
            if (inflight.empty() && launched >= needed)
                return fail(PM_RUNTIME_ERROR, "batched solver stopped without a final state");
        }
    }
    std::vector<BatchState> host(static_cast<size_t>(b.systems));
    PM_HIP(hipMemcpyAsync(host.data(), b.st, sizeof(BatchState) * b.systems, hipMemcpyDeviceToHost, stream));
    PM_HIP(hipStreamSynchronize(stream));
    PM_HIP(hipEventRecord(ws->done_event, stream));
    for (int64_t s = 0; s < b.systems; ++s)
    {
        pm_batch_report& r = reports[s];
        r.status = host[s].done;
        r.converged = host[s].done == kConverged ? 1 : 0;
        r.iterations = host[s].iteration;
        r.initial_residual = std::sqrt(host[s].rr0);
        r.final_residual = std::sqrt(host[s].rr);
        r.bad_row = host[s].bad_row == ~0ull ? -1 : static_cast<int64_t>(host[s].bad_row);
    }
    return PM_OK;
}

// Validation, host layout (per-chunk system ids, per-system row end and
// chunk range, copied to the device) and workspace carving shared by the
// batched solve and its refinement loop.
int batch_setup(const pm_csr* a, int64_t systems, const int64_t* row_begin, const int64_t* row_end, const void* rhs,
                void* x, const pm_solver_options* options, bool have_reports, pm_workspace* ws, hipStream_t stream,
                BatchProblem& b)
{
    PM_TRY(check_csr(a, "batched pcg"));
    if (!options || !ws || !have_reports || systems < 0 || (systems > 0 && (!row_begin || !row_end)))
        return fail(PM_INVALID_ARGUMENT, "batched pcg: null argument");
    if (a->rows != a->cols) return fail(PM_INVALID_ARGUMENT, "batched pcg: matrix must be square");
    if (options->max_iterations < 0 || !std::isfinite(options->relative_tolerance) ||
        options->relative_tolerance < 0.0 || options->relative_tolerance > 1.0 ||
        !std::isfinite(options->absolute_tolerance) || options->absolute_tolerance < 0.0)
        return fail(PM_INVALID_ARGUMENT, "iterative solver options are invalid");
    if (a->rows > 0 && (!rhs || !x)) return fail(PM_INVALID_ARGUMENT, "batched pcg: null vectors");
    if (rhs && rhs == x) return fail(PM_INVALID_ARGUMENT, "batched pcg: rhs and solution must not alias");
    if (systems >= (int64_t(1) << 31)) return fail(PM_OVERFLOW, "batched pcg: too many systems");
    const int64_t n = a->rows;
    const int64_t chunks = chunk_count(n);
    // host layout -> per-chunk system ids and per-system chunk ranges
    std::vector<int32_t> chunk_sys(static_cast<size_t>(chunks), -1);
    std::vector<int64_t> sys(static_cast<size_t>(3 * std::max<int64_t>(systems, 1)));
    int64_t prev_end = 0;
    for (int64_t s = 0; s < systems; ++s)
    {
        const int64_t b0 = row_begin[s], b1 = row_end[s];
        if (b0 % kChunkRows != 0)
            return fail(PM_INVALID_ARGUMENT, "batched pcg: system %lld starts at row %lld, not a multiple of %d",
                        static_cast<long long>(s), static_cast<long long>(b0), kChunkRows);
        if (b0 < prev_end || b1 < b0 || b1 > n)
            return fail(PM_INVALID_ARGUMENT, "batched pcg: system %lld row range [%lld, %lld) is out of order",
                        static_cast<long long>(s), static_cast<long long>(b0), static_cast<long long>(b1));
        prev_end = ceil_div(b1, kChunkRows) * kChunkRows;
        const int64_t c0 = b0 / kChunkRows, c1 = ceil_div(b1, kChunkRows);
        for (int64_t c = c0; c < c1; ++c) chunk_sys[static_cast<size_t>(c)] = static_cast<int32_t>(s);
        sys[static_cast<size_t>(s)] = b1;
        sys[static_cast<size_t>(systems + s)] = c0;
        sys[static_cast<size_t>(2 * systems + s)] = c1;
    }
    PM_TRY(ws->bind(stream));
    PM_TRY(ws->ensure_events());
    const size_t es = a->dtype == PM_F64 ? 8 : 4;
    const size_t vec = ((es * static_cast<size_t>(n > 0 ? n : 1)) + 255) & ~size_t(255);
    PM_TRY(ws->vectors.reserve(vec * 7));
    if (n > ws->capacity_rows) ws->capacity_rows = n;
    auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
    const size_t s1 = std::max<int64_t>(systems, 1);
    const size_t c1 = std::max<int64_t>(chunks, 1);
    const size_t bytes = al(sizeof(int32_t) * c1) + 3 * al(sizeof(int64_t) * s1) + al(sizeof(BatchState) * s1) +
                         4 * al(sizeof(double) * c1) + 256 + al(sizeof(int64_t) * c1) + 256;
    PM_TRY(ws->batch.reserve(bytes));
    if (!ws->host_batch)
        PM_HIP(hipHostMalloc(reinterpret_cast<void**>(&ws->host_batch), 2 * sizeof(unsigned long long),
                             hipHostMallocCoherent));
    b = BatchProblem{};
    b.a = a;
    b.rhs = rhs;
    b.x = x;
    b.jacobi = options->use_jacobi != 0;
    b.systems = systems;
    b.chunks = chunks;
    b.max_iterations = options->max_iterations;
    b.rel_tol = options->relative_tolerance;
    b.abs_tol = options->absolute_tolerance;
    // the fused direction measured faster on the 512-system batch (537 vs
    // 640 ms); PLAMATRIX_AMD_BATCH_SPLIT=1 selects the split form
    const char* split = std::getenv("PLAMATRIX_AMD_BATCH_SPLIT");
    b.split = split != nullptr && split[0] == '1';
    // Jacobi / CG / block-Jacobi batches as the single-reduction form
    // (PLAMATRIX_AMD_BATCH_SINGLE_REDUCTION=0 keeps direction / update)
    const char* single = std::getenv("PLAMATRIX_AMD_BATCH_SINGLE_REDUCTION");
    b.single_reduction = single == nullptr || single[0] != '0';
    // lanes per long row: PLAMATRIX_AMD_BATCH_LANES (8, 16, 32, 64), else tuned
    // on the first solve of this matrix
    const char* lanes = std::getenv("PLAMATRIX_AMD_BATCH_LANES");
    const int lv = lanes ? std::atoi(lanes) : 0;
    b.long_lanes = (lv == 8 || lv == 16 || lv == 32 || lv == 64) ? lv : 0;
    b.grid_chunks = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(chunks, kBatchWaves), kBatchGrid)));
    b.grid_systems = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(systems, kBatchWaves), 1024)));
    char* vb = ws->vectors.as<char>();
    b.r = vb;
    b.z = vb + vec;
    b.q = vb + 2 * vec;
    b.p0 = vb + 3 * vec;
    b.p1 = vb + 4 * vec;
    b.inv = vb + 5 * vec;
    b.t = vb + 6 * vec;
    char* p = ws->batch.as<char>();
    b.chunk_sys = reinterpret_cast<int32_t*>(p);
    p += al(sizeof(int32_t) * c1);
    b.sys_end = reinterpret_cast<int64_t*>(p);
    b.sys_chunk0 = b.sys_end + systems;
    b.sys_chunk1 = b.sys_end + 2 * systems;
    p += 3 * al(sizeof(int64_t) * s1);
    b.st = reinterpret_cast<BatchState*>(p);
    p += al(sizeof(BatchState) * s1);
    b.pden = reinterpret_cast<double*>(p);
    p += al(sizeof(double) * c1);
    b.prr = reinterpret_cast<double*>(p);
    p += al(sizeof(double) * c1);
    b.prz = reinterpret_cast<double*>(p);
    p += al(sizeof(double) * c1);
    b.pus = reinterpret_cast<double*>(p);
    p += al(sizeof(double) * c1);
    b.done_count = reinterpret_cast<unsigned long long*>(p);
    p += 256;
    b.list = reinterpret_cast<int64_t*>(p);
    p += al(sizeof(int64_t) * c1);
    b.list_count = reinterpret_cast<unsigned long long*>(p);
    // the layout arrays are consumed by the first kernels: copy them
    // synchronously (they live in host vectors that go out of scope)
    PM_HIP(hipStreamSynchronize(stream));
    if (chunks > 0)
        PM_HIP(hipMemcpy(b.chunk_sys, chunk_sys.data(), sizeof(int32_t) * chunks, hipMemcpyHostToDevice));
    if (systems > 0) PM_HIP(hipMemcpy(b.sys_end, sys.data(), sizeof(int64_t) * 3 * systems, hipMemcpyHostToDevice));
    return PM_OK;
}

// Block-Jacobi sizes of a batch (host, one per system, 1 to kMaxBatchBlock;
// 1 is Jacobi): every block-Jacobi system's size must be a multiple of its
// block size.  A batch whose sizes are all 1 is a Jacobi batch.
int batch_attach_blocks(BatchProblem& b, const int64_t* row_begin, const int64_t* row_end,
                        const int32_t* block_sizes, pm_workspace* ws, hipStream_t stream)
{
    if (b.systems > 0 && !block_sizes) return fail(PM_INVALID_ARGUMENT, "batched block pcg: null block sizes");
    int bmax = 1;
    for (int64_t s = 0; s < b.systems; ++s)
    {
        const int bs = block_sizes[s];
        if (bs < 1 || bs > kMaxBatchBlock)
            return fail(PM_INVALID_ARGUMENT, "batched block pcg: system %lld block size %d is outside [1, %d]",
                        static_cast<long long>(s), bs, kMaxBatchBlock);
        if ((row_end[s] - row_begin[s]) % bs != 0)
            return fail(PM_INVALID_ARGUMENT, "batched block pcg: system %lld has %lld rows, not a multiple of %d",
                        static_cast<long long>(s), static_cast<long long>(row_end[s] - row_begin[s]), bs);
        bmax = std::max(bmax, bs);
    }
    b.jacobi = true;
    if (bmax == 1) return PM_OK;
    auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
    const size_t es = b.a->dtype == PM_F64 ? 8 : 4;
    const size_t rows = static_cast<size_t>(std::max<int64_t>(b.chunks, 1)) * kChunkRows;
    const size_t sizes_bytes = al(sizeof(int32_t) * static_cast<size_t>(b.systems));
    const size_t inv_bytes = al(es * rows * static_cast<size_t>(bmax));
    PM_TRY(ws->blocks.reserve(sizes_bytes + inv_bytes + al(sizeof(double) * rows * static_cast<size_t>(bmax))));
    char* base = ws->blocks.as<char>();
    b.sys_bs = reinterpret_cast<const int32_t*>(base);
    b.binv = base + sizes_bytes;
    b.block_rows = reinterpret_cast<double*>(base + sizes_bytes + inv_bytes);
    b.bmax = bmax;
    // the host array is consumed by the first kernels: copy it synchronously
    PM_HIP(hipStreamSynchronize(stream));
    PM_HIP(hipMemcpy(base, block_sizes, sizeof(int32_t) * static_cast<size_t>(b.systems), hipMemcpyHostToDevice));
    return PM_OK;
}

constexpr int64_t kFuseStepEntriesPerRow = 32;

// The factored preconditioner of a batch: G lower-triangular and G^T, square
// like a, with a's dtype and index widths (32-bit mirrors when a uses them).
int batch_attach_factors(BatchProblem& b, const pm_csr* g, const pm_csr* gt, const uint8_t* sys_jacobi)
{
    if (!g || !gt) return fail(PM_INVALID_ARGUMENT, "batched factored pcg: null factor");
    PM_TRY(check_csr(g, "batched factored pcg (G)"));
    PM_TRY(check_csr(gt, "batched factored pcg (G^T)"));
    const pm_csr* a = b.a;
    for (const pm_csr* m : {g, gt})
    {
        if (m->rows != a->rows || m->cols != a->cols)
            return fail(PM_INVALID_ARGUMENT, "batched factored pcg: factor shape %lld x %lld does not match A",
                        static_cast<long long>(m->rows), static_cast<long long>(m->cols));
        if (m->dtype != a->dtype) return fail(PM_INVALID_ARGUMENT, "batched factored pcg: factor dtype differs from A");
        if (use_compact(a) ? (!m->row_offsets32 || !m->col_indices32) : (!m->row_offsets || !m->col_indices))
            return fail(PM_INVALID_ARGUMENT, "batched factored pcg: factor lacks the index arrays of A's width");
    }
    if (g->nnz != gt->nnz) return fail(PM_INVALID_ARGUMENT, "batched factored pcg: G and G^T differ in nnz");
    b.g = g;
    b.gt = gt;
    b.sys_jacobi = sys_jacobi;
    b.jacobi = false;
    // Folding the step into the products costs a second gather (of q) per
    // factor entry and saves one pass over x, p, q, r and a launch.  On the
    // held-out batch (256 diffusion boxes with 4-entry factor rows, 256 reduced
    // Schur systems with about 70) both forms measured the same (70.5 vs
    // 71.0 ms), so the fold is kept unless factor rows average 32 entries or
    // more.  PLAMATRIX_AMD_FUSE_STEP=0 / 1 forces either form; both give the
    // same bits.
    const char* fuse = std::getenv("PLAMATRIX_AMD_FUSE_STEP");
    b.fuse_step = fuse != nullptr ? fuse[0] != '0' : g->nnz < kFuseStepEntriesPerRow * g->rows;
    return PM_OK;
}

// ---- batched FP64 iterative refinement ----------------------------------------------
//
// Per outer step: r = b - A x with FP64 accumulation over FP64 x (per-chunk
// partial r.r and b.b), one wavefront per system reduces them in chunk order
// and decides converged (relres <= target), out of outer steps, or active with
// scale 1/||r||; the active systems' scaled residuals are solved by the FP32
// batched PCG (inactive systems get a zero right-hand side and stop at its
// initialisation) and x += ||r|| d.  The host reads the per-system records
// once per outer step.

struct BatchRefineState
{
    double norm;
    double initial;
    double relres;
    int32_t active;
    int32_t converged;
    int32_t outer;
    int32_t pad;
};

template <typename I>
__global__ __launch_bounds__(kBlock) void k_batch_residual64(int64_t chunks, const int32_t* __restrict__ chunk_sys,
                                                             const int64_t* __restrict__ sys_end,
                                                             const I* __restrict__ ro, const I* __restrict__ ci,
                                                             const float* __restrict__ v,
                                                             const double* __restrict__ b,
                                                             const double* __restrict__ x, double* __restrict__ r,
                                                             const BatchRefineState* __restrict__ rs,
                                                             double* __restrict__ prr, double* __restrict__ pbb)
{
    __shared__ float lv_all[kBatchWaves][kStreamCap];
    __shared__ I lc_all[kBatchWaves][kStreamCap];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    PM_FOR_CHUNKS(c, chunks, kBatchWaves)
    {
        const int32_t s = chunk_sys[c];
        if (s < 0 || !rs[s].active)
        {
            if (lane == 0) prr[c] = pbb[c] = 0.0;
            continue;
        }
        const int64_t end = sys_end[s];
        const int64_t row0 = c * kChunkRows;
        int64_t ks, ke;
        const double acc = chunk_row_sums<double>(row0, end, ro, ci, lv_all[wave], lc_all[wave],
                                                  Fp64Source<float, double>{v, x}, ks, ke);
        const int64_t row = row0 + lane;
        double rr = 0.0, bb = 0.0;
        if (row < end)
        {
            const double bi = b[row];
            const double ri = bi - acc;
            r[row] = ri;
            rr = ri * ri;
            bb = bi * bi;
        }
        rr = wave_sum(rr);
        bb = wave_sum(bb);
        if (lane == 0)
        {
            prr[c] = rr;
            pbb[c] = bb;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_batch_refine_init(int64_t systems, BatchRefineState* rs)
{
    for (int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; s < systems;
         s += static_cast<int64_t>(gridDim.x) * blockDim.x)
    {
        BatchRefineState w{};
        w.active = 1;
        rs[s] = w;
    }
}

// One wavefront per active system: relative residual, stop decision, scale.
__global__ __launch_bounds__(kBlock) void k_batch_refine_decide(int64_t systems, const int64_t* __restrict__ sys_chunk0,
                                                                const int64_t* __restrict__ sys_chunk1,
                                                                const double* __restrict__ prr,
                                                                const double* __restrict__ pbb, BatchRefineState* rs,
                                                                int32_t outer, double target, int32_t max_outer)
{
    const int lane = threadIdx.x % kWave;
    const int64_t waves = static_cast<int64_t>(gridDim.x) * (blockDim.x / kWave);
    for (int64_t s = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave; s < systems;
         s += waves)
    {
        if (!rs[s].active) continue;
        double rr = 0.0, bb = 0.0;
        for (int64_t c = sys_chunk0[s] + lane; c < sys_chunk1[s]; c += kWave)
        {
            rr += prr[c];
            bb += pbb[c];
        }
        rr = wave_sum(rr);
        bb = wave_sum(bb);
        if (lane != 0) continue;
        BatchRefineState w = rs[s];
        const double relres = bb > 0.0 ? sqrt(rr / bb) : sqrt(rr);
        if (outer == 0) w.initial = relres;
        w.relres = relres;
        w.outer = outer;
        if (!isfinite(rr)) w.active = 0;
        else if (relres <= target || rr == 0.0)
        {
            w.active = 0;
            w.converged = 1;
        }
        else if (outer >= max_outer) w.active = 0;
        else w.norm = sqrt(rr);
        rs[s] = w;
    }
}

// Active systems: rf = r / ||r||, d = 0; others: rf = 0, d = 0.
__global__ __launch_bounds__(kBlock) void k_batch_refine_scale(int64_t chunks, const int32_t* __restrict__ chunk_sys,
                                                               const int64_t* __restrict__ sys_end,
                                                               const double* __restrict__ r,
                                                               const BatchRefineState* __restrict__ rs,
                                                               float* __restrict__ rf, float* __restrict__ d)
{
    const int lane = threadIdx.x % kWave;
    PM_FOR_CHUNKS(c, chunks, kBatchWaves)
    {
        const int32_t s = chunk_sys[c];
        const int64_t row = c * kChunkRows + lane;
        const bool act = s >= 0 && rs[s].active && row < sys_end[s];
        rf[row] = act ? static_cast<float>(r[row] / rs[s].norm) : 0.f;
        d[row] = 0.f;
    }
}

// Active systems: x += ||r|| d.
__global__ __launch_bounds__(kBlock) void k_batch_refine_accumulate(int64_t chunks,
                                                                    const int32_t* __restrict__ chunk_sys,
                                                                    const int64_t* __restrict__ sys_end,
                                                                    const BatchRefineState* __restrict__ rs,
                                                                    const float* __restrict__ d,
                                                                    double* __restrict__ x)
{
    const int lane = threadIdx.x % kWave;
    PM_FOR_CHUNKS(c, chunks, kBatchWaves)
    {
        const int32_t s = chunk_sys[c];
        const int64_t row = c * kChunkRows + lane;
        if (s >= 0 && rs[s].active && row < sys_end[s])
            x[row] = __builtin_fma(rs[s].norm, static_cast<double>(d[row]), x[row]);
    }
}

int batch_refine(BatchProblem& b, const double* rhs, double* x, double target, int32_t max_outer,
                 const BatchPolicy& pol, pm_batch_refine_report* reports, pm_workspace* ws, hipStream_t stream)
{
    const int64_t n = b.a->rows, systems = b.systems, chunks = b.chunks;
    auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
    const size_t rows_alloc = static_cast<size_t>(std::max<int64_t>(chunks, 1)) * kChunkRows;
    const size_t c1 = static_cast<size_t>(std::max<int64_t>(chunks, 1));
    const size_t s1 = static_cast<size_t>(std::max<int64_t>(systems, 1));
    const size_t bytes = al(sizeof(double) * rows_alloc) + 2 * al(sizeof(float) * rows_alloc) +
                         2 * al(sizeof(double) * c1) + al(sizeof(BatchRefineState) * s1) + al(sizeof(double) * s1);
    PM_TRY(ws->refine.reserve(bytes));
    char* p = ws->refine.as<char>();
    double* r64 = reinterpret_cast<double*>(p);
    p += al(sizeof(double) * rows_alloc);
    float* rf = reinterpret_cast<float*>(p);
    p += al(sizeof(float) * rows_alloc);
    float* d = reinterpret_cast<float*>(p);
    p += al(sizeof(float) * rows_alloc);
    double* prr = reinterpret_cast<double*>(p);
    p += al(sizeof(double) * c1);
    double* pbb = reinterpret_cast<double*>(p);
    p += al(sizeof(double) * c1);
    BatchRefineState* rs = reinterpret_cast<BatchRefineState*>(p);
    p += al(sizeof(BatchRefineState) * s1);
    double* sys_rel = reinterpret_cast<double*>(p);
    std::vector<double> host_rel(s1, b.rel_tol);
    b.rhs = rf;
    b.x = d;
    b.sys_rel = sys_rel;
    const int gs = static_cast<int>(std::min<int64_t>(ceil_div(systems, kBlock), 1024));
    hipLaunchKernelGGL(k_batch_refine_init, dim3(gs), dim3(kBlock), 0, stream, systems, rs);
    std::vector<BatchRefineState> host(s1);
    std::vector<pm_batch_report> inner(s1);
    std::vector<int64_t> inner_total(s1, 0);
    std::vector<int32_t> last_status(s1, 0);
    for (int32_t outer = 0;; ++outer)
    {
        if (use_compact(b.a))
            hipLaunchKernelGGL(k_batch_residual64<int32_t>, dim3(b.grid_chunks), dim3(kBlock), 0, stream, chunks,
                               b.chunk_sys, b.sys_end, b.a->row_offsets32, b.a->col_indices32,
                               static_cast<const float*>(b.a->values), rhs, x, r64, rs, prr, pbb);
        else
            hipLaunchKernelGGL(k_batch_residual64<int64_t>, dim3(b.grid_chunks), dim3(kBlock), 0, stream, chunks,
                               b.chunk_sys, b.sys_end, b.a->row_offsets, b.a->col_indices,
                               static_cast<const float*>(b.a->values), rhs, x, r64, rs, prr, pbb);
        hipLaunchKernelGGL(k_batch_refine_decide, dim3(b.grid_systems), dim3(kBlock), 0, stream, systems,
                           b.sys_chunk0, b.sys_chunk1, prr, pbb, rs, outer, target, max_outer);
        PM_HIP(hipGetLastError());
        PM_HIP(hipMemcpyAsync(host.data(), rs, sizeof(BatchRefineState) * systems, hipMemcpyDeviceToHost, stream));
        PM_HIP(hipStreamSynchronize(stream));
        int64_t active = 0;
        for (int64_t s = 0; s < systems; ++s) active += host[s].active ? 1 : 0;
        if (active == 0) break;
        hipLaunchKernelGGL(k_batch_refine_scale, dim3(b.grid_chunks), dim3(kBlock), 0, stream, chunks, b.chunk_sys,
                           b.sys_end, r64, rs, rf, d);
        PM_HIP(hipGetLastError());
        // each system's inner tolerance for this pass (the host copy of the
        // states above is current)
        for (int64_t s = 0; s < systems; ++s)
            host_rel[s] = refine_pass_tolerance(b.rel_tol, host[s].relres, target);
        PM_HIP(hipMemcpyAsync(sys_rel, host_rel.data(), sizeof(double) * systems, hipMemcpyHostToDevice, stream));
        PM_TRY(batch_solve<float>(b, pol, inner.data(), ws, stream));
        // the host array stays unchanged until the stream has read it: the
        // next pass first synchronises on its state copy
        // later inner solves reuse this matrix's Jacobi and block inverses
        b.inverses_ready = true;
        for (int64_t s = 0; s < systems; ++s)
            if (host[s].active)
            {
                inner_total[s] += inner[s].iterations;
                last_status[s] = inner[s].status;
            }
        hipLaunchKernelGGL(k_batch_refine_accumulate, dim3(b.grid_chunks), dim3(kBlock), 0, stream, chunks,
                           b.chunk_sys, b.sys_end, rs, d, x);
        PM_HIP(hipGetLastError());
    }
    PM_HIP(hipEventRecord(ws->done_event, stream));
    for (int64_t s = 0; s < systems; ++s)
    {
        pm_batch_refine_report& o = reports[s];
        o = pm_batch_refine_report{};
        o.converged = host[s].converged;
        o.outer_iterations = host[s].outer;
        o.inner_iterations = inner_total[s];
        o.initial_relres = host[s].initial;
        o.final_relres = host[s].relres;
        o.last_inner_status = last_status[s];
    }
    (void)n;
    return PM_OK;
}

} // namespace pm

using namespace pm;

extern "C" {

int pm_batch_layout(int64_t systems, const int64_t* sizes, int64_t* row_begin, int64_t* total_rows)
{
    if (systems < 0 || (systems > 0 && (!sizes || !row_begin)) || !total_rows)
        return fail(PM_INVALID_ARGUMENT, "pm_batch_layout: invalid arguments");
    int64_t cursor = 0;
    for (int64_t s = 0; s < systems; ++s)
    {
        if (sizes[s] < 0) return fail(PM_INVALID_ARGUMENT, "pm_batch_layout: negative system size");
        row_begin[s] = cursor;
        cursor += ceil_div(sizes[s], kChunkRows) * kChunkRows;
    }
    *total_rows = cursor;
    return PM_OK;
}

int pm_batch_pcg(const pm_csr* a, int64_t systems, const int64_t* row_begin, const int64_t* row_end,
                 const void* rhs, void* x, const pm_solver_options* options, const pm_solve_policy* policy,
                 pm_batch_report* reports, pm_workspace* ws, hipStream_t stream)
{
    BatchProblem b;
    PM_TRY(batch_setup(a, systems, row_begin, row_end, rhs, x, options, reports != nullptr, ws, stream, b));
    if (systems == 0) return PM_OK;
    const BatchPolicy pol = resolve_batch_policy(policy);
    if (a->dtype == PM_F32) return batch_solve<float>(b, pol, reports, ws, stream);
    return batch_solve<double>(b, pol, reports, ws, stream);
}

static int batch_refine_entry(const pm_csr* a, int64_t systems, const int64_t* row_begin, const int64_t* row_end,
                              const double* rhs, double* x, const pm_csr* g, const pm_csr* gt,
                              const uint8_t* sys_jacobi, const pm_solver_options* inner, double target_relres,
                              int32_t max_outer, const pm_solve_policy* policy, pm_batch_refine_report* reports,
                              pm_workspace* ws, hipStream_t stream, const int32_t* block_sizes = nullptr)
{
    if (!(target_relres >= 0.0) || max_outer < 0)
        return fail(PM_INVALID_ARGUMENT, "batched refinement: invalid target or outer limit");
    PM_TRY(check_csr(a, "batched refinement"));
    if (a->dtype != PM_F32) return fail(PM_INVALID_ARGUMENT, "batched refinement expects an FP32 matrix");
    if (!inner) return fail(PM_INVALID_ARGUMENT, "batched refinement: null inner options");
    pm_solver_options in = *inner;
    in.absolute_tolerance = 0.0;
    in.require_convergence = 0;
    // the inner solves run on workspace vectors; rhs / x here are FP64 and
    // only checked for presence
    BatchProblem b;
    PM_TRY(batch_setup(a, systems, row_begin, row_end, rhs, x, &in, reports != nullptr, ws, stream, b));
    if (g || gt) PM_TRY(batch_attach_factors(b, g, gt, sys_jacobi));
    if (block_sizes) PM_TRY(batch_attach_blocks(b, row_begin, row_end, block_sizes, ws, stream));
    if (systems == 0) return PM_OK;
    return batch_refine(b, rhs, x, target_relres, max_outer, resolve_batch_policy(policy), reports, ws, stream);
}

int pm_batch_pcg_refine(const pm_csr* a, int64_t systems, const int64_t* row_begin, const int64_t* row_end,
                        const double* rhs, double* x, const pm_solver_options* inner, double target_relres,
                        int32_t max_outer, const pm_solve_policy* policy, pm_batch_refine_report* reports,
                        pm_workspace* ws, hipStream_t stream)
{
    return batch_refine_entry(a, systems, row_begin, row_end, rhs, x, nullptr, nullptr, nullptr, inner, target_relres,
                              max_outer, policy, reports, ws, stream);
}

int pm_batch_pcg_factored(const pm_csr* a, int64_t systems, const int64_t* row_begin, const int64_t* row_end,
                          const void* rhs, void* x, const pm_csr* g, const pm_csr* gt, const uint8_t* jacobi_systems,
                          const pm_solver_options* options, const pm_solve_policy* policy, pm_batch_report* reports,
                          pm_workspace* ws, hipStream_t stream)
{
    BatchProblem b;
    PM_TRY(batch_setup(a, systems, row_begin, row_end, rhs, x, options, reports != nullptr, ws, stream, b));
    PM_TRY(batch_attach_factors(b, g, gt, jacobi_systems));
    if (systems == 0) return PM_OK;
    const BatchPolicy pol = resolve_batch_policy(policy);
    if (a->dtype == PM_F32) return batch_solve<float>(b, pol, reports, ws, stream);
    return batch_solve<double>(b, pol, reports, ws, stream);
}

int pm_batch_pcg_refine_factored(const pm_csr* a, int64_t systems, const int64_t* row_begin, const int64_t* row_end,
                                 const double* rhs, double* x, const pm_csr* g, const pm_csr* gt,
                                 const uint8_t* jacobi_systems, const pm_solver_options* inner, double target_relres,
                                 int32_t max_outer, const pm_solve_policy* policy, pm_batch_refine_report* reports,
                                 pm_workspace* ws, hipStream_t stream)
{
    if (!g || !gt) return fail(PM_INVALID_ARGUMENT, "batched factored refinement: null factor");
    return batch_refine_entry(a, systems, row_begin, row_end, rhs, x, g, gt, jacobi_systems, inner, target_relres,
                              max_outer, policy, reports, ws, stream);
}

int pm_batch_block_pcg(const pm_csr* a, int64_t systems, const int64_t* row_begin, const int64_t* row_end,
                       const int32_t* block_sizes, const void* rhs, void* x, const pm_solver_options* options,
                       const pm_solve_policy* policy, pm_batch_report* reports, pm_workspace* ws, hipStream_t stream)
{
    BatchProblem b;
    PM_TRY(batch_setup(a, systems, row_begin, row_end, rhs, x, options, reports != nullptr, ws, stream, b));
    PM_TRY(batch_attach_blocks(b, row_begin, row_end, block_sizes, ws, stream));
    if (systems == 0) return PM_OK;
    const BatchPolicy pol = resolve_batch_policy(policy);
    if (a->dtype == PM_F32) return batch_solve<float>(b, pol, reports, ws, stream);
    return batch_solve<double>(b, pol, reports, ws, stream);
}

int pm_batch_block_pcg_refine(const pm_csr* a, int64_t systems, const int64_t* row_begin, const int64_t* row_end,
                              const int32_t* block_sizes, const double* rhs, double* x,
                              const pm_solver_options* inner, double target_relres, int32_t max_outer,
                              const pm_solve_policy* policy, pm_batch_refine_report* reports, pm_workspace* ws,
                              hipStream_t stream)
{
    if (systems > 0 && !block_sizes) return fail(PM_INVALID_ARGUMENT, "batched block refinement: null block sizes");
    return batch_refine_entry(a, systems, row_begin, row_end, rhs, x, nullptr, nullptr, nullptr, inner, target_relres,
                              max_outer, policy, reports, ws, stream, block_sizes);
}

} // extern "C"
