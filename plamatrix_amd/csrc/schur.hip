// Damped Schur-complement solve of block LM normal equations on gfx950.
//
// Replaces solveDampedSchurComplement with the accelerated backend
// (block_schur.cpp:53-459, block_schur_accelerated.cpp:80-299,
// block_schur_device_assembly.cu:19-256) with a pipeline that keeps every
// array in HBM:
//
//   normal equations  From a residual batch (one primary and one eliminated
//                     block per residual) with Huber IRLS weights.  Residuals
//                     are grouped by primary, by eliminated and by (primary,
//                     eliminated) pair with the stable radix sort; each output
//                     entry is one thread that folds its group in input order,
//                     so B, C, W, g come out in the operation order of
//                     BlockNormalEquations::addResidualBlocks
//                     (block_normal_equations.cpp:135-312) with no atomics.
//                     Cross blocks are numbered in first-appearance order, as
//                     findOrCreateCrossBlock numbers them.
//   pattern           Built on the device from the cross-block topology and
//                     cached in the workspace until the topology changes
//                     (prepareSchurTopology, block_schur_sparse_assembly.h:
//                     23-272).  Every (left, right) primary pair of every
//                     eliminated block, the diagonal and the primary cross
//                     blocks become 2-D block keys; one radix sort orders them
//                     like the reference's std::map; runs of equal keys are the
//                     block slots, and each slot's terms stay in (eliminated,
//                     left, right) order.
//   prologue          Damping, Cholesky inverses of the eliminated blocks, the
//                     reduced right-hand side and the transformed cross
//                     blocks W C^-1, one thread per block or per row.
//   assembly          One wavefront per block slot, one lane per entry of the
//                     ps x ps block: the slot's terms are folded in order, an
//                     atomic-free segmented reduction over the fixed pattern.
//   solve             Block-Jacobi PCG (pm_block_pcg) on the scalar CSR of the
//                     reduced system with the inverted diagonal blocks.
//   back-substitution One thread per eliminated block.
//
// Kernels that restate a reference loop keep its operation order and are
// compiled without floating-point contraction, so their results equal the
// CPU path's bit for bit given the same inputs.
#include <algorithm>
#include <cmath>
#include <limits>
#include <cstdlib>
#include <unordered_map>
#include <cstring>

#include "dense_cholesky.hpp"
#include "pm_common.hpp"
#include "solver_internal.hpp"
#include "sort.hpp"
#include "workspace.hpp"

namespace pm
{

constexpr int kMaxSchurBlock = 16;  // ps and es limit
constexpr int kSmallSchurBlock = 8;  // kernels with per-thread block arrays use this size when it suffices

struct SchurStatus
{
    unsigned long long bad_primary;     // lowest residual with an out-of-range primary block
    unsigned long long bad_eliminated;  // lowest residual with an out-of-range eliminated block
    unsigned long long bad_value;       // lowest residual with a non-finite Jacobian / residual entry
    unsigned long long bad_weight;      // lowest residual with a non-finite or negative weight
    unsigned long long bad_inverse;     // lowest eliminated block that is not positive definite
    unsigned long long bad_preconditioner;  // lowest primary block whose Schur diagonal is not PD
    unsigned long long mismatch;        // topology differences from the cached copy
    unsigned long long count0;          // small device-side counts read back by the host
    unsigned long long count1;
    unsigned long long pad[7];
};

__global__ void k_schur_status_init(SchurStatus* st)
{
    st->bad_primary = ~0ull;
    st->bad_eliminated = ~0ull;
    st->bad_value = ~0ull;
    st->bad_weight = ~0ull;
    st->bad_inverse = ~0ull;
    st->bad_preconditioner = ~0ull;
    st->mismatch = 0;
    st->count0 = 0;
    st->count1 = 0;
}

__device__ __forceinline__ int64_t gtid() { return static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; }
__device__ __forceinline__ int64_t gstride() { return static_cast<int64_t>(gridDim.x) * blockDim.x; }

inline int grid_for(int64_t n)
{
    return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, kBlock), 16384)));
}

// ---- small dense algebra in the reference's operation order -------------------

// evaluateHuberLoss(s, delta).weight (robust_loss.h:27-51).
__device__ double huber_weight(double s, double delta)
{
    if (delta == 0.0 || s <= delta * delta) return 1.0;
    return delta / sqrt(s);
}

// invertPositiveDefiniteInto (block_schur_linear_algebra.h:92-147): Cholesky
// factor, then one forward and one backward substitution per column.
// N > 0 fixes the size at compile time: the loops unroll and the arrays stay
// in registers instead of per-lane scratch memory.
template <int MB, int N = 0>
__device__ bool invert_pd(const double* m, int n_rt, double* inv)
{
#pragma clang fp contract(off)
    const int n = N > 0 ? N : n_rt;
    double lower[MB * MB];
    double mid[MB];
#pragma unroll
    for (int i = 0; i < n * n; ++i) lower[i] = 0.0;
#pragma unroll
    for (int row = 0; row < n; ++row)
#pragma unroll
        for (int col = 0; col <= row; ++col)
        {
            double value = m[row * n + col];
#pragma unroll
            for (int k = 0; k < col; ++k) value -= lower[row * n + k] * lower[col * n + k];
            if (row == col)
            {
                if (!(value > 0.0) || !isfinite(value)) return false;
                lower[row * n + col] = sqrt(value);
            }
            else
            {
                lower[row * n + col] = value / lower[col * n + col];
            }
        }
#pragma unroll
    for (int i = 0; i < n * n; ++i) inv[i] = 0.0;
#pragma unroll
    for (int col = 0; col < n; ++col)
    {
#pragma unroll
        for (int row = 0; row < n; ++row)
        {
            double value = row == col ? 1.0 : 0.0;
#pragma unroll
            for (int k = 0; k < row; ++k) value -= lower[row * n + k] * mid[k];
            mid[row] = value / lower[row * n + row];
        }
#pragma unroll
        for (int off = 0; off < n; ++off)
        {
            const int row = n - 1 - off;
            double value = mid[row];
#pragma unroll
            for (int k = row + 1; k < n; ++k) value -= lower[k * n + row] * inv[k * n + col];
            inv[row * n + col] = value / lower[row * n + row];
        }
    }
    return true;
}

// ---- generic helpers ------------------------------------------------------------

// out[x] = number of sorted keys whose (key >> shift) is below x, x in [0, m].
__global__ __launch_bounds__(kBlock) void k_group_offsets(const unsigned long long* __restrict__ keys, int64_t n,
                                                          int shift, int64_t m, int64_t* __restrict__ out)
{
    for (int64_t x = gtid(); x <= m; x += gstride())
    {
        int64_t lo = 0, hi = n;
        while (lo < hi)
        {
            const int64_t mid = (lo + hi) / 2;
            if (static_cast<int64_t>(keys[mid] >> shift) < x) lo = mid + 1;
            else hi = mid;
        }
        out[x] = lo;
    }
}

__global__ __launch_bounds__(kBlock) void k_count_mismatch(const int64_t* __restrict__ a,
                                                           const int64_t* __restrict__ b, int64_t n,
                                                           SchurStatus* st)
{
    bool diff = false;
    for (int64_t i = gtid(); i < n; i += gstride()) diff |= a[i] != b[i];
    if (diff) atomicAdd(&st->mismatch, 1ull);
}

__global__ __launch_bounds__(kBlock) void k_copy_i64(const int64_t* __restrict__ src, int64_t n,
                                                     int64_t* __restrict__ dst)
{
    for (int64_t i = gtid(); i < n; i += gstride()) dst[i] = src[i];
}

__global__ __launch_bounds__(kBlock) void k_fill_f64(double* __restrict__ dst, int64_t n, double v)
{
    for (int64_t i = gtid(); i < n; i += gstride()) dst[i] = v;
}

// FP64 -> FP32 (rounded to nearest) and FP32 -> FP64 copies for the
// mixed-precision seed solve.
__global__ __launch_bounds__(kBlock) void k_f64_to_f32(const double* __restrict__ src, int64_t n,
                                                       float* __restrict__ dst)
{
    for (int64_t i = gtid(); i < n; i += gstride()) dst[i] = static_cast<float>(src[i]);
}

__global__ __launch_bounds__(kBlock) void k_f32_to_f64(const float* __restrict__ src, int64_t n,
                                                       double* __restrict__ dst)
{
    for (int64_t i = gtid(); i < n; i += gstride()) dst[i] = static_cast<double>(src[i]);
}

__global__ __launch_bounds__(kBlock) void k_u32_to_i64(const unsigned* __restrict__ src, int64_t n,
                                                       int64_t* __restrict__ dst)
{
    for (int64_t i = gtid(); i < n; i += gstride()) dst[i] = static_cast<int64_t>(src[i]);
}

// ---- normal equations from residuals -------------------------------------------

// Validation, weights and the three grouping keys of every residual.
// RS, PS, ES > 0: the residual and block sizes at compile time (unrolled
// loops, same arithmetic).
template <int RS = 0, int PS = 0, int ES = 0>
__global__ __launch_bounds__(kBlock) void k_residual_keys(
    int64_t count, int64_t pc, int64_t ec, int rs_rt, int ps_rt, int es_rt, int eb, const int64_t* __restrict__ primary,
    const int64_t* __restrict__ eliminated, const double* __restrict__ jp, const double* __restrict__ je,
    const double* __restrict__ r, const double* __restrict__ weights, double delta, double* __restrict__ w_out,
    unsigned long long* __restrict__ key_p, unsigned long long* __restrict__ key_e,
    unsigned long long* __restrict__ key_pe, unsigned* __restrict__ val_p, unsigned* __restrict__ val_e,
    unsigned* __restrict__ val_pe, SchurStatus* st)
{
#pragma clang fp contract(off)
    const int rs = RS > 0 ? RS : rs_rt, ps = PS > 0 ? PS : ps_rt, es = ES > 0 ? ES : es_rt;
    for (int64_t i = gtid(); i < count; i += gstride())
    {
        const int64_t p = primary[i], e = eliminated[i];
        const bool p_ok = p >= 0 && p < pc;
        const bool e_ok = e >= 0 && e < ec;
        if (!p_ok) atomicMin(&st->bad_primary, static_cast<unsigned long long>(i));
        if (!e_ok) atomicMin(&st->bad_eliminated, static_cast<unsigned long long>(i));
        bool finite = true;
        double s = 0.0;
        for (int k = 0; k < rs; ++k)
        {
            const double v = r[i * rs + k];
            finite &= isfinite(v);
            s += v * v;
        }
        for (int k = 0; k < rs * ps; ++k) finite &= isfinite(jp[i * rs * ps + k]);
        for (int k = 0; k < rs * es; ++k) finite &= isfinite(je[i * rs * es + k]);
        if (!finite) atomicMin(&st->bad_value, static_cast<unsigned long long>(i));
        double w;
        if (weights)
        {
            w = weights[i];
            if (!isfinite(w) || w < 0.0) atomicMin(&st->bad_weight, static_cast<unsigned long long>(i));
        }
        else
        {
            w = huber_weight(s, delta);
        }
        w_out[i] = w;
        const unsigned long long kp = p_ok ? static_cast<unsigned long long>(p) : 0ull;
        const unsigned long long ke = e_ok ? static_cast<unsigned long long>(e) : 0ull;
        key_p[i] = kp;
        key_e[i] = ke;
        key_pe[i] = (kp << eb) | ke;
        val_p[i] = val_e[i] = val_pe[i] = static_cast<unsigned>(i);
    }
}

// Heads of the (primary, eliminated) runs: first[i] = 1 for the first
// residual (in input order) of each pair.
__global__ __launch_bounds__(kBlock) void k_pair_heads(const unsigned long long* __restrict__ keys,
                                                       const unsigned* __restrict__ perm, int64_t n,
                                                       unsigned* __restrict__ head, unsigned* __restrict__ first)
{
    for (int64_t i = gtid(); i < n; i += gstride())
    {
        const bool h = i == 0 || keys[i] != keys[i - 1];
        head[i] = h ? 1u : 0u;
        if (h) first[perm[i]] = 1u;
    }
}

__global__ __launch_bounds__(kBlock) void k_zero_u32(unsigned* __restrict__ p, int64_t n)
{
    for (int64_t i = gtid(); i < n; i += gstride()) p[i] = 0u;
}

// Per run (in sorted order): its start, and the cross index (creation order)
// with the block ids of the cross block.
__global__ __launch_bounds__(kBlock) void k_pair_runs(const unsigned long long* __restrict__ keys,
                                                      const unsigned* __restrict__ perm,
                                                      const unsigned* __restrict__ head,
                                                      const unsigned* __restrict__ run_id,
                                                      const unsigned* __restrict__ creation, int64_t n, int eb,
                                                      int64_t* __restrict__ run_start,
                                                      unsigned* __restrict__ run_cross,
                                                      int64_t* __restrict__ cross_primary,
                                                      int64_t* __restrict__ cross_eliminated)
{
    const unsigned long long emask = (1ull << eb) - 1ull;
    for (int64_t i = gtid(); i < n; i += gstride())
    {
        if (!head[i]) continue;
        const int64_t run = run_id[i];
        const unsigned c = creation[perm[i]];
        run_start[run] = i;
        run_cross[run] = c;
        cross_primary[c] = static_cast<int64_t>(keys[i] >> eb);
        cross_eliminated[c] = static_cast<int64_t>(keys[i] & emask);
    }
}

// B_p and g_p: thread per (primary, entry); entries [0, ps*ps) are B, then g.
// RS > 0 (residual size known at compile time) loads the operands of
// kNeBatch residuals before accumulating them in order; RS == 0 takes rs at
// run time, one residual at a time.
constexpr int kNeBatch = 8;
template <int RS>
__global__ __launch_bounds__(kBlock) void k_ne_primary(int64_t pc, int rs_runtime, int ps,
                                                       const int64_t* __restrict__ offs,
                                                       const unsigned* __restrict__ perm,
                                                       const double* __restrict__ jp, const double* __restrict__ r,
                                                       const double* __restrict__ w, double* __restrict__ B,
                                                       double* __restrict__ g)
{
#pragma clang fp contract(off)
    const int rs = RS > 0 ? RS : rs_runtime;
    const int entries = ps * ps + ps;
    for (int64_t t = gtid(); t < pc * entries; t += gstride())
    {
        const int64_t p = t / entries;
        const int en = static_cast<int>(t % entries);
        const bool is_b = en < ps * ps;
        const int column = is_b ? en / ps : en - ps * ps;
        const int other = is_b ? en % ps : 0;
        double acc = 0.0;
        const int64_t k1 = offs[p + 1];
        if (RS > 0)
        {
            for (int64_t k = offs[p]; k < k1; k += kNeBatch)
            {
                double wi[kNeBatch], a[kNeBatch][RS > 0 ? RS : 1], b[kNeBatch][RS > 0 ? RS : 1];
#pragma unroll
                for (int u = 0; u < kNeBatch; ++u)
                {
                    const bool in = k + u < k1;
                    const int64_t i = in ? perm[k + u] : 0;
                    const double* J = jp + i * RS * ps;
                    wi[u] = in ? w[i] : 0.0;
#pragma unroll
                    for (int row = 0; row < RS; ++row)
                    {
                        a[u][row] = in ? J[row * ps + column] : 0.0;
                        b[u][row] = in ? (is_b ? J[row * ps + other] : r[i * RS + row]) : 0.0;
                    }
                }
#pragma unroll
                for (int u = 0; u < kNeBatch; ++u)
                {
                    if (k + u >= k1) break;
#pragma unroll
                    for (int row = 0; row < RS; ++row) acc += wi[u] * a[u][row] * b[u][row];
                }
            }
        }
        else
        {
            for (int64_t k = offs[p]; k < k1; ++k)
            {
                const int64_t i = perm[k];
                const double wi = w[i];
                const double* J = jp + i * rs * ps;
                for (int row = 0; row < rs; ++row)
                {
                    const double value = J[row * ps + column];
                    acc += wi * value * (is_b ? J[row * ps + other] : r[i * rs + row]);
                }
            }
        }
        if (is_b) B[p * ps * ps + en] = acc;
        else g[p * ps + column] = acc;
    }
}

// B_p and g_p with one wavefront per primary (lane = entry, ps*ps + ps of
// them, RS x PS Jacobians): the wavefront stages the Jacobian rows, residuals
// and weights of kNeStage residuals into LDS with coalesced loads, then each
// lane accumulates its entry over them in the primary's residual order, the
// arithmetic of k_ne_primary.  With one wavefront per primary the kernel is
// latency-bound (2,000 wavefronts on the LM system), so each round stages 64
// residuals: 124 us with 16 per round, 213 us with 128.
constexpr int kNeStage = 64;
template <int RS, int PS>
__global__ __launch_bounds__(kBlock) void k_ne_primary_staged(int64_t pc, const int64_t* __restrict__ offs,
                                                              const unsigned* __restrict__ perm,
                                                              const double* __restrict__ jp,
                                                              const double* __restrict__ r,
                                                              const double* __restrict__ w, double* __restrict__ B,
                                                              double* __restrict__ g)
{
#pragma clang fp contract(off)
    static_assert(PS * PS + PS <= kWave, "one lane per entry");
    constexpr int kPer = RS * PS + RS + 1;  // J rows, residual, weight
    constexpr int kVals = kNeStage * kPer;
    constexpr int kPerLane = (kVals + kWave - 1) / kWave;
    __shared__ double stage_all[kBlock / kWave][kVals];
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    double* st = stage_all[wave];
    const bool active = lane < PS * PS + PS;
    const bool is_b = lane < PS * PS;
    const int column = is_b ? lane / PS : lane - PS * PS;
    const int other = is_b ? lane % PS : 0;
    const int64_t waves = static_cast<int64_t>(gridDim.x) * (kBlock / kWave);
    for (int64_t p = static_cast<int64_t>(blockIdx.x) * (kBlock / kWave) + wave; p < pc; p += waves)
    {
        const int64_t k0 = offs[p], k1 = offs[p + 1];
        double acc = 0.0;
        for (int64_t k = k0; k < k1; k += kNeStage)
        {
            const int nr = static_cast<int>(k1 - k < kNeStage ? k1 - k : kNeStage);
            const int total = nr * kPer;
            unsigned id[kPerLane];
#pragma unroll
            for (int m = 0; m < kPerLane; ++m)
            {
                const int f = lane + m * kWave;
                id[m] = f < total ? perm[k + f / kPer] : 0u;
            }
            double val[kPerLane];
#pragma unroll
            for (int m = 0; m < kPerLane; ++m)
            {
                const int f = lane + m * kWave;
                const int within = f % kPer;
                const int64_t i = id[m];
                val[m] = 0.0;
                if (f < total)
                    val[m] = within < RS * PS ? jp[i * RS * PS + within]
                                              : (within < RS * PS + RS ? r[i * RS + (within - RS * PS)] : w[i]);
            }
#pragma unroll
            for (int m = 0; m < kPerLane; ++m)
            {
                const int f = lane + m * kWave;
                if (f < total) st[f] = val[m];
            }
            wave_lds_sync();
            if (active)
            {
                for (int u = 0; u < nr; ++u)
                {
                    const double* J = st + u * kPer;
                    const double* res = J + RS * PS;
                    const double wi = J[RS * PS + RS];
#pragma unroll
                    for (int row = 0; row < RS; ++row)
                    {
                        const double value = J[row * PS + column];
                        acc += wi * value * (is_b ? J[row * PS + other] : res[row]);
                    }
                }
            }
            wave_lds_sync();
        }
        if (active)
        {
            if (is_b) B[p * PS * PS + lane] = acc;
            else g[p * PS + column] = acc;
        }
    }
}

// C_e and g_e: thread per (eliminated, entry).
__global__ __launch_bounds__(kBlock) void k_ne_eliminated(int64_t ec, int rs, int es,
                                                          const int64_t* __restrict__ offs,
                                                          const unsigned* __restrict__ perm,
                                                          const double* __restrict__ je,
                                                          const double* __restrict__ r,
                                                          const double* __restrict__ w, double* __restrict__ C,
                                                          double* __restrict__ g)
{
#pragma clang fp contract(off)
    const int entries = es * es + es;
    for (int64_t t = gtid(); t < ec * entries; t += gstride())
    {
        const int64_t e = t / entries;
        const int en = static_cast<int>(t % entries);
        const bool is_c = en < es * es;
        const int column = is_c ? en / es : en - es * es;
        const int other = is_c ? en % es : 0;
        double acc = 0.0;
        for (int64_t k = offs[e]; k < offs[e + 1]; ++k)
        {
            const int64_t i = perm[k];
            const double wi = w[i];
            const double* J = je + i * rs * es;
            for (int row = 0; row < rs; ++row)
            {
                const double value = J[row * es + column];
                acc += wi * value * (is_c ? J[row * es + other] : r[i * rs + row]);
            }
        }
        if (is_c) C[e * es * es + en] = acc;
        else g[e * es + column] = acc;
    }
}

// k_ne_eliminated with the sizes at compile time and one thread per
// eliminated block: its residuals are loaded once for all ES * ES + ES
// entries (each summed in the same order as above, so the same bits).
template <int RS, int ES>
__global__ __launch_bounds__(kBlock) void k_ne_eliminated_points(int64_t ec, const int64_t* __restrict__ offs,
                                                                 const unsigned* __restrict__ perm,
                                                                 const double* __restrict__ je,
                                                                 const double* __restrict__ r,
                                                                 const double* __restrict__ w,
                                                                 double* __restrict__ C, double* __restrict__ g)
{
#pragma clang fp contract(off)
    for (int64_t e = gtid(); e < ec; e += gstride())
    {
        double acc[ES * ES + ES];
#pragma unroll
        for (int en = 0; en < ES * ES + ES; ++en) acc[en] = 0.0;
        for (int64_t k = offs[e]; k < offs[e + 1]; ++k)
        {
            const int64_t i = perm[k];
            const double wi = w[i];
            double J[RS * ES], res[RS];
#pragma unroll
            for (int j = 0; j < RS * ES; ++j) J[j] = je[i * RS * ES + j];
#pragma unroll
            for (int j = 0; j < RS; ++j) res[j] = r[i * RS + j];
#pragma unroll
            for (int column = 0; column < ES; ++column)
            {
#pragma unroll
                for (int other = 0; other < ES; ++other)
#pragma unroll
                    for (int row = 0; row < RS; ++row)
                        acc[column * ES + other] += wi * J[row * ES + column] * J[row * ES + other];
#pragma unroll
                for (int row = 0; row < RS; ++row) acc[ES * ES + column] += wi * J[row * ES + column] * res[row];
            }
        }
#pragma unroll
        for (int en = 0; en < ES * ES; ++en) C[e * ES * ES + en] = acc[en];
#pragma unroll
        for (int column = 0; column < ES; ++column) g[e * ES + column] = acc[ES * ES + column];
    }
}

// W blocks (ps x es, row-major), thread per (run, entry), stored at the run's
// cross index.
template <int RS = 0, int PS = 0, int ES = 0>
__global__ __launch_bounds__(kBlock) void k_ne_cross(int64_t runs, int64_t n, int rs_rt, int ps_rt, int es_rt,
                                                     const int64_t* __restrict__ run_start,
                                                     const unsigned* __restrict__ run_cross,
                                                     const unsigned* __restrict__ perm,
                                                     const double* __restrict__ jp, const double* __restrict__ je,
                                                     const double* __restrict__ w, double* __restrict__ W)
{
#pragma clang fp contract(off)
    const int rs = RS > 0 ? RS : rs_rt, ps = PS > 0 ? PS : ps_rt, es = ES > 0 ? ES : es_rt;
    const int entries = ps * es;
    for (int64_t t = gtid(); t < runs * entries; t += gstride())
    {
        const int64_t run = t / entries;
        const int en = static_cast<int>(t % entries);
        const int pcol = en / es, ecol = en % es;
        const int64_t end = run + 1 < runs ? run_start[run + 1] : n;
        double acc = 0.0;
        for (int64_t k = run_start[run]; k < end; ++k)
        {
            const int64_t i = perm[k];
            const double wi = w[i];
            for (int row = 0; row < rs; ++row)
                acc += wi * jp[i * rs * ps + row * ps + pcol] * je[i * rs * es + row * es + ecol];
        }
        W[static_cast<int64_t>(run_cross[run]) * entries + en] = acc;
    }
}

// Per residual: its cross block when it is the only residual of its
// (primary, eliminated) pair, ~0u otherwise.
__global__ __launch_bounds__(kBlock) void k_residual_cross(int64_t runs, int64_t n,
                                                           const int64_t* __restrict__ run_start,
                                                           const unsigned* __restrict__ perm,
                                                           const unsigned* __restrict__ run_cross,
                                                           unsigned* __restrict__ res_cross)
{
    for (int64_t run = gtid(); run < runs; run += gstride())
    {
        const int64_t b = run_start[run], e = run + 1 < runs ? run_start[run + 1] : n;
        const unsigned c = e - b == 1 ? run_cross[run] : ~0u;
        for (int64_t k = b; k < e; ++k) res_cross[perm[k]] = c;
    }
}

// The cross blocks of single-residual pairs, a thread per residual in
// residual order: the Jacobian rows are read in order, and cross blocks are
// numbered in the order residuals first reach them, so the stores follow
// too.  k_ne_cross_runs' arithmetic for a run of one residual.
template <int RS, int PS, int ES>
__global__ __launch_bounds__(kBlock) void k_ne_cross_single(int64_t n, const unsigned* __restrict__ res_cross,
                                                            const double* __restrict__ jp,
                                                            const double* __restrict__ je,
                                                            const double* __restrict__ w, double* __restrict__ W)
{
#pragma clang fp contract(off)
    for (int64_t i = gtid(); i < n; i += gstride())
    {
        const unsigned c = res_cross[i];
        if (c == ~0u) continue;
        const double wi = w[i];
        double a[RS * PS], b[RS * ES];
#pragma unroll
        for (int j = 0; j < RS * PS; ++j) a[j] = jp[i * RS * PS + j];
#pragma unroll
        for (int j = 0; j < RS * ES; ++j) b[j] = je[i * RS * ES + j];
        double* out = W + static_cast<int64_t>(c) * PS * ES;
#pragma unroll
        for (int pcol = 0; pcol < PS; ++pcol)
#pragma unroll
            for (int ecol = 0; ecol < ES; ++ecol)
            {
                double acc = 0.0;
#pragma unroll
                for (int row = 0; row < RS; ++row) acc += wi * a[row * PS + pcol] * b[row * ES + ecol];
                out[pcol * ES + ecol] = acc;
            }
    }
}

// k_ne_cross with the sizes at compile time and one thread per run: the
// run's residuals are loaded once for all PS * ES entries of its block (each
// entry summed in the same order as above, so the same bits).
template <int RS, int PS, int ES>
__global__ __launch_bounds__(kBlock) void k_ne_cross_runs(int64_t runs, int64_t n,
                                                          const int64_t* __restrict__ run_start,
                                                          const unsigned* __restrict__ run_cross,
                                                          const unsigned* __restrict__ perm,
                                                          const double* __restrict__ jp,
                                                          const double* __restrict__ je,
                                                          const double* __restrict__ w, double* __restrict__ W,
                                                          bool skip_single)
{
#pragma clang fp contract(off)
    for (int64_t run = gtid(); run < runs; run += gstride())
    {
        const int64_t end = run + 1 < runs ? run_start[run + 1] : n;
        if (skip_single && end - run_start[run] == 1) continue;  // k_ne_cross_single's
        double acc[PS * ES];
#pragma unroll
        for (int en = 0; en < PS * ES; ++en) acc[en] = 0.0;
        for (int64_t k = run_start[run]; k < end; ++k)
        {
            const int64_t i = perm[k];
            const double wi = w[i];
            double a[RS * PS], b[RS * ES];
#pragma unroll
            for (int j = 0; j < RS * PS; ++j) a[j] = jp[i * RS * PS + j];
#pragma unroll
            for (int j = 0; j < RS * ES; ++j) b[j] = je[i * RS * ES + j];
#pragma unroll
            for (int pcol = 0; pcol < PS; ++pcol)
#pragma unroll
                for (int ecol = 0; ecol < ES; ++ecol)
#pragma unroll
                    for (int row = 0; row < RS; ++row)
                        acc[pcol * ES + ecol] += wi * a[row * PS + pcol] * b[row * ES + ecol];
        }
        double* out = W + static_cast<int64_t>(run_cross[run]) * PS * ES;
#pragma unroll
        for (int en = 0; en < PS * ES; ++en) out[en] = acc[en];
    }
}

// ---- Schur pattern ---------------------------------------------------------------

// Range checks of caller-supplied topology: cross ids in range, primary
// cross blocks with row < col < pc.
__global__ __launch_bounds__(kBlock) void k_validate_topology(int64_t nc, int64_t npc, int64_t pc, int64_t ec,
                                                              const int64_t* __restrict__ cp,
                                                              const int64_t* __restrict__ ce,
                                                              const int64_t* __restrict__ pr,
                                                              const int64_t* __restrict__ pcl, SchurStatus* st)
{
    for (int64_t c = gtid(); c < nc; c += gstride())
    {
        if (cp[c] < 0 || cp[c] >= pc) atomicMin(&st->bad_primary, static_cast<unsigned long long>(c));
        if (ce[c] < 0 || ce[c] >= ec) atomicMin(&st->bad_eliminated, static_cast<unsigned long long>(c));
    }
    for (int64_t c = gtid(); c < npc; c += gstride())
        if (pr[c] < 0 || pcl[c] <= pr[c] || pcl[c] >= pc) atomicMin(&st->bad_value, static_cast<unsigned long long>(c));
}

// Sort keys of the cross blocks: eliminated id (shift 0) or (primary, eliminated).
__global__ __launch_bounds__(kBlock) void k_cross_keys(int64_t nc, const int64_t* __restrict__ cp,
                                                       const int64_t* __restrict__ ce, int shift,
                                                       unsigned long long* __restrict__ keys,
                                                       unsigned* __restrict__ vals)
{
    for (int64_t c = gtid(); c < nc; c += gstride())
    {
        const unsigned long long e = static_cast<unsigned long long>(ce[c]);
        keys[c] = shift ? ((static_cast<unsigned long long>(cp[c]) << shift) | e) : e;
        vals[c] = static_cast<unsigned>(c);
    }
}

__global__ __launch_bounds__(kBlock) void k_fill_i64(int64_t* __restrict__ dst, int64_t n, int64_t v)
{
    for (int64_t i = gtid(); i < n; i += gstride()) dst[i] = v;
}

// count0 = total pair terms sum_e k_e^2 (64-bit; integer atomics are exact).
__global__ __launch_bounds__(kBlock) void k_sum_pairs(int64_t ec, const int64_t* __restrict__ adj_offs,
                                                      SchurStatus* st)
{
    unsigned long long total = 0;
    for (int64_t e = gtid(); e < ec; e += gstride())
    {
        const unsigned long long k = static_cast<unsigned long long>(adj_offs[e + 1] - adj_offs[e]);
        total += k * k;
    }
    if (total) atomicAdd(&st->count0, total);
}

// Pair-term count per eliminated block (k_e^2).
__global__ __launch_bounds__(kBlock) void k_pair_counts(int64_t ec, const int64_t* __restrict__ adj_offs,
                                                        unsigned* __restrict__ counts)
{
    for (int64_t e = gtid(); e < ec; e += gstride())
    {
        const int64_t k = adj_offs[e + 1] - adj_offs[e];
        counts[e] = static_cast<unsigned>(k * k);
    }
}

// Keys of every pair term (eliminated-major, then left, then right), then the
// diagonal and both orientations of each primary cross block.
__global__ __launch_bounds__(kBlock) void k_pattern_keys(int64_t ec, int64_t pc, int64_t npc, int pb,
                                                         const int64_t* __restrict__ adj_offs,
                                                         const unsigned* __restrict__ adj,
                                                         const unsigned* __restrict__ pair_offs,
                                                         const int64_t* __restrict__ cross_primary,
                                                         const int64_t* __restrict__ pc_rows,
                                                         const int64_t* __restrict__ pc_cols, int64_t terms,
                                                         unsigned long long* __restrict__ keys,
                                                         unsigned* __restrict__ vals)
{
    for (int64_t e = gtid(); e < ec; e += gstride())
    {
        const int64_t a0 = adj_offs[e], k = adj_offs[e + 1] - a0;
        int64_t t = pair_offs[e];
        for (int64_t l = 0; l < k; ++l)
        {
            const unsigned long long pl = static_cast<unsigned long long>(cross_primary[adj[a0 + l]]);
            for (int64_t rr = 0; rr < k; ++rr)
            {
                const unsigned long long pr = static_cast<unsigned long long>(cross_primary[adj[a0 + rr]]);
                keys[t] = (pl << pb) | pr;
                vals[t] = static_cast<unsigned>(t);
                ++t;
            }
        }
    }
    for (int64_t b = gtid(); b < pc; b += gstride())
    {
        keys[terms + b] = (static_cast<unsigned long long>(b) << pb) | static_cast<unsigned long long>(b);
        vals[terms + b] = static_cast<unsigned>(terms + b);
    }
    for (int64_t c = gtid(); c < npc; c += gstride())
    {
        const unsigned long long row = static_cast<unsigned long long>(pc_rows[c]);
        const unsigned long long col = static_cast<unsigned long long>(pc_cols[c]);
        const int64_t base = terms + pc + 2 * c;
        keys[base] = (row << pb) | col;
        vals[base] = static_cast<unsigned>(base);
        keys[base + 1] = (col << pb) | row;
        vals[base + 1] = static_cast<unsigned>(base + 1);
    }
}

__global__ __launch_bounds__(kBlock) void k_key_heads(const unsigned long long* __restrict__ keys, int64_t n,
                                                      unsigned* __restrict__ heads)
{
    for (int64_t i = gtid(); i < n; i += gstride()) heads[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// Slot records from the run heads: key, run start.
__global__ __launch_bounds__(kBlock) void k_slots(const unsigned long long* __restrict__ keys,
                                                  const unsigned* __restrict__ heads,
                                                  const unsigned* __restrict__ slot_id, int64_t n,
                                                  unsigned long long* __restrict__ slot_key,
                                                  int64_t* __restrict__ slot_start)
{
    for (int64_t i = gtid(); i < n; i += gstride())
    {
        if (!heads[i]) continue;
        const int64_t s = slot_id[i];
        slot_key[s] = keys[i];
        slot_start[s] = i;
    }
}

// Decode each sorted entry: term -> (eliminated, left cross, right cross);
// diagonal / primary-cross markers -> kind in term_e (~0u diagonal, ~1u
// primary cross, with the primary cross index and orientation in term_left).
__global__ __launch_bounds__(kBlock) void k_decode_terms(const unsigned* __restrict__ vals, int64_t n,
                                                         int64_t terms, int64_t pc, int64_t ec,
                                                         const unsigned* __restrict__ pair_offs,
                                                         const int64_t* __restrict__ adj_offs,
                                                         const unsigned* __restrict__ adj,
                                                         unsigned* __restrict__ term_e,
                                                         unsigned* __restrict__ term_left,
                                                         unsigned* __restrict__ term_right)
{
    for (int64_t i = gtid(); i < n; i += gstride())
    {
        const int64_t t = vals[i];
        if (t < terms)
        {
            int64_t lo = 0, hi = ec;  // last e with pair_offs[e] <= t
            while (lo < hi)
            {
                const int64_t mid = (lo + hi + 1) / 2;
                if (static_cast<int64_t>(pair_offs[mid]) <= t) lo = mid;
                else hi = mid - 1;
            }
            const int64_t e = lo;
            const int64_t k = adj_offs[e + 1] - adj_offs[e];
            const int64_t local = t - static_cast<int64_t>(pair_offs[e]);
            term_e[i] = static_cast<unsigned>(e);
            term_left[i] = adj[adj_offs[e] + local / k];
            term_right[i] = adj[adj_offs[e] + local % k];
        }
        else if (t < terms + pc)
        {
            term_e[i] = ~0u;
            term_left[i] = 0;
            term_right[i] = 0;
        }
        else
        {
            const int64_t q = t - terms - pc;
            term_e[i] = ~1u;
            term_left[i] = static_cast<unsigned>(q / 2);
            term_right[i] = static_cast<unsigned>(q % 2);
        }
    }
}

// Scalar CSR structure: block row br has slots [bro[br], bro[br+1]); scalar
// row br*ps + lr starts at bro[br]*ps*ps + lr*(slots in row)*ps; entry
// (slot, lr, lc) sits at ordinal*ps + lc within that row
// (block_schur_sparse_assembly.h:129-154).
__global__ __launch_bounds__(kBlock) void k_scalar_rows(int64_t pc, int ps, const int64_t* __restrict__ bro,
                                                        int64_t* __restrict__ ro, int32_t* __restrict__ ro32)
{
    for (int64_t r = gtid(); r <= pc * ps; r += gstride())
    {
        int64_t v;
        if (r == pc * ps)
        {
            v = bro[pc] * ps * ps;
        }
        else
        {
            const int64_t br = r / ps, lr = r % ps;
            v = bro[br] * ps * ps + lr * (bro[br + 1] - bro[br]) * ps;
        }
        ro[r] = v;
        ro32[r] = static_cast<int32_t>(v);
    }
}

__global__ __launch_bounds__(kBlock) void k_scalar_cols(int64_t slots, int ps, int pb,
                                                        const unsigned long long* __restrict__ slot_key,
                                                        const int64_t* __restrict__ bro,
                                                        const int64_t* __restrict__ ro, int64_t* __restrict__ ci,
                                                        int32_t* __restrict__ ci32)
{
    const unsigned long long mask = (1ull << pb) - 1ull;
    for (int64_t t = gtid(); t < slots * ps * ps; t += gstride())
    {
        const int64_t s = t / (ps * ps);
        const int lr = static_cast<int>((t / ps) % ps), lc = static_cast<int>(t % ps);
        const int64_t row = static_cast<int64_t>(slot_key[s] >> pb);
        const int64_t col = static_cast<int64_t>(slot_key[s] & mask);
        const int64_t pos = ro[row * ps + lr] + (s - bro[row]) * ps + lc;
        ci[pos] = col * ps + lc;
        ci32[pos] = static_cast<int32_t>(col * ps + lc);
    }
}

// ---- numerics ----------------------------------------------------------------------

// D_ii += lambda * max(|D_ii|, minimum) (block_schur.cpp:100-128) on a copy.
__global__ __launch_bounds__(kBlock) void k_damp(const double* __restrict__ src, int64_t blocks, int size,
                                                 double damping, double minimum, double* __restrict__ dst)
{
#pragma clang fp contract(off)
    const int64_t n = blocks * size * size;
    for (int64_t i = gtid(); i < n; i += gstride())
    {
        double v = src[i];
        const int64_t local = i % (size * size);
        if (local / size == local % size)
        {
            const double magnitude = fabs(v);
            v += damping * (magnitude < minimum ? minimum : magnitude);
        }
        dst[i] = v;
    }
}

// C^-1 per eliminated block and y_e = C^-1 g_e.
template <int MB, int N = 0>
__global__ __launch_bounds__(kBlock) void k_eliminated_inverse(int64_t ec, int es_rt, const double* __restrict__ Cd,
                                                               const double* __restrict__ ge,
                                                               double* __restrict__ Cinv, double* __restrict__ y,
                                                               SchurStatus* st)
{
#pragma clang fp contract(off)
    const int es = N > 0 ? N : es_rt;
    for (int64_t e = gtid(); e < ec; e += gstride())
    {
        double m[MB * MB], inv[MB * MB];
#pragma unroll
        for (int i = 0; i < es * es; ++i) m[i] = Cd[e * es * es + i];
        if (!invert_pd<MB, N>(m, es, inv))
        {
            atomicMin(&st->bad_inverse, static_cast<unsigned long long>(e));
#pragma unroll
            for (int i = 0; i < es * es; ++i) inv[i] = 0.0;
        }
#pragma unroll
        for (int i = 0; i < es * es; ++i) Cinv[e * es * es + i] = inv[i];
#pragma unroll
        for (int row = 0; row < es; ++row)
        {
            double value = 0.0;
#pragma unroll
            for (int col = 0; col < es; ++col) value += inv[row * es + col] * ge[e * es + col];
            y[e * es + row] = value;
        }
    }
}

// Per cross block c and row: v_c[row] = W_c[row] . y_e (one thread each),
// stored in the primaries' summation order (position k of prim_cross), so
// that k_reduced_rhs streams each primary's products contiguously.
// PS, ES > 0: the block sizes at compile time (the divisions by ps become
// multiplications), same arithmetic.
template <int PS = 0, int ES = 0>
__global__ __launch_bounds__(kBlock) void k_cross_rhs(int64_t nc, int ps_rt, int es_rt,
                                                      const unsigned* __restrict__ prim_cross,
                                                      const int64_t* __restrict__ cross_eliminated,
                                                      const double* __restrict__ W, const double* __restrict__ y,
                                                      double* __restrict__ crv)
{
#pragma clang fp contract(off)
    const int ps = PS > 0 ? PS : ps_rt, es = ES > 0 ? ES : es_rt;
    for (int64_t t = gtid(); t < nc * ps; t += gstride())
    {
        const int64_t c = prim_cross[t / ps];
        const int row = static_cast<int>(t % ps);
        const double* m = W + c * ps * es + row * es;
        const double* v = y + cross_eliminated[c] * es;
        double value = 0.0;
        for (int col = 0; col < es; ++col) value += m[col] * v[col];
        crv[t] = value;
    }
}

// k_transform_cross and k_cross_rhs in one pass over W for PS x ES blocks:
// thread (c, row) reads row `row` of cross block c once and forms its W C^-1
// row and its W C^-1 g product, the latter stored at the block's position in
// the primaries' order (cross_pos, the inverse of prim_cross).  Same sums in
// the same order as the two kernels.
template <int PS, int ES>
__global__ __launch_bounds__(kBlock) void k_transform_rhs(int64_t nc, const int64_t* __restrict__ cross_eliminated,
                                                          const unsigned* __restrict__ cross_pos,
                                                          const double* __restrict__ Cinv,
                                                          const double* __restrict__ y, const double* __restrict__ W,
                                                          double* __restrict__ TW, double* __restrict__ crv)
{
#pragma clang fp contract(off)
    for (int64_t t = gtid(); t < nc * PS; t += gstride())
    {
        const int64_t c = t / PS;
        const int row = static_cast<int>(t % PS);
        const int64_t e = cross_eliminated[c];
        double w[ES], v[ES];
#pragma unroll
        for (int j = 0; j < ES; ++j)
        {
            w[j] = W[c * PS * ES + row * ES + j];
            v[j] = y[e * ES + j];
        }
#pragma unroll
        for (int i = 0; i < ES; ++i)
        {
            const double* inv = Cinv + e * ES * ES + i * ES;
            double value = 0.0;
#pragma unroll
            for (int j = 0; j < ES; ++j) value += inv[j] * w[j];
            TW[c * PS * ES + row * ES + i] = value;
        }
        double value = 0.0;
#pragma unroll
        for (int col = 0; col < ES; ++col) value += w[col] * v[col];
        crv[static_cast<int64_t>(cross_pos[c]) * PS + row] = value;
    }
}

// cross_pos[prim_cross[k]] = k.
__global__ __launch_bounds__(kBlock) void k_invert_order(int64_t n, const unsigned* __restrict__ order,
                                                         unsigned* __restrict__ pos)
{
    for (int64_t k = gtid(); k < n; k += gstride()) pos[order[k]] = static_cast<unsigned>(k);
}

// rhs_p = -g_p + sum over the primary's cross blocks in (primary, eliminated)
// order of W C^-1 g_e (block_schur.cpp:158-181), the products precomputed by
// k_cross_rhs; the ordered sum loads kRhsBatch products ahead of the adds.
constexpr int kRhsBatch = 64;  // one thread per rhs entry (pc * ps of them): deep prefetch for the few chains
__global__ __launch_bounds__(kBlock) void k_reduced_rhs(int64_t pc, int ps, const double* __restrict__ gp,
                                                        const int64_t* __restrict__ prim_offs,
                                                        const double* __restrict__ crv, double* __restrict__ rhs)
{
#pragma clang fp contract(off)
    for (int64_t t = gtid(); t < pc * ps; t += gstride())
    {
        const int64_t p = t / ps;
        const int row = static_cast<int>(t % ps);
        double acc = -gp[t];
        const int64_t k1 = prim_offs[p + 1];
        for (int64_t k = prim_offs[p]; k < k1; k += kRhsBatch)
        {
            double v[kRhsBatch];
#pragma unroll
            for (int u = 0; u < kRhsBatch; ++u)
                v[u] = k + u < k1 ? crv[(k + u) * ps + row] : 0.0;
#pragma unroll
            for (int u = 0; u < kRhsBatch; ++u)
                if (k + u < k1) acc += 1.0 * v[u];
        }
        rhs[t] = acc;
    }
}

// TW_c[row][i] = sum_j Cinv_e[i][j] W_c[row][j] (multiplyMatrixVector per row).
template <int PS = 0, int ES = 0>
__global__ __launch_bounds__(kBlock) void k_transform_cross(int64_t nc, int ps_rt, int es_rt,
                                                            const int64_t* __restrict__ cross_eliminated,
                                                            const double* __restrict__ Cinv,
                                                            const double* __restrict__ W, double* __restrict__ TW)
{
#pragma clang fp contract(off)
    const int ps = PS > 0 ? PS : ps_rt, es = ES > 0 ? ES : es_rt;
    // one thread per output entry (c, row, i): consecutive threads store
    // consecutive entries, and the threads of one row read the same W row
    const int64_t bv = static_cast<int64_t>(ps) * es;
    for (int64_t t = gtid(); t < nc * bv; t += gstride())
    {
        const int64_t c = t / bv;
        const int rem = static_cast<int>(t % bv);
        const int row = rem / es, i = rem % es;
        const double* inv = Cinv + cross_eliminated[c] * es * es + i * es;
        const double* w = W + c * bv + row * es;
        double value = 0.0;
        for (int j = 0; j < es; ++j) value += inv[j] * w[j];
        TW[t] = value;
    }
}

// One wavefront per block slot, lane = (lr, lc).  Value = damped B on the
// diagonal, + each primary cross block (forward, or transposed for the
// mirrored slot), then - (W_l C^-1)(W_r)^T for every term in order
// (block_schur_sparse_assembly.h:817-903).  ES > 0 (the eliminated block
// size known at compile time): the whole wavefront stages the W C^-1 and W
// blocks of kTermStage terms into LDS with coalesced loads (indices first,
// then data), then each lane accumulates its entry over them in term order.
// ES == 0 takes es at run time, one term at a time.
constexpr int kTermStage = 8;
template <int ES, int PS>
__global__ __launch_bounds__(kBlock) void k_assemble_slots(
    int64_t slots, int ps, int es_runtime, int pb, const unsigned long long* __restrict__ slot_key,
    const int64_t* __restrict__ slot_start, int64_t entries, const unsigned* __restrict__ term_e,
    const unsigned* __restrict__ term_left, const unsigned* __restrict__ term_right,
    const int64_t* __restrict__ bro, const int64_t* __restrict__ ro, const double* __restrict__ Bd,
    const double* __restrict__ pcv, const double* __restrict__ TW, const double* __restrict__ W,
    double* __restrict__ values)
{
#pragma clang fp contract(off)
    static_assert(ES == 0 || PS > 0, "the staged path needs both block sizes");
    constexpr int kStageVals = kTermStage * 2 * (PS > 0 ? PS : 1) * (ES > 0 ? ES : 1);
    constexpr int kStagePerLane = (kStageVals + kWave - 1) / kWave;
    __shared__ double stage_all[ES > 0 ? kBlock / kWave : 1][ES > 0 ? kStageVals : 1];
    const int es = ES > 0 ? ES : es_runtime;
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    const unsigned long long mask = (1ull << pb) - 1ull;
    // each workgroup owns a contiguous, XCD-local range of slots: neighbouring
    // slots share their W C^-1 and W blocks, which then stay in one L2
    const int64_t wpb = blockDim.x / kWave;
    const int64_t per = ceil_div(ceil_div(slots, static_cast<int64_t>(gridDim.x)), wpb) * wpb;
    const int64_t sbegin = xcd_logical_block(blockIdx.x, gridDim.x) * per;
    const int64_t send = sbegin + per < slots ? sbegin + per : slots;
    for (int64_t s = sbegin + wave; s < send; s += wpb)
    {
        const int64_t row = static_cast<int64_t>(slot_key[s] >> pb);
        const int64_t col = static_cast<int64_t>(slot_key[s] & mask);
        const int64_t begin = slot_start[s];
        const int64_t end = s + 1 < slots ? slot_start[s + 1] : entries;
        // blocks above 8 x 8 have more entries than lanes: 64 at a time
        for (int e0 = 0; e0 < ps * ps; e0 += kWave)
        {
        const int ent = e0 + lane;
        const bool active = ent < ps * ps;
        const int lr = active ? ent / ps : 0, lc = active ? ent % ps : 0;
        double acc = 0.0;
        // the run holds terms first, then the diagonal marker, then primary
        // cross markers in index order
        if (active && row == col) acc += Bd[row * ps * ps + lr * ps + lc];
        // the markers are the few entries at the run's end
        int64_t tend = end;
        while (tend > begin && term_e[tend - 1] >= ~1u) --tend;
        if (active)
        {
            for (int64_t k = tend; k < end; ++k)
            {
                if (term_e[k] != ~1u) continue;
                const int64_t q = term_left[k];
                const double* m = pcv + q * ps * ps;
                acc += term_right[k] == 0 ? m[lr * ps + lc] : m[lc * ps + lr];
            }
        }
        if (ES > 0)
        {
            double* st = stage_all[ES > 0 ? wave : 0];
            constexpr int bvals = (PS > 0 ? PS : 1) * (ES > 0 ? ES : 1);  // doubles of one W C^-1 or W block
            constexpr int per_term = 2 * bvals;                                // W C^-1 block, then W block
            for (int64_t k = begin; k < tend; k += kTermStage)
            {
                const int nt = static_cast<int>(tend - k < kTermStage ? tend - k : kTermStage);
                const int total = nt * per_term;
                unsigned id[kStagePerLane];
#pragma unroll
                for (int m = 0; m < kStagePerLane; ++m)
                {
                    const int f = lane + m * kWave;
                    id[m] = 0u;
                    if (f < total) id[m] = (f % per_term) < bvals ? term_left[k + f / per_term] : term_right[k + f / per_term];
                }
                double val[kStagePerLane];
#pragma unroll
                for (int m = 0; m < kStagePerLane; ++m)
                {
                    const int f = lane + m * kWave;
                    const int within = f % per_term;
                    val[m] = 0.0;
                    if (f < total)
                        val[m] = within < bvals ? TW[static_cast<int64_t>(id[m]) * bvals + within]
                                                : W[static_cast<int64_t>(id[m]) * bvals + (within - bvals)];
                }
#pragma unroll
                for (int m = 0; m < kStagePerLane; ++m)
                {
                    const int f = lane + m * kWave;
                    if (f < total) st[f] = val[m];
                }
                wave_lds_sync();
                if (active)
                {
                    for (int u = 0; u < nt; ++u)
                    {
                        const double* t = st + u * per_term + lr * ES;
                        const double* w = st + u * per_term + bvals + lc * ES;
                        double value = 0.0;
#pragma unroll
                        for (int i = 0; i < ES; ++i) value += t[i] * w[i];
                        acc += -value;
                    }
                }
                wave_lds_sync();
            }
        }
        else if (active)
        {
            for (int64_t k = begin; k < tend; ++k)
            {
                const double* t = TW + static_cast<int64_t>(term_left[k]) * ps * es + lr * es;
                const double* w = W + static_cast<int64_t>(term_right[k]) * ps * es + lc * es;
                double value = 0.0;
                for (int i = 0; i < es; ++i) value += t[i] * w[i];
                acc += -value;
            }
        }
        if (active) values[ro[row * ps + lr] + (s - bro[row]) * ps + lc] = acc;
        }
    }
}

// ---- segmented assembly ------------------------------------------------------------
//
// Each slot's terms are cut into segments of at most kSegTerms (in the slot's
// eliminated order).  k_assemble_segments gives every segment one wavefront:
// the segment's W C^-1 and W blocks are staged into LDS with all loads in
// flight, and each lane (one entry of the 6 x 6 block) sums -(W_l C^-1) W_r^T
// over the segment's terms in order.  k_assemble_reduce then forms each value
// as damped B, the primary cross markers and the segment sums in segment
// order.  The summation order is fixed (deterministic run to run) but groups
// the terms by segment, so values differ from the CPU accumulation
// (block_schur_sparse_assembly.h:817-903) by rounding only; the slot kernel
// above keeps the CPU order (PLAMATRIX_AMD_SCHUR_ASSEMBLY=slots).  Long slots
// (the diagonal ones hold every secondary of their primary) no longer run as
// one serial chain, and a launch has about 30 times more wavefronts.
//
// Only slots on and above the block diagonal get segments: S is symmetric
// (C^-1 is), so the value of slot (l, r) below the diagonal takes the segment
// sums of its mirror (r, l) transposed, which halves the gathered terms of the
// off-diagonal slots, and the lower triangle of a diagonal block takes its
// upper triangle's value (damped B included, which the normal equations leave
// symmetric only to rounding).  S is then symmetric bit for bit, and each
// value still differs from the CPU accumulation by rounding only.
constexpr int kSegTerms = 16;
// Segments a wavefront of k_assemble_segments sums.  One per wavefront, over
// a grid of every segment, measured fastest on the 2,000-primary LM system:
// 0.88 ms for the assembly against 0.90 with two and 0.94 with four (their
// loads in flight together cost occupancy: 134 VGPRs at four), and 1.01 ms
// for two segments per wavefront in sequence under a 65,536-workgroup grid.
constexpr int kSegsPerWave = 1;

// Per slot: the number of terms (the run minus its trailing markers) and of
// segments (none below the block diagonal); seg_count[slots] = 0 closes the
// scan.
__global__ __launch_bounds__(kBlock) void k_slot_segments(int64_t slots, int pb,
                                                          const unsigned long long* __restrict__ slot_key,
                                                          const int64_t* __restrict__ slot_start, int64_t entries,
                                                          const unsigned* __restrict__ term_e,
                                                          unsigned* __restrict__ slot_terms,
                                                          unsigned* __restrict__ seg_count)
{
    const unsigned long long mask = (1ull << pb) - 1ull;
    for (int64_t s = gtid(); s <= slots; s += gstride())
    {
        if (s == slots)
        {
            seg_count[s] = 0u;
            continue;
        }
        const int64_t begin = slot_start[s];
        const int64_t end = s + 1 < slots ? slot_start[s + 1] : entries;
        int64_t te = end;
        while (te > begin && term_e[te - 1] >= ~1u) --te;
        slot_terms[s] = static_cast<unsigned>(te - begin);
        const bool lower = (slot_key[s] >> pb) > (slot_key[s] & mask);
        seg_count[s] = lower ? 0u : static_cast<unsigned>((te - begin + kSegTerms - 1) / kSegTerms);
    }
}

// The slot of the mirrored block (c, r) of every slot (r, c) (the pattern is
// symmetric: every term and primary cross block lands in both).
__global__ __launch_bounds__(kBlock) void k_slot_mirror(int64_t slots, int pb,
                                                        const unsigned long long* __restrict__ slot_key,
                                                        unsigned* __restrict__ mirror)
{
    const unsigned long long mask = (1ull << pb) - 1ull;
    for (int64_t s = gtid(); s < slots; s += gstride())
    {
        const unsigned long long key = ((slot_key[s] & mask) << pb) | (slot_key[s] >> pb);
        int64_t lo = 0, hi = slots;
        while (lo < hi)
        {
            const int64_t mid = (lo + hi) / 2;
            if (slot_key[mid] < key) lo = mid + 1;
            else hi = mid;
        }
        mirror[s] = static_cast<unsigned>(lo < slots && slot_key[lo] == key ? lo : s);
    }
}

__global__ __launch_bounds__(kBlock) void k_segment_slots(int64_t slots, const unsigned* __restrict__ seg_offs,
                                                          unsigned* __restrict__ seg_slot)
{
    for (int64_t s = gtid(); s < slots; s += gstride())
        for (unsigned g = seg_offs[s]; g < seg_offs[s + 1]; ++g) seg_slot[g] = static_cast<unsigned>(s);
}

// Every segment's (left, right) term pairs at kSegTerms per segment, padded
// with kNoTerm: a wavefront of k_assemble_segments reaches its operand blocks
// in two dependent loads (pairs, then blocks) instead of walking segment ->
// slot -> term run first.  Built once per pattern.  On the 2,000-primary LM
// system: 284 us for the assembly kernel against 291 with the walk; the
// kernel stays bound by the 1.27 GB of gathered operand blocks.
constexpr unsigned kNoTerm = ~0u;
__global__ __launch_bounds__(kBlock) void k_segment_terms(int64_t nseg, const unsigned* __restrict__ seg_slot,
                                                          const unsigned* __restrict__ seg_offs,
                                                          const int64_t* __restrict__ slot_start,
                                                          const unsigned* __restrict__ slot_terms,
                                                          const unsigned* __restrict__ term_left,
                                                          const unsigned* __restrict__ term_right,
                                                          uint2* __restrict__ seg_terms)
{
    for (int64_t t = gtid(); t < nseg * kSegTerms; t += gstride())
    {
        const int64_t g = t / kSegTerms;
        const int u = static_cast<int>(t % kSegTerms);
        const unsigned s = seg_slot[g];
        const int64_t j = g - static_cast<int64_t>(seg_offs[s]);
        const int64_t k = j * kSegTerms + u;
        uint2 pair = make_uint2(kNoTerm, kNoTerm);
        if (k < static_cast<int64_t>(slot_terms[s]))
        {
            const int64_t e = slot_start[s] + k;
            pair = make_uint2(term_left[e], term_right[e]);
        }
        seg_terms[t] = pair;
    }
}

template <int ES, int PS>
__global__ __launch_bounds__(kBlock) void k_assemble_segments(int64_t nseg, const uint2* __restrict__ seg_terms,
                                                              const double* __restrict__ TW,
                                                              const double* __restrict__ W,
                                                              double* __restrict__ partial)
{
#pragma clang fp contract(off)
    constexpr int BV = PS * ES;              // doubles of one cross block
    constexpr int PER = 2 * BV;              // W C^-1 block, then W block
    constexpr int STAGE = kSegTerms * PER;
    constexpr int LOADS = (STAGE + kWave - 1) / kWave;
    constexpr int kWaves = kBlock / kWave;
    constexpr int SEGS = kSegsPerWave;
    __shared__ double stage_all[kWaves][STAGE];
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    double* st = stage_all[wave];
    const bool active = lane < PS * PS;
    const int lr = active ? lane / PS : 0, lc = active ? lane % PS : 0;
    // each wavefront owns SEGS consecutive segments and each workgroup a
    // contiguous, XCD-local range: neighbouring segments share slots and
    // blocks, which then stay in one L2.  The term pairs of all SEGS
    // segments are loaded before the first block load; a segment's valid
    // pairs are a prefix of its kSegTerms.
    const int64_t g0 = (xcd_logical_block(blockIdx.x, gridDim.x) * kWaves + wave) * SEGS;
    unsigned idl[SEGS], idr[SEGS];
#pragma unroll
    for (int q = 0; q < SEGS; ++q)
    {
        uint2 pair = make_uint2(kNoTerm, kNoTerm);
        if (lane < kSegTerms && g0 + q < nseg) pair = seg_terms[(g0 + q) * kSegTerms + lane];
        idl[q] = pair.x;
        idr[q] = pair.y;
    }
    int nt[SEGS];
#pragma unroll
    for (int q = 0; q < SEGS; ++q) nt[q] = static_cast<int>(__popcll(__ballot(idl[q] != kNoTerm)));
    double val[SEGS][LOADS];
#pragma unroll
    for (int q = 0; q < SEGS; ++q)
    {
#pragma unroll
        for (int m = 0; m < LOADS; ++m)
        {
            const int f = lane + m * kWave;
            const int u = f / PER, within = f % PER;
            const unsigned a = __shfl(idl[q], u < kWave ? u : 0, kWave);
            const unsigned b = __shfl(idr[q], u < kWave ? u : 0, kWave);
            val[q][m] = 0.0;
            if (f < STAGE && u < nt[q])
                val[q][m] = within < BV ? TW[static_cast<int64_t>(a) * BV + within]
                                        : W[static_cast<int64_t>(b) * BV + (within - BV)];
        }
    }
#pragma unroll
    for (int q = 0; q < SEGS; ++q)
    {
        if (nt[q] == 0) break;  // wave-uniform: later segments are past the end too
#pragma unroll
        for (int m = 0; m < LOADS; ++m)
        {
            const int f = lane + m * kWave;
            if (f < STAGE) st[f] = val[q][m];
        }
        wave_lds_sync();
        if (active)
        {
            double acc = 0.0;
            for (int u = 0; u < nt[q]; ++u)
            {
                const double* t = st + u * PER + lr * ES;
                const double* w = st + u * PER + BV + lc * ES;
                double value = 0.0;
#pragma unroll
                for (int i = 0; i < ES; ++i) value += t[i] * w[i];
                acc += -value;
            }
            partial[(g0 + q) * (PS * PS) + lane] = acc;
        }
        wave_lds_sync();
    }
}

// One wavefront per slot: damped B on the diagonal, the primary cross
// markers, then the segment sums in segment order (below the block diagonal:
// the mirrored slot's, transposed).
template <int PS>
__global__ __launch_bounds__(kBlock) void k_assemble_reduce(
    int64_t slots, int pb, const unsigned long long* __restrict__ slot_key, const int64_t* __restrict__ slot_start,
    int64_t entries, const unsigned* __restrict__ term_e, const unsigned* __restrict__ term_left,
    const unsigned* __restrict__ term_right, const unsigned* __restrict__ seg_offs,
    const unsigned* __restrict__ mirror, const int64_t* __restrict__ bro, const int64_t* __restrict__ ro,
    const double* __restrict__ Bd, const double* __restrict__ pcv, const double* __restrict__ partial,
    double* __restrict__ values)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x % kWave;
    const bool active = lane < PS * PS;
    const int lr = active ? lane / PS : 0, lc = active ? lane % PS : 0;
    const unsigned long long mask = (1ull << pb) - 1ull;
    const int64_t waves = static_cast<int64_t>(gridDim.x) * (kBlock / kWave);
    for (int64_t s = static_cast<int64_t>(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave; s < slots;
         s += waves)
    {
        if (!active) continue;
        const int64_t row = static_cast<int64_t>(slot_key[s] >> pb);
        const int64_t col = static_cast<int64_t>(slot_key[s] & mask);
        const int64_t end = s + 1 < slots ? slot_start[s + 1] : entries;
        // the lower triangle of a diagonal block is its upper triangle's
        // value: (er, ec) is the entry computed
        const bool diag_lower = row == col && lr > lc;
        const int er = diag_lower ? lc : lr, ec = diag_lower ? lr : lc;
        double acc = 0.0;
        if (row == col) acc += Bd[row * PS * PS + er * PS + ec];
        int64_t te = end;
        while (te > slot_start[s] && term_e[te - 1] >= ~1u) --te;
        for (int64_t k = te; k < end; ++k)
        {
            if (term_e[k] != ~1u) continue;
            const double* m = pcv + static_cast<int64_t>(term_left[k]) * PS * PS;
            acc += term_right[k] == 0 ? m[er * PS + ec] : m[ec * PS + er];
        }
        const bool lower = row > col;
        const int64_t src = lower ? static_cast<int64_t>(mirror[s]) : s;
        const int entry = lower ? ec * PS + er : er * PS + ec;
        const int64_t g1 = seg_offs[src + 1];
        constexpr int U = 8;
        for (int64_t g = seg_offs[src]; g < g1; g += U)
        {
            double v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = g + u < g1 ? partial[(g + u) * (PS * PS) + entry] : 0.0;
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (g + u < g1) acc += v[u];
        }
        values[ro[row * PS + lr] + (s - bro[row]) * PS + lc] = acc;
    }
}

// The segmented kernels serve 6 x 6 primaries with 3 x 3 eliminated blocks;
// PLAMATRIX_AMD_SCHUR_ASSEMBLY=slots keeps the slot kernel and with it the
// CPU accumulation order bit for bit.
bool segmented_assembly_wanted()
{
    const char* env = std::getenv("PLAMATRIX_AMD_SCHUR_ASSEMBLY");
    return env == nullptr || std::strcmp(env, "slots") != 0;
}

// Preconditioner: inverse of each assembled diagonal block.
template <int MB, int N = 0>
__global__ __launch_bounds__(kBlock) void k_preconditioner(int64_t pc, int ps_rt, const int64_t* __restrict__ diag_slot,
                                                           const int64_t* __restrict__ bro,
                                                           const int64_t* __restrict__ ro,
                                                           const double* __restrict__ values,
                                                           double* __restrict__ pinv, SchurStatus* st)
{
    const int ps = N > 0 ? N : ps_rt;
    for (int64_t b = gtid(); b < pc; b += gstride())
    {
        double m[MB * MB], inv[MB * MB];
        const int64_t s = diag_slot[b];
#pragma unroll
        for (int lr = 0; lr < ps; ++lr)
#pragma unroll
            for (int lc = 0; lc < ps; ++lc) m[lr * ps + lc] = values[ro[b * ps + lr] + (s - bro[b]) * ps + lc];
        if (!invert_pd<MB, N>(m, ps, inv))
        {
            atomicMin(&st->bad_preconditioner, static_cast<unsigned long long>(b));
#pragma unroll
            for (int i = 0; i < ps * ps; ++i) inv[i] = 0.0;
        }
#pragma unroll
        for (int i = 0; i < ps * ps; ++i) pinv[b * ps * ps + i] = inv[i];
    }
}

// Slot of the diagonal block of every primary.
__global__ __launch_bounds__(kBlock) void k_diag_slots(int64_t pc, int pb, int64_t slots,
                                                       const unsigned long long* __restrict__ slot_key,
                                                       int64_t* __restrict__ diag_slot)
{
    for (int64_t b = gtid(); b < pc; b += gstride())
    {
        const unsigned long long key = (static_cast<unsigned long long>(b) << pb) | static_cast<unsigned long long>(b);
        int64_t lo = 0, hi = slots;
        while (lo < hi)
        {
            const int64_t mid = (lo + hi) / 2;
            if (slot_key[mid] < key) lo = mid + 1;
            else hi = mid;
        }
        diag_slot[b] = lo;
    }
}

// Back-substitution, first half: t_c[col] = sum_row W_c[row][col] dp_p[row]
// for every cross block c (a thread per (c, col): W is read in order).
// PS, ES > 0: block sizes at compile time (same arithmetic).
template <int PS = 0, int ES = 0>
__global__ __launch_bounds__(kBlock) void k_cross_point(int64_t nc, int ps_rt, int es_rt,
                                                        const int64_t* __restrict__ cross_primary,
                                                        const double* __restrict__ W, const double* __restrict__ dp,
                                                        double* __restrict__ t)
{
#pragma clang fp contract(off)
    const int ps = PS > 0 ? PS : ps_rt, es = ES > 0 ? ES : es_rt;
    for (int64_t i = gtid(); i < nc * es; i += gstride())
    {
        const int64_t c = i / es;
        const int column = static_cast<int>(i % es);
        const double* values = W + c * ps * es + column;
        const double* primary = dp + cross_primary[c] * ps;
        double value = 0.0;
        for (int row = 0; row < ps; ++row) value += values[row * es] * primary[row];
        t[i] = value;
    }
}

// Second half: point = -g_e - t_c over the eliminated block's crosses in
// adjacency order, de_e = C^-1 point.  Together the arithmetic of
// k_back_substitute, operation for operation.
template <int MB, int N = 0>
__global__ __launch_bounds__(kBlock) void k_back_substitute_point(int64_t ec, int es_rt,
                                                                  const int64_t* __restrict__ adj_offs,
                                                                  const unsigned* __restrict__ adj,
                                                                  const double* __restrict__ t,
                                                                  const double* __restrict__ Cinv,
                                                                  const double* __restrict__ ge,
                                                                  double* __restrict__ de)
{
#pragma clang fp contract(off)
    const int es = N > 0 ? N : es_rt;
    for (int64_t e = gtid(); e < ec; e += gstride())
    {
        double point[MB];
#pragma unroll
        for (int i = 0; i < es; ++i) point[i] = -ge[e * es + i];
        const int64_t k1 = adj_offs[e + 1];
        for (int64_t k = adj_offs[e]; k < k1; ++k)
        {
            const double* tc = t + static_cast<int64_t>(adj[k]) * es;
#pragma unroll
            for (int column = 0; column < es; ++column) point[column] -= tc[column];
        }
        const double* inv = Cinv + e * es * es;
#pragma unroll
        for (int row = 0; row < es; ++row)
        {
            double value = 0.0;
#pragma unroll
            for (int col = 0; col < es; ++col) value += inv[row * es + col] * point[col];
            de[e * es + row] = value;
        }
    }
}

// Back-substitution (block_schur.cpp:422-454): de_e = C^-1 (-g_e - sum W^T dp).
template <int MB>
__global__ __launch_bounds__(kBlock) void k_back_substitute(int64_t ec, int ps, int es,
                                                            const int64_t* __restrict__ adj_offs,
                                                            const unsigned* __restrict__ adj,
                                                            const int64_t* __restrict__ cross_primary,
                                                            const double* __restrict__ W,
                                                            const double* __restrict__ Cinv,
                                                            const double* __restrict__ ge,
                                                            const double* __restrict__ dp, double* __restrict__ de)
{
#pragma clang fp contract(off)
    for (int64_t e = gtid(); e < ec; e += gstride())
    {
        double point[MB];
        for (int i = 0; i < es; ++i) point[i] = -ge[e * es + i];
        const int64_t k0 = adj_offs ? adj_offs[e] : 0;
        const int64_t k1 = adj_offs ? adj_offs[e + 1] : 0;
        for (int64_t k = k0; k < k1; ++k)
        {
            const int64_t c = adj[k];
            const double* values = W + c * ps * es;
            const double* primary = dp + cross_primary[c] * ps;
            for (int column = 0; column < es; ++column)
            {
                double value = 0.0;
                for (int row = 0; row < ps; ++row) value += values[row * es + column] * primary[row];
                point[column] -= value;
            }
        }
        const double* inv = Cinv + e * es * es;
        for (int row = 0; row < es; ++row)
        {
            double value = 0.0;
            for (int col = 0; col < es; ++col) value += inv[row * es + col] * point[col];
            de[e * es + row] = value;
        }
    }
}

// ---- host side ---------------------------------------------------------------------

// Bump allocator over one device buffer: a sizing pass (base null) then a
// carving pass.
struct Carve
{
    char* base = nullptr;
    size_t off = 0;
    template <typename T>
    T* take(int64_t n)
    {
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += (sizeof(T) * static_cast<size_t>(n > 0 ? n : 1) + 255) & ~size_t(255);
        return p;
    }
};

struct SortSpace
{
    unsigned long long* keys_a;
    unsigned long long* keys_b;
    unsigned* vals_a;
    unsigned* vals_b;
    unsigned* hist;
    unsigned* scratch;
    unsigned* total;
};

SortSpace carve_sort(Carve& c, int64_t n)
{
    SortSpace s{};
    const int64_t hist = radix_hist_words(n);
    s.keys_a = c.take<unsigned long long>(n);
    s.keys_b = c.take<unsigned long long>(n);
    s.vals_a = c.take<unsigned>(n);
    s.vals_b = c.take<unsigned>(n);
    s.hist = c.take<unsigned>(hist);
    s.scratch = c.take<unsigned>(scan_scratch_words(std::max<int64_t>(hist, n)));
    s.total = c.take<unsigned>(4);
    return s;
}

// ---- primary-only residuals (addPrimaryTerms, block_normal_equations.cpp:135-222) ----
//
// A batch of residuals, each linking k distinct primary blocks (k x residual
// size x primary size Jacobians).  The groupings (which terms land in which
// block, and the primary cross blocks in their order of creation) are built on
// the host from the ids; the sums run on the device, per output entry over
// its terms in residual order, starting from the entry's current value, with
// the reference's operation order ((w * a) * b, then +=, no contraction): the
// result is bit for bit that of calling addPrimaryTerms for every residual in
// order after the equations already held.

// Values check and weights: flags the lowest residual with a non-finite
// Jacobian or residual entry (bad_value) or an invalid caller weight
// (bad_weight); weights == null takes the Huber weight of ||r||^2.
__global__ __launch_bounds__(kBlock) void k_pt_check(int64_t count, int k, int rs, int ps,
                                                     const double* __restrict__ jac, const double* __restrict__ r,
                                                     const double* __restrict__ weights, double delta,
                                                     double* __restrict__ w_out, SchurStatus* st)
{
#pragma clang fp contract(off)
    for (int64_t i = gtid(); i < count; i += gstride())
    {
        bool finite = true;
        double s = 0.0;
        for (int q = 0; q < rs; ++q)
        {
            const double v = r[i * rs + q];
            finite &= isfinite(v);
            s += v * v;
        }
        const int64_t per = static_cast<int64_t>(k) * rs * ps;
        for (int64_t q = 0; q < per; ++q) finite &= isfinite(jac[i * per + q]);
        if (!finite) atomicMin(&st->bad_value, static_cast<unsigned long long>(i));
        double w;
        if (weights)
        {
            w = weights[i];
            if (!isfinite(w) || w < 0.0) atomicMin(&st->bad_weight, static_cast<unsigned long long>(i));
        }
        else w = huber_weight(s, delta);
        w_out[i] = w;
    }
}

// Diagonal blocks and gradients: thread per (block, entry) over the blocks
// with terms; terms[offs[b] .. offs[b+1]) are i * k + slot in residual order.
__global__ __launch_bounds__(kBlock) void k_pt_diag(int64_t touched, const int64_t* __restrict__ blocks,
                                                    const int64_t* __restrict__ offs,
                                                    const int64_t* __restrict__ terms, int k, int rs, int ps,
                                                    const double* __restrict__ jac, const double* __restrict__ r,
                                                    const double* __restrict__ w, double* __restrict__ B,
                                                    double* __restrict__ g)
{
#pragma clang fp contract(off)
    const int entries = ps * ps + ps;
    for (int64_t t = gtid(); t < touched * entries; t += gstride())
    {
        const int64_t x = t / entries;
        const int en = static_cast<int>(t % entries);
        const bool is_b = en < ps * ps;
        const int column = is_b ? en / ps : en - ps * ps;
        const int other = is_b ? en % ps : 0;
        const int64_t blk = blocks[x];
        double* out = is_b ? B + blk * ps * ps + column * ps + other : g + blk * ps + column;
        double acc = *out;
        for (int64_t q = offs[x]; q < offs[x + 1]; ++q)
        {
            const int64_t term = terms[q];
            const int64_t i = term / k;
            const double* J = jac + term * rs * ps;
            const double wi = w[i];
            for (int row = 0; row < rs; ++row)
                acc += wi * J[row * ps + column] * (is_b ? J[row * ps + other] : r[i * rs + row]);
        }
        *out = acc;
    }
}

// Primary cross blocks: thread per (cross block, entry) over the blocks with
// terms; a term is (i * k + row slot) * k + column slot.
__global__ __launch_bounds__(kBlock) void k_pt_cross(int64_t touched, const int64_t* __restrict__ index,
                                                     const int64_t* __restrict__ offs,
                                                     const int64_t* __restrict__ terms, int k, int rs, int ps,
                                                     const double* __restrict__ jac, const double* __restrict__ w,
                                                     double* __restrict__ values)
{
#pragma clang fp contract(off)
    const int entries = ps * ps;
    for (int64_t t = gtid(); t < touched * entries; t += gstride())
    {
        const int64_t x = t / entries;
        const int en = static_cast<int>(t % entries);
        const int row = en / ps, column = en % ps;
        double* out = values + index[x] * entries + en;
        double acc = *out;
        for (int64_t q = offs[x]; q < offs[x + 1]; ++q)
        {
            const int64_t term = terms[q];
            const int64_t cs = term % k;
            const int64_t rsl = (term / k) % k;
            const int64_t i = term / k / k;
            const double* Jr = jac + (i * k + rsl) * rs * ps;
            const double* Jc = jac + (i * k + cs) * rs * ps;
            const double wi = w[i];
            for (int rr = 0; rr < rs; ++rr) acc += wi * Jr[rr * ps + row] * Jc[rr * ps + column];
        }
        *out = acc;
    }
}

} // namespace pm

using namespace pm;

struct pm_schur_workspace
{
    hipStream_t stream = nullptr;
    bool bound = false;
    int device = -1;
    SchurStatus* host_status = nullptr;  // pinned
    DeviceBuffer status;

    // normal equations assembled from residuals
    DeviceBuffer ne_topology;  // residual grouping (reused while the linkage is unchanged)
    DeviceBuffer ne_values;    // B, C, g, W, cross ids, weights
    DeviceBuffer ne_sort;
    int64_t ne_sig[6] = {-1, -1, -1, -1, -1, -1};  // count, pc, ec, ps, es, rs
    int64_t ne_runs = 0;
    int64_t ne_builds = 0;
    int64_t* ne_primary_copy = nullptr;
    int64_t* ne_eliminated_copy = nullptr;
    int64_t* ne_offs_p = nullptr;
    int64_t* ne_offs_e = nullptr;
    unsigned* ne_perm_p = nullptr;
    unsigned* ne_perm_e = nullptr;
    unsigned* ne_perm_pe = nullptr;
    int64_t* ne_run_start = nullptr;
    unsigned* ne_run_cross = nullptr;
    int64_t* ne_cross_primary = nullptr;
    int64_t* ne_cross_eliminated = nullptr;
    unsigned* ne_res_cross = nullptr;  // per residual: its cross block if the only residual of its pair, else ~0u

    // primary-only residuals (addPrimaryTerms): groupings, weights, primary
    // cross blocks
    DeviceBuffer pt_meta;
    DeviceBuffer pt_cross;

    // Schur pattern (reused while the cross-block topology is unchanged)
    DeviceBuffer pattern;
    DeviceBuffer pattern_sort;
    int64_t sig[6] = {-1, -1, -1, -1, -1, -1};  // pc, ec, ps, es, nc, npc
    int64_t pattern_builds = 0;
    int64_t slots = 0, values_count = 0, terms = 0, entries = 0;
    int pb = 0, eb = 0;
    int64_t* cp_copy = nullptr;   // cross_primary
    int64_t* ce_copy = nullptr;   // cross_eliminated
    int64_t* pr_copy = nullptr;   // primary cross rows
    int64_t* pcc_copy = nullptr;  // primary cross cols
    int64_t* adj_offs = nullptr;
    unsigned* adj = nullptr;
    int64_t* prim_offs = nullptr;
    unsigned* prim_cross = nullptr;
    unsigned* cross_pos = nullptr;   // position of each cross block in prim_cross
    unsigned long long* slot_key = nullptr;
    int64_t* slot_start = nullptr;
    unsigned* term_e = nullptr;
    unsigned* term_left = nullptr;
    unsigned* term_right = nullptr;
    unsigned* seg_offs = nullptr;    // segmented assembly: first segment of each slot (slots + 1)
    unsigned* slot_terms = nullptr;  // terms of each slot (markers excluded)
    DeviceBuffer segment_slot;       // slot of each segment
    DeviceBuffer segment_terms;      // (left, right) pairs of each segment, kSegTerms each, padded
    DeviceBuffer slot_mirror;        // slot of the mirrored block of each slot
    int64_t segments = 0;
    int64_t* bro = nullptr;
    int64_t* ro = nullptr;
    int32_t* ro32 = nullptr;
    int64_t* ci = nullptr;
    int32_t* ci32 = nullptr;
    int64_t* diag_slot = nullptr;

    // FP32 copies of the reduced system for the mixed-precision seed solve
    DeviceBuffer seed;

    // per-solve numerics
    DeviceBuffer numerics;
    double* Bd = nullptr;
    double* Cd = nullptr;
    double* Cinv = nullptr;
    double* y = nullptr;
    double* rhs = nullptr;
    double* TW = nullptr;
    double* values = nullptr;
    double* pinv = nullptr;
    double* crv = nullptr;  // per (cross block, row) products of the reduced rhs, then of the back-substitution

    int bind(hipStream_t s)
    {
        int dev = 0;
        PM_HIP(hipGetDevice(&dev));
        if (!bound)
        {
            stream = s;
            device = dev;
            bound = true;
            PM_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_status), sizeof(SchurStatus), hipHostMallocDefault));
            PM_TRY(status.reserve(sizeof(SchurStatus)));
            return PM_OK;
        }
        if (s != stream)
            return fail(PM_LOGIC_ERROR, "SchurComplementSolverWorkspace is bound to another stream");
        if (dev != device)
            return fail(PM_LOGIC_ERROR, "SchurComplementSolverWorkspace is bound to another device");
        return PM_OK;
    }

    SchurStatus* dstatus() const { return status.as<SchurStatus>(); }

    int read_status(hipStream_t s)
    {
        PM_HIP(hipMemcpyAsync(host_status, dstatus(), sizeof(SchurStatus), hipMemcpyDeviceToHost, s));
        PM_HIP(hipStreamSynchronize(s));
        return PM_OK;
    }

    ~pm_schur_workspace()
    {
        if (bound) (void)hipStreamSynchronize(stream);
        if (host_status) (void)hipHostFree(host_status);
        status.release();
        ne_topology.release();
        ne_values.release();
        ne_sort.release();
        pattern.release();
        pattern_sort.release();
        segment_slot.release();
        segment_terms.release();
        slot_mirror.release();
        numerics.release();
        pt_meta.release();
        pt_cross.release();
    }
};

namespace pm
{

int sort_by_key(SortSpace& sp, int64_t n, int bits, hipStream_t stream, unsigned long long** keys, unsigned** vals)
{
    bool in_b = false;
    PM_TRY(radix_sort_pairs(sp.keys_a, sp.vals_a, sp.keys_b, sp.vals_b, n, bits, sp.hist, sp.scratch, stream, &in_b));
    *keys = in_b ? sp.keys_b : sp.keys_a;
    *vals = in_b ? sp.vals_b : sp.vals_a;
    return PM_OK;
}

int check_schur_sizes(int64_t pc, int64_t ec, int64_t ps, int64_t es)
{
    if (pc < 0 || ec < 0 || ps < 0 || es < 0)
        return fail(PM_INVALID_ARGUMENT, "BlockNormalEquations: block counts and sizes must be non-negative");
    if ((pc > 0 && ps <= 0) || (ec > 0 && es <= 0))
        return fail(PM_INVALID_ARGUMENT, "BlockNormalEquations: block sizes must be positive");
    if (ps > kMaxSchurBlock || es > kMaxSchurBlock)
        return fail(PM_INVALID_ARGUMENT, "the HIP Schur path supports block sizes up to %d", kMaxSchurBlock);
    if (pc >= (int64_t(1) << 31) || ec >= (int64_t(1) << 31))
        return fail(PM_OVERFLOW, "Schur block counts must be below 2^31");
    return PM_OK;
}

// Device pattern build from the cross-block topology (cached copies in w).
int build_pattern(pm_schur_workspace* w, const pm_normal_equations* eq, hipStream_t stream)
{
    const int64_t pc = eq->primary_count, ec = eq->eliminated_count, nc = eq->cross_count;
    const int64_t npc = eq->primary_cross_count;
    const int ps = static_cast<int>(eq->primary_size);
    const int pb = std::max(1, bits_for(pc)), eb = std::max(1, bits_for(ec));
    w->pb = pb;
    w->eb = eb;
    SchurStatus* st = w->dstatus();

    // ids must address existing blocks before they index anything
    hipLaunchKernelGGL(k_validate_topology, dim3(grid_for(std::max<int64_t>(std::max(nc, npc), 1))), dim3(kBlock), 0,
                       stream, nc, npc, pc, ec, eq->cross_primary, eq->cross_eliminated, eq->primary_cross_rows,
                       eq->primary_cross_cols, st);
    PM_HIP(hipGetLastError());
    PM_TRY(w->read_status(stream));
    if (w->host_status->bad_primary != ~0ull)
        return fail(PM_OUT_OF_RANGE, "solveDampedSchurComplement: cross block %llu has an out-of-range primary block",
                    w->host_status->bad_primary);
    if (w->host_status->bad_eliminated != ~0ull)
        return fail(PM_OUT_OF_RANGE,
                    "solveDampedSchurComplement: cross block %llu has an out-of-range eliminated block",
                    w->host_status->bad_eliminated);
    if (w->host_status->bad_value != ~0ull)
        return fail(PM_OUT_OF_RANGE,
                    "solveDampedSchurComplement: primary cross block %llu needs row < column < primary count",
                    w->host_status->bad_value);

    // stage 1: adjacency by eliminated and by primary, pair counts (the sort
    // space also serves the scan over ec + 1 pair counts)
    const int64_t n1 = std::max<int64_t>(std::max<int64_t>(nc, ec + 1), 1);
    Carve s1z;
    carve_sort(s1z, n1);
    const size_t stage1 = s1z.off;
    Carve pz;
    pz.take<int64_t>(nc);
    pz.take<int64_t>(nc);
    pz.take<int64_t>(npc);
    pz.take<int64_t>(npc);
    pz.take<int64_t>(ec + 1);
    pz.take<unsigned>(nc);
    pz.take<int64_t>(pc + 1);
    pz.take<unsigned>(nc);
    pz.take<unsigned>(ec + 1);
    const size_t head_bytes = pz.off;
    PM_TRY(w->pattern_sort.reserve(std::max(stage1, head_bytes)));
    // the first part of the pattern buffer holds the topology copies and
    // adjacency; the rest is sized once the term count is known
    Carve sc;
    sc.base = w->pattern_sort.as<char>();
    SortSpace sp = carve_sort(sc, n1);

    // temporary holder for the head arrays until the final pattern buffer exists
    DeviceBuffer head;
    struct Release
    {
        DeviceBuffer& b;
        ~Release() { b.release(); }
    } release_head{head};
    PM_TRY(head.reserve(head_bytes));
    Carve hc;
    hc.base = head.as<char>();
    int64_t* cp = hc.take<int64_t>(nc);
    int64_t* ce = hc.take<int64_t>(nc);
    int64_t* pr = hc.take<int64_t>(npc);
    int64_t* pcl = hc.take<int64_t>(npc);
    int64_t* adj_offs = hc.take<int64_t>(ec + 1);
    unsigned* adj = hc.take<unsigned>(nc);
    int64_t* prim_offs = hc.take<int64_t>(pc + 1);
    unsigned* prim_cross = hc.take<unsigned>(nc);
    const size_t prefix_bytes = hc.off;  // same carving as the front of the final pattern buffer
    unsigned* pair_offs = hc.take<unsigned>(ec + 1);

    const int gc = grid_for(n1);
    if (nc > 0)
    {
        hipLaunchKernelGGL(k_copy_i64, dim3(gc), dim3(kBlock), 0, stream, eq->cross_primary, nc, cp);
        hipLaunchKernelGGL(k_copy_i64, dim3(gc), dim3(kBlock), 0, stream, eq->cross_eliminated, nc, ce);
    }
    if (npc > 0)
    {
        hipLaunchKernelGGL(k_copy_i64, dim3(grid_for(npc)), dim3(kBlock), 0, stream, eq->primary_cross_rows, npc, pr);
        hipLaunchKernelGGL(k_copy_i64, dim3(grid_for(npc)), dim3(kBlock), 0, stream, eq->primary_cross_cols, npc, pcl);
    }
    // keys: eliminated id (adjacency, creation order kept by stability)
    unsigned long long* sk = nullptr;
    unsigned* sv = nullptr;
    if (nc > 0)
    {
        hipLaunchKernelGGL(k_cross_keys, dim3(gc), dim3(kBlock), 0, stream, nc, cp, ce, 0, sp.keys_a, sp.vals_a);
        PM_TRY(sort_by_key(sp, nc, eb, stream, &sk, &sv));
        hipLaunchKernelGGL(k_group_offsets, dim3(grid_for(ec + 1)), dim3(kBlock), 0, stream, sk, nc, 0, ec, adj_offs);
        PM_HIP(hipMemcpyAsync(adj, sv, sizeof(unsigned) * nc, hipMemcpyDeviceToDevice, stream));
        // (primary, eliminated) order for the reduced right-hand side
        hipLaunchKernelGGL(k_cross_keys, dim3(gc), dim3(kBlock), 0, stream, nc, cp, ce, eb, sp.keys_a, sp.vals_a);
        PM_TRY(sort_by_key(sp, nc, pb + eb, stream, &sk, &sv));
        hipLaunchKernelGGL(k_group_offsets, dim3(grid_for(pc + 1)), dim3(kBlock), 0, stream, sk, nc, eb, pc,
                           prim_offs);
        PM_HIP(hipMemcpyAsync(prim_cross, sv, sizeof(unsigned) * nc, hipMemcpyDeviceToDevice, stream));
    }
    else
    {
        hipLaunchKernelGGL(k_fill_i64, dim3(grid_for(ec + 1)), dim3(kBlock), 0, stream, adj_offs, ec + 1, 0);
        hipLaunchKernelGGL(k_fill_i64, dim3(grid_for(pc + 1)), dim3(kBlock), 0, stream, prim_offs, pc + 1, 0);
    }
    hipLaunchKernelGGL(k_pair_counts, dim3(grid_for(ec)), dim3(kBlock), 0, stream, ec, adj_offs, pair_offs);
    // pair counts may exceed u32 in total only past 2^32 terms; checked below in 64 bits
    hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(kBlock), 0, stream, ec, adj_offs, st);
    PM_TRY(scan_u32(pair_offs, pair_offs, ec + 1, sp.scratch, nullptr, stream));
    PM_HIP(hipGetLastError());
    PM_TRY(w->read_status(stream));
    const int64_t terms = static_cast<int64_t>(w->host_status->count0);
    const int64_t entries = terms + pc + 2 * npc;
    if (entries >= (int64_t(1) << 31))
        return fail(PM_OVERFLOW, "Schur pattern has %lld terms; the HIP path supports fewer than 2^31",
                    static_cast<long long>(terms));

    // stage 2: sort all block keys
    Carve s2z;
    carve_sort(s2z, std::max<int64_t>(entries, 1));
    s2z.take<unsigned>(entries);  // heads
    s2z.take<unsigned>(entries);  // slot ids
    PM_TRY(w->pattern_sort.reserve(std::max(s2z.off, stage1)));
    Carve sc2;
    sc2.base = w->pattern_sort.as<char>();
    SortSpace sp2 = carve_sort(sc2, std::max<int64_t>(entries, 1));
    unsigned* heads = sc2.take<unsigned>(entries);
    unsigned* slot_id = sc2.take<unsigned>(entries);
    hipLaunchKernelGGL(k_pattern_keys, dim3(grid_for(std::max(ec, std::max(pc, npc)))), dim3(kBlock), 0, stream, ec,
                       pc, npc, pb, adj_offs, adj, pair_offs, cp, pr, pcl, terms, sp2.keys_a, sp2.vals_a);
    PM_HIP(hipGetLastError());
    PM_TRY(sort_by_key(sp2, entries, 2 * pb, stream, &sk, &sv));
    const int ge = grid_for(entries);
    hipLaunchKernelGGL(k_key_heads, dim3(ge), dim3(kBlock), 0, stream, sk, entries, heads);
    PM_TRY(scan_u32(heads, slot_id, entries, sp2.scratch, sp2.total, stream));
    unsigned slots32 = 0;
    PM_HIP(hipMemcpyAsync(&slots32, sp2.total, sizeof(unsigned), hipMemcpyDeviceToHost, stream));
    PM_HIP(hipStreamSynchronize(stream));
    const int64_t slots = slots32;
    const int64_t values = slots * ps * ps;
    const int64_t dim = pc * ps;
    if (values >= (int64_t(1) << 31)) return fail(PM_OVERFLOW, "reduced Schur system has too many values");

    // final pattern buffer
    Carve fz;
    fz.take<int64_t>(nc);
    fz.take<int64_t>(nc);
    fz.take<int64_t>(npc);
    fz.take<int64_t>(npc);
    fz.take<int64_t>(ec + 1);
    fz.take<unsigned>(nc);
    fz.take<int64_t>(pc + 1);
    fz.take<unsigned>(nc);
    fz.take<unsigned long long>(slots);
    fz.take<int64_t>(slots + 1);
    fz.take<unsigned>(entries);
    fz.take<unsigned>(entries);
    fz.take<unsigned>(entries);
    fz.take<int64_t>(pc + 1);
    fz.take<int64_t>(dim + 1);
    fz.take<int32_t>(dim + 1);
    fz.take<int64_t>(values);
    fz.take<int32_t>(values);
    fz.take<int64_t>(pc);
    fz.take<unsigned>(slots + 1);
    fz.take<unsigned>(slots + 1);
    fz.take<unsigned>(slots);
    fz.take<unsigned>(nc);
    PM_TRY(w->pattern.reserve(fz.off));
    Carve f;
    f.base = w->pattern.as<char>();
    w->cp_copy = f.take<int64_t>(nc);
    w->ce_copy = f.take<int64_t>(nc);
    w->pr_copy = f.take<int64_t>(npc);
    w->pcc_copy = f.take<int64_t>(npc);
    w->adj_offs = f.take<int64_t>(ec + 1);
    w->adj = f.take<unsigned>(nc);
    w->prim_offs = f.take<int64_t>(pc + 1);
    w->prim_cross = f.take<unsigned>(nc);
    w->slot_key = f.take<unsigned long long>(slots);
    w->slot_start = f.take<int64_t>(slots + 1);
    w->term_e = f.take<unsigned>(entries);
    w->term_left = f.take<unsigned>(entries);
    w->term_right = f.take<unsigned>(entries);
    w->bro = f.take<int64_t>(pc + 1);
    w->ro = f.take<int64_t>(dim + 1);
    w->ro32 = f.take<int32_t>(dim + 1);
    w->ci = f.take<int64_t>(values);
    w->ci32 = f.take<int32_t>(values);
    w->diag_slot = f.take<int64_t>(pc);
    unsigned* seg_count = f.take<unsigned>(slots + 1);
    w->seg_offs = f.take<unsigned>(slots + 1);
    w->slot_terms = f.take<unsigned>(slots);
    w->cross_pos = f.take<unsigned>(nc);
    // the head block was carved in the same order and sizes as the front of
    // the pattern buffer, so one copy moves cp, ce, pr, pcl, adjacency and
    // the primary lists
    PM_HIP(hipMemcpyAsync(w->cp_copy, cp, prefix_bytes, hipMemcpyDeviceToDevice, stream));
    if (nc > 0)
        hipLaunchKernelGGL(k_invert_order, dim3(grid_for(nc)), dim3(kBlock), 0, stream, nc,
                           static_cast<const unsigned*>(w->prim_cross), w->cross_pos);
    hipLaunchKernelGGL(k_slots, dim3(ge), dim3(kBlock), 0, stream, sk, heads, slot_id, entries, w->slot_key,
                       w->slot_start);
    hipLaunchKernelGGL(k_decode_terms, dim3(ge), dim3(kBlock), 0, stream, sv, entries, terms, pc, ec, pair_offs,
                       adj_offs, adj, w->term_e, w->term_left, w->term_right);
    hipLaunchKernelGGL(k_group_offsets, dim3(grid_for(pc + 1)), dim3(kBlock), 0, stream, w->slot_key, slots, pb, pc,
                       w->bro);
    // segments of the slots' term runs (segmented assembly)
    hipLaunchKernelGGL(k_slot_segments, dim3(grid_for(slots + 1)), dim3(kBlock), 0, stream, slots, pb, w->slot_key,
                       w->slot_start, entries, w->term_e, w->slot_terms, seg_count);
    PM_TRY(w->slot_mirror.reserve(sizeof(unsigned) * std::max<int64_t>(slots, 1)));
    hipLaunchKernelGGL(k_slot_mirror, dim3(grid_for(slots)), dim3(kBlock), 0, stream, slots, pb, w->slot_key,
                       w->slot_mirror.as<unsigned>());
    PM_TRY(scan_u32(seg_count, w->seg_offs, slots, sp2.scratch, w->seg_offs + slots, stream));
    unsigned nseg32 = 0;
    PM_HIP(hipMemcpyAsync(&nseg32, w->seg_offs + slots, sizeof(unsigned), hipMemcpyDeviceToHost, stream));
    PM_HIP(hipStreamSynchronize(stream));
    w->segments = nseg32;
    PM_TRY(w->segment_slot.reserve(sizeof(unsigned) * std::max<int64_t>(w->segments, 1)));
    hipLaunchKernelGGL(k_segment_slots, dim3(grid_for(slots)), dim3(kBlock), 0, stream, slots, w->seg_offs,
                       w->segment_slot.as<unsigned>());
    PM_TRY(w->segment_terms.reserve(sizeof(uint2) * kSegTerms * std::max<int64_t>(w->segments, 1)));
    if (w->segments > 0)
        hipLaunchKernelGGL(k_segment_terms, dim3(grid_for(w->segments * kSegTerms)), dim3(kBlock), 0, stream,
                           w->segments, static_cast<const unsigned*>(w->segment_slot.as<unsigned>()),
                           static_cast<const unsigned*>(w->seg_offs), static_cast<const int64_t*>(w->slot_start),
                           static_cast<const unsigned*>(w->slot_terms), static_cast<const unsigned*>(w->term_left),
                           static_cast<const unsigned*>(w->term_right), w->segment_terms.as<uint2>());
    hipLaunchKernelGGL(k_scalar_rows, dim3(grid_for(dim + 1)), dim3(kBlock), 0, stream, pc, ps, w->bro, w->ro,
                       w->ro32);
    hipLaunchKernelGGL(k_scalar_cols, dim3(grid_for(values)), dim3(kBlock), 0, stream, slots, ps, pb, w->slot_key,
                       w->bro, w->ro, w->ci, w->ci32);
    hipLaunchKernelGGL(k_diag_slots, dim3(grid_for(pc)), dim3(kBlock), 0, stream, pc, pb, slots, w->slot_key,
                       w->diag_slot);
    PM_HIP(hipGetLastError());
    PM_HIP(hipStreamSynchronize(stream));
    w->slots = slots;
    w->values_count = values;
    w->terms = terms;
    w->entries = entries;
    ++w->pattern_builds;
    return PM_OK;
}

} // namespace pm

namespace pm
{

int residual_pass(pm_schur_workspace* w, const pm_residual_batch* rb, int64_t pc, int64_t ec, int ps, int es,
                  double* weights_out, hipStream_t stream, bool sort_needed)
{
    const int64_t n = rb->count;
    const int pbits = std::max(1, bits_for(pc)), ebits = std::max(1, bits_for(ec));
    Carve sizing;
    carve_sort(sizing, n);
    sizing.take<unsigned long long>(n);
    sizing.take<unsigned long long>(n);
    sizing.take<unsigned>(n);
    sizing.take<unsigned>(n);
    sizing.take<unsigned>(n);
    sizing.take<unsigned>(n);
    sizing.take<unsigned>(n);
    PM_TRY(w->ne_sort.reserve(sizing.off));
    Carve c;
    c.base = w->ne_sort.as<char>();
    SortSpace sp = carve_sort(c, n);
    unsigned long long* key_e = c.take<unsigned long long>(n);
    unsigned long long* key_pe = c.take<unsigned long long>(n);
    unsigned* val_e = c.take<unsigned>(n);
    unsigned* val_pe = c.take<unsigned>(n);
    unsigned* heads = c.take<unsigned>(n);
    unsigned* run_id = c.take<unsigned>(n);
    unsigned* first = c.take<unsigned>(n);

    SchurStatus* st = w->dstatus();
    const int g = grid_for(n);
    if (n > 0)
    {
        const bool lm = rb->residual_size == 2 && ps == 6 && es == 3;
        hipLaunchKernelGGL((lm ? k_residual_keys<2, 6, 3> : k_residual_keys<>), dim3(g), dim3(kBlock), 0, stream, n,
                           pc, ec,
                           static_cast<int>(rb->residual_size), ps, es, ebits, rb->primary, rb->eliminated,
                           rb->primary_jacobians, rb->eliminated_jacobians, rb->residuals, rb->weights,
                           rb->huber_delta, weights_out, sp.keys_a, key_e, key_pe, sp.vals_a, val_e, val_pe, st);
        PM_HIP(hipGetLastError());
    }
    PM_TRY(w->read_status(stream));
    const SchurStatus& hs = *w->host_status;
    if (hs.bad_primary != ~0ull)
        return fail(PM_OUT_OF_RANGE, "BlockNormalEquations: primary block index out of range (residual %llu)",
                    hs.bad_primary);
    if (hs.bad_eliminated != ~0ull)
        return fail(PM_OUT_OF_RANGE,
                    "BlockNormalEquations::addResidualBlocks: eliminated block out of range (residual %llu)",
                    hs.bad_eliminated);
    if (hs.bad_value != ~0ull)
        return fail(PM_INVALID_ARGUMENT,
                    "BlockNormalEquations::addResidualBlocks: Jacobian and residual values must be finite "
                    "(residual %llu)",
                    hs.bad_value);
    if (hs.bad_weight != ~0ull)
        return fail(PM_INVALID_ARGUMENT,
                    "BlockNormalEquations::addResidualBlocks: weight must be finite and non-negative (residual %llu)",
                    hs.bad_weight);
    if (!sort_needed) return PM_OK;

    // persistent grouping
    Carve tz;
    tz.take<int64_t>(n);
    tz.take<int64_t>(n);
    tz.take<int64_t>(pc + 1);
    tz.take<int64_t>(ec + 1);
    tz.take<unsigned>(n);
    tz.take<unsigned>(n);
    tz.take<unsigned>(n);
    tz.take<int64_t>(n + 1);
    tz.take<unsigned>(n);
    tz.take<int64_t>(n);
    tz.take<int64_t>(n);
    tz.take<unsigned>(n);
    PM_TRY(w->ne_topology.reserve(tz.off));
    Carve t;
    t.base = w->ne_topology.as<char>();
    w->ne_primary_copy = t.take<int64_t>(n);
    w->ne_eliminated_copy = t.take<int64_t>(n);
    w->ne_offs_p = t.take<int64_t>(pc + 1);
    w->ne_offs_e = t.take<int64_t>(ec + 1);
    w->ne_perm_p = t.take<unsigned>(n);
    w->ne_perm_e = t.take<unsigned>(n);
    w->ne_perm_pe = t.take<unsigned>(n);
    w->ne_run_start = t.take<int64_t>(n + 1);
    w->ne_run_cross = t.take<unsigned>(n);
    w->ne_cross_primary = t.take<int64_t>(n);
    w->ne_cross_eliminated = t.take<int64_t>(n);
    w->ne_res_cross = t.take<unsigned>(n);
    if (n == 0)
    {
        hipLaunchKernelGGL(k_fill_i64, dim3(grid_for(pc + 1)), dim3(kBlock), 0, stream, w->ne_offs_p, pc + 1, 0);
        hipLaunchKernelGGL(k_fill_i64, dim3(grid_for(ec + 1)), dim3(kBlock), 0, stream, w->ne_offs_e, ec + 1, 0);
        PM_HIP(hipGetLastError());
        w->ne_runs = 0;
        ++w->ne_builds;
        return PM_OK;
    }
    hipLaunchKernelGGL(k_copy_i64, dim3(g), dim3(kBlock), 0, stream, rb->primary, n, w->ne_primary_copy);
    hipLaunchKernelGGL(k_copy_i64, dim3(g), dim3(kBlock), 0, stream, rb->eliminated, n, w->ne_eliminated_copy);
    unsigned long long* sk;
    unsigned* sv;
    // by primary
    PM_TRY(sort_by_key(sp, n, pbits, stream, &sk, &sv));
    hipLaunchKernelGGL(k_group_offsets, dim3(grid_for(pc + 1)), dim3(kBlock), 0, stream, sk, n, 0, pc, w->ne_offs_p);
    PM_HIP(hipMemcpyAsync(w->ne_perm_p, sv, sizeof(unsigned) * n, hipMemcpyDeviceToDevice, stream));
    // by eliminated
    PM_HIP(hipMemcpyAsync(sp.keys_a, key_e, sizeof(unsigned long long) * n, hipMemcpyDeviceToDevice, stream));
    PM_HIP(hipMemcpyAsync(sp.vals_a, val_e, sizeof(unsigned) * n, hipMemcpyDeviceToDevice, stream));
    PM_TRY(sort_by_key(sp, n, ebits, stream, &sk, &sv));
    hipLaunchKernelGGL(k_group_offsets, dim3(grid_for(ec + 1)), dim3(kBlock), 0, stream, sk, n, 0, ec, w->ne_offs_e);
    PM_HIP(hipMemcpyAsync(w->ne_perm_e, sv, sizeof(unsigned) * n, hipMemcpyDeviceToDevice, stream));
    // by (primary, eliminated): runs are the cross blocks
    PM_HIP(hipMemcpyAsync(sp.keys_a, key_pe, sizeof(unsigned long long) * n, hipMemcpyDeviceToDevice, stream));
    PM_HIP(hipMemcpyAsync(sp.vals_a, val_pe, sizeof(unsigned) * n, hipMemcpyDeviceToDevice, stream));
    PM_TRY(sort_by_key(sp, n, pbits + ebits, stream, &sk, &sv));
    PM_HIP(hipMemcpyAsync(w->ne_perm_pe, sv, sizeof(unsigned) * n, hipMemcpyDeviceToDevice, stream));
    hipLaunchKernelGGL(k_zero_u32, dim3(g), dim3(kBlock), 0, stream, first, n);
    hipLaunchKernelGGL(k_pair_heads, dim3(g), dim3(kBlock), 0, stream, sk, sv, n, heads, first);
    PM_TRY(scan_u32(heads, run_id, n, sp.scratch, sp.total, stream));
    PM_TRY(scan_u32(first, first, n, sp.scratch, nullptr, stream));
    hipLaunchKernelGGL(k_pair_runs, dim3(g), dim3(kBlock), 0, stream, sk, sv, heads, run_id, first, n, ebits,
                       w->ne_run_start, w->ne_run_cross, w->ne_cross_primary, w->ne_cross_eliminated);
    PM_HIP(hipGetLastError());
    unsigned runs = 0;
    PM_HIP(hipMemcpyAsync(&runs, sp.total, sizeof(unsigned), hipMemcpyDeviceToHost, stream));
    PM_HIP(hipStreamSynchronize(stream));
    if (runs > 0)
        hipLaunchKernelGGL(k_residual_cross, dim3(grid_for(runs)), dim3(kBlock), 0, stream,
                           static_cast<int64_t>(runs), n, static_cast<const int64_t*>(w->ne_run_start),
                           static_cast<const unsigned*>(w->ne_perm_pe), static_cast<const unsigned*>(w->ne_run_cross),
                           w->ne_res_cross);
    PM_HIP(hipGetLastError());
    w->ne_runs = runs;
    ++w->ne_builds;
    return PM_OK;
}

int check_normal_equations(const pm_normal_equations* eq)
{
    if (!eq) return fail(PM_INVALID_ARGUMENT, "solveDampedSchurComplement: null normal equations");
    PM_TRY(check_schur_sizes(eq->primary_count, eq->eliminated_count, eq->primary_size, eq->eliminated_size));
    if (eq->cross_count < 0 || eq->primary_cross_count < 0)
        return fail(PM_INVALID_ARGUMENT, "solveDampedSchurComplement: negative cross-block count");
    const int64_t pd = eq->primary_count * eq->primary_size, ed = eq->eliminated_count * eq->eliminated_size;
    if ((pd > 0 && (!eq->primary_diagonal || !eq->primary_gradient)) ||
        (ed > 0 && (!eq->eliminated_diagonal || !eq->eliminated_gradient)) ||
        (eq->cross_count > 0 && (!eq->cross_primary || !eq->cross_eliminated || !eq->cross_values)) ||
        (eq->primary_cross_count > 0 &&
         (!eq->primary_cross_rows || !eq->primary_cross_cols || !eq->primary_cross_values)))
        return fail(PM_INVALID_ARGUMENT, "solveDampedSchurComplement: null normal-equation arrays");
    return PM_OK;
}

// Pattern (built or reused), damping, eliminated inverses, reduced rhs, Schur
// values and preconditioner inverses.  report->status 1 / 2 flags a block
// that is not positive definite (the reference reports it the same way,
// block_schur.cpp:140-146, 240-246).
int schur_assemble(pm_schur_workspace* w, const pm_normal_equations* eq, double damping, double minimum_diagonal,
                   pm_schur_report* report, hipStream_t stream)
{
    const int64_t pc = eq->primary_count, ec = eq->eliminated_count;
    const int ps = static_cast<int>(eq->primary_size), es = static_cast<int>(eq->eliminated_size);
    const int64_t pd = pc * ps, nc = eq->cross_count, npc = eq->primary_cross_count;
    PM_TRY(w->bind(stream));
    SchurStatus* st = w->dstatus();
    hipLaunchKernelGGL(k_schur_status_init, dim3(1), dim3(1), 0, stream, st);

    // pattern reuse: compare the cross-block topology with the cached copy
    const int64_t sig[6] = {pc, ec, ps, es, nc, npc};
    bool reuse = w->pattern_builds > 0 && std::equal(sig, sig + 6, w->sig);
    if (reuse)
    {
        if (nc > 0)
        {
            hipLaunchKernelGGL(k_count_mismatch, dim3(grid_for(nc)), dim3(kBlock), 0, stream, eq->cross_primary,
                               w->cp_copy, nc, st);
            hipLaunchKernelGGL(k_count_mismatch, dim3(grid_for(nc)), dim3(kBlock), 0, stream, eq->cross_eliminated,
                               w->ce_copy, nc, st);
        }
        if (npc > 0)
        {
            hipLaunchKernelGGL(k_count_mismatch, dim3(grid_for(npc)), dim3(kBlock), 0, stream,
                               eq->primary_cross_rows, w->pr_copy, npc, st);
            hipLaunchKernelGGL(k_count_mismatch, dim3(grid_for(npc)), dim3(kBlock), 0, stream,
                               eq->primary_cross_cols, w->pcc_copy, npc, st);
        }
        PM_HIP(hipGetLastError());
        PM_TRY(w->read_status(stream));
        reuse = w->host_status->mismatch == 0;
    }
    if (!reuse && pd > 0)
    {
        hipLaunchKernelGGL(k_schur_status_init, dim3(1), dim3(1), 0, stream, st);
        PM_TRY(build_pattern(w, eq, stream));
        std::copy(sig, sig + 6, w->sig);
    }
    report->pattern_reused = reuse ? 1 : 0;
    report->pattern_build_count = w->pattern_builds;

    // numerics
    Carve nz;
    nz.take<double>(pc * ps * ps);
    nz.take<double>(ec * es * es);
    nz.take<double>(ec * es * es);
    nz.take<double>(ec * es);
    nz.take<double>(pd);
    nz.take<double>(nc * ps * es);
    nz.take<double>(pd > 0 ? w->values_count : 0);
    nz.take<double>(pc * ps * ps);
    nz.take<double>(nc * (ps > es ? ps : es));
    nz.take<double>(w->segments * ps * ps);
    PM_TRY(w->numerics.reserve(nz.off));
    Carve n;
    n.base = w->numerics.as<char>();
    w->Bd = n.take<double>(pc * ps * ps);
    w->Cd = n.take<double>(ec * es * es);
    w->Cinv = n.take<double>(ec * es * es);
    w->y = n.take<double>(ec * es);
    w->rhs = n.take<double>(pd);
    w->TW = n.take<double>(nc * ps * es);
    w->values = n.take<double>(pd > 0 ? w->values_count : 0);
    w->pinv = n.take<double>(pc * ps * ps);
    w->crv = n.take<double>(nc * (ps > es ? ps : es));
    double* seg_partial = n.take<double>(w->segments * ps * ps);

    hipLaunchKernelGGL(k_schur_status_init, dim3(1), dim3(1), 0, stream, st);
    if (pc > 0)
        hipLaunchKernelGGL(k_damp, dim3(grid_for(pc * ps * ps)), dim3(kBlock), 0, stream, eq->primary_diagonal, pc,
                           ps, damping, minimum_diagonal, w->Bd);
    if (ec > 0)
    {
        hipLaunchKernelGGL(k_damp, dim3(grid_for(ec * es * es)), dim3(kBlock), 0, stream, eq->eliminated_diagonal, ec,
                           es, damping, minimum_diagonal, w->Cd);
        const auto inverse_kernel = es == 3                  ? k_eliminated_inverse<3, 3>
                                    : es <= kSmallSchurBlock ? k_eliminated_inverse<kSmallSchurBlock>
                                                             : k_eliminated_inverse<kMaxSchurBlock>;
        hipLaunchKernelGGL(inverse_kernel, dim3(grid_for(ec)), dim3(kBlock), 0, stream, ec, es, w->Cd,
                           eq->eliminated_gradient, w->Cinv, w->y, st);
    }
    if (pd > 0)
    {
        // 6 x 3 blocks: W C^-1 and W C^-1 g in one pass over W
        const bool fused = ps == 6 && es == 3;
        if (nc > 0 && fused)
            hipLaunchKernelGGL((k_transform_rhs<6, 3>), dim3(grid_for(nc * ps)), dim3(kBlock), 0, stream, nc,
                               eq->cross_eliminated, static_cast<const unsigned*>(w->cross_pos), w->Cinv, w->y,
                               eq->cross_values, w->TW, w->crv);
        else if (nc > 0)
            hipLaunchKernelGGL((k_cross_rhs<>), dim3(grid_for(nc * ps)), dim3(kBlock), 0, stream, nc, ps, es,
                               w->prim_cross, eq->cross_eliminated, eq->cross_values, w->y, w->crv);
        hipLaunchKernelGGL(k_reduced_rhs, dim3(static_cast<unsigned>(ceil_div(pd, kWave))), dim3(kWave), 0, stream, pc,
                           ps, eq->primary_gradient, w->prim_offs, static_cast<const double*>(w->crv), w->rhs);
        if (es == 3 && ps == 6 && segmented_assembly_wanted())
        {
            // W C^-1 came with the rhs products
            const int64_t waves_per_block = kBlock / kWave;
            if (w->segments > 0)
            {
                const int64_t seg_blocks = ceil_div(w->segments, waves_per_block * kSegsPerWave);
                if (seg_blocks > INT32_MAX) return fail(PM_OVERFLOW, "Schur assembly: too many segments");
                const int seg_grid = static_cast<int>(std::max<int64_t>(1, seg_blocks));
                hipLaunchKernelGGL((k_assemble_segments<3, 6>), dim3(seg_grid), dim3(kBlock), 0, stream, w->segments,
                                   static_cast<const uint2*>(w->segment_terms.as<uint2>()),
                                   static_cast<const double*>(w->TW), eq->cross_values, seg_partial);
            }
            const int slot_grid =
                static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(w->slots, waves_per_block), 65536)));
            hipLaunchKernelGGL(k_assemble_reduce<6>, dim3(slot_grid), dim3(kBlock), 0, stream, w->slots, w->pb,
                               w->slot_key, w->slot_start, w->entries, w->term_e, w->term_left, w->term_right,
                               static_cast<const unsigned*>(w->seg_offs), w->slot_mirror.as<unsigned>(), w->bro,
                               w->ro, w->Bd, eq->primary_cross_values, static_cast<const double*>(seg_partial),
                               w->values);
        }
        else
        {
            if (nc > 0 && !fused)
                hipLaunchKernelGGL((k_transform_cross<>), dim3(grid_for(nc * ps * es)), dim3(kBlock), 0, stream, nc, ps,
                                   es, eq->cross_eliminated, w->Cinv, eq->cross_values, w->TW);
            const int64_t waves_per_block = kBlock / kWave;
            const int slot_grid =
                static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(w->slots, waves_per_block), 65536)));
            const auto assemble = es == 3 && ps == 6 ? k_assemble_slots<3, 6> : k_assemble_slots<0, 0>;
            hipLaunchKernelGGL(assemble, dim3(slot_grid), dim3(kBlock), 0,
                               stream, w->slots, ps, es, w->pb, w->slot_key, w->slot_start, w->entries, w->term_e,
                               w->term_left, w->term_right, w->bro, w->ro, w->Bd, eq->primary_cross_values, w->TW,
                               eq->cross_values, w->values);
        }
        const auto preconditioner_kernel = ps == 6                  ? k_preconditioner<6, 6>
                                           : ps <= kSmallSchurBlock ? k_preconditioner<kSmallSchurBlock>
                                                                    : k_preconditioner<kMaxSchurBlock>;
        hipLaunchKernelGGL(preconditioner_kernel, dim3(grid_for(pc)), dim3(kBlock), 0, stream, pc, ps, w->diag_slot, w->bro,
                           w->ro, w->values, w->pinv, st);
    }
    PM_HIP(hipGetLastError());
    PM_TRY(w->read_status(stream));
    report->assembly_on_device = 1;
    if (w->host_status->bad_inverse != ~0ull)
    {
        std::snprintf(report->message, sizeof(report->message), "eliminated block is not positive definite");
        report->status = 1;
        return PM_OK;
    }
    if (w->host_status->bad_preconditioner != ~0ull)
    {
        std::snprintf(report->message, sizeof(report->message), "Schur preconditioner block is not positive definite");
        report->status = 2;
        return PM_OK;
    }
    return PM_OK;
}

} // namespace pm

extern "C" {

int pm_schur_workspace_create(pm_schur_workspace** out)
{
    if (!out) return fail(PM_INVALID_ARGUMENT, "pm_schur_workspace_create: null output");
    *out = new (std::nothrow) pm_schur_workspace();
    if (!*out) return fail(PM_RUNTIME_ERROR, "pm_schur_workspace_create: out of host memory");
    return PM_OK;
}

int pm_schur_workspace_destroy(pm_schur_workspace* w)
{
    delete w;
    return PM_OK;
}

int64_t pm_schur_pattern_build_count(const pm_schur_workspace* w) { return w ? w->pattern_builds : 0; }

int pm_schur_pattern_size(const pm_schur_workspace* w, int64_t* slots, int64_t* values, int64_t* terms)
{
    if (!w) return fail(PM_INVALID_ARGUMENT, "pm_schur_pattern_size: null workspace");
    if (slots) *slots = w->slots;
    if (values) *values = w->values_count;
    if (terms) *terms = w->terms;
    return PM_OK;
}

int pm_normal_equations_add_primary(pm_schur_workspace* w, const pm_normal_equations* in,
                                    const pm_primary_residual_batch* pb, pm_normal_equations* out,
                                    hipStream_t stream)
{
    if (!w || !in || !pb || !out) return fail(PM_INVALID_ARGUMENT, "pm_normal_equations_add_primary: null argument");
    PM_TRY(check_normal_equations(in));
    const int64_t pc = in->primary_count, ps = in->primary_size;
    const int64_t n = pb->count, k = pb->blocks_per_residual, rs = pb->residual_size;
    if (n < 0 || k <= 0 || rs <= 0)
        return fail(PM_INVALID_ARGUMENT,
                    "BlockNormalEquations: primary block and Jacobian lists must have equal non-zero size");
    if (n > 0 && (!pb->primary || !pb->jacobians || !pb->residuals))
        return fail(PM_INVALID_ARGUMENT,
                    "BlockNormalEquations::addPrimaryTerms: pointers must be non-null and residual size positive");
    if (n > 0 && pc * ps == 0)
        return fail(PM_OUT_OF_RANGE, "BlockNormalEquations: primary block index out of range");
    if (!pb->weights && (!std::isfinite(pb->huber_delta) || pb->huber_delta < 0.0))
        return fail(PM_INVALID_ARGUMENT, "evaluateHuberLoss: delta must be finite and non-negative");
    if (n * k >= (int64_t(1) << 40)) return fail(PM_OVERFLOW, "primary residual batch is too large");
    PM_TRY(w->bind(stream));
    // ids and the existing primary cross blocks on the host
    std::vector<int64_t> ids(static_cast<size_t>(n * k));
    const int64_t n0 = in->primary_cross_count;
    std::vector<int64_t> rows0(static_cast<size_t>(n0)), cols0(static_cast<size_t>(n0));
    if (n > 0) PM_HIP(hipMemcpyAsync(ids.data(), pb->primary, sizeof(int64_t) * n * k, hipMemcpyDeviceToHost, stream));
    if (n0 > 0)
    {
        PM_HIP(hipMemcpyAsync(rows0.data(), in->primary_cross_rows, sizeof(int64_t) * n0, hipMemcpyDeviceToHost,
                              stream));
        PM_HIP(hipMemcpyAsync(cols0.data(), in->primary_cross_cols, sizeof(int64_t) * n0, hipMemcpyDeviceToHost,
                              stream));
    }
    PM_HIP(hipStreamSynchronize(stream));
    int64_t bad_id = -1;
    bool bad_range = false;
    for (int64_t i = 0; i < n && bad_id < 0; ++i)
        for (int64_t a = 0; a < k; ++a)
        {
            const int64_t b = ids[static_cast<size_t>(i * k + a)];
            if (b < 0 || b >= pc)
            {
                bad_id = i;
                bad_range = true;
                break;
            }
            for (int64_t prior = 0; prior < a; ++prior)
                if (ids[static_cast<size_t>(i * k + prior)] == b) bad_id = i;
            if (bad_id >= 0) break;
        }
    // values and weights
    Carve mz;
    mz.take<double>(n);
    PM_TRY(w->pt_meta.reserve(mz.off));
    double* weights = w->pt_meta.as<double>();
    hipLaunchKernelGGL(k_schur_status_init, dim3(1), dim3(1), 0, stream, w->dstatus());
    if (n > 0)
        hipLaunchKernelGGL(k_pt_check, dim3(grid_for(n)), dim3(kBlock), 0, stream, n, static_cast<int>(k),
                           static_cast<int>(rs), static_cast<int>(ps), pb->jacobians, pb->residuals, pb->weights,
                           pb->huber_delta, weights, w->dstatus());
    PM_HIP(hipGetLastError());
    PM_TRY(w->read_status(stream));
    const SchurStatus& hs = *w->host_status;
    const unsigned long long bad_val = std::min(hs.bad_value, hs.bad_weight);
    // the lowest offending residual decides the error, as a sequence of
    // addPrimaryTerms calls would stop at it (ids are checked first within it)
    if (bad_id >= 0 && static_cast<unsigned long long>(bad_id) <= bad_val)
    {
        if (bad_range)
            return fail(PM_OUT_OF_RANGE, "BlockNormalEquations: primary block index out of range (residual %lld)",
                        static_cast<long long>(bad_id));
        return fail(PM_INVALID_ARGUMENT, "BlockNormalEquations: primary block indices must be unique (residual %lld)",
                    static_cast<long long>(bad_id));
    }
    if (hs.bad_value != ~0ull && hs.bad_value <= hs.bad_weight)
        return fail(PM_INVALID_ARGUMENT,
                    "BlockNormalEquations::addPrimaryTerms: Jacobian and residual values must be finite (residual "
                    "%llu)",
                    hs.bad_value);
    if (hs.bad_weight != ~0ull)
        return fail(PM_INVALID_ARGUMENT,
                    "BlockNormalEquations::addPrimaryTerms: weight must be finite and non-negative (residual %llu)",
                    hs.bad_weight);

    // groupings: diagonal terms by block (terms in residual order), primary
    // cross blocks found or created in the reference's order
    std::vector<int64_t> order(static_cast<size_t>(n * k));
    for (int64_t t = 0; t < n * k; ++t) order[static_cast<size_t>(t)] = t;
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t x, int64_t y) { return ids[static_cast<size_t>(x)] < ids[static_cast<size_t>(y)]; });
    std::vector<int64_t> dblocks, doffs{0};
    for (int64_t q = 0; q < n * k; ++q)
    {
        const int64_t b = ids[static_cast<size_t>(order[static_cast<size_t>(q)])];
        if (dblocks.empty() || dblocks.back() != b)
        {
            if (!dblocks.empty()) doffs.push_back(q);
            dblocks.push_back(b);
        }
    }
    if (!dblocks.empty()) doffs.push_back(n * k);
    std::unordered_map<int64_t, int64_t> where;
    std::vector<int64_t> crows(rows0), ccols(cols0);
    for (int64_t x = 0; x < n0; ++x) where.emplace(rows0[static_cast<size_t>(x)] * pc + cols0[static_cast<size_t>(x)], x);
    std::vector<std::vector<int64_t>> cterms(static_cast<size_t>(n0));
    for (int64_t i = 0; i < n; ++i)
        for (int64_t l = 0; l < k; ++l)
            for (int64_t r = l + 1; r < k; ++r)
            {
                const int64_t bl = ids[static_cast<size_t>(i * k + l)], br = ids[static_cast<size_t>(i * k + r)];
                const bool ordered = bl < br;
                const int64_t row = ordered ? bl : br, col = ordered ? br : bl;
                const int64_t rslot = ordered ? l : r, cslot = ordered ? r : l;
                auto it = where.find(row * pc + col);
                int64_t x;
                if (it == where.end())
                {
                    x = static_cast<int64_t>(crows.size());
                    where.emplace(row * pc + col, x);
                    crows.push_back(row);
                    ccols.push_back(col);
                    cterms.emplace_back();
                }
                else x = it->second;
                cterms[static_cast<size_t>(x)].push_back((i * k + rslot) * k + cslot);
            }
    const int64_t ncross = static_cast<int64_t>(crows.size());
    std::vector<int64_t> cindex, coffs{0}, cflat;
    for (int64_t x = 0; x < ncross; ++x)
    {
        const auto& v = cterms[static_cast<size_t>(x)];
        if (v.empty()) continue;
        cindex.push_back(x);
        cflat.insert(cflat.end(), v.begin(), v.end());
        coffs.push_back(static_cast<int64_t>(cflat.size()));
    }
    const int64_t dtouched = static_cast<int64_t>(dblocks.size()), ctouched = static_cast<int64_t>(cindex.size());

    // outputs: B and g copied from the input, then summed into; the cross
    // blocks' ids, with the existing values copied and the new ones zero
    Carve oz;
    oz.take<double>(pc * ps * ps);
    oz.take<double>(pc * ps);
    oz.take<int64_t>(ncross);
    oz.take<int64_t>(ncross);
    oz.take<double>(ncross * ps * ps);
    oz.take<int64_t>(dtouched);
    oz.take<int64_t>(dtouched + 1);
    oz.take<int64_t>(n * k);
    oz.take<int64_t>(ctouched);
    oz.take<int64_t>(ctouched + 1);
    oz.take<int64_t>(static_cast<int64_t>(cflat.size()));
    DeviceBuffer fresh;
    PM_TRY(fresh.reserve(oz.off));
    Carve o;
    o.base = fresh.as<char>();
    double* B = o.take<double>(pc * ps * ps);
    double* g = o.take<double>(pc * ps);
    int64_t* prow = o.take<int64_t>(ncross);
    int64_t* pcol = o.take<int64_t>(ncross);
    double* pval = o.take<double>(ncross * ps * ps);
    int64_t* d_blocks = o.take<int64_t>(dtouched);
    int64_t* d_offs = o.take<int64_t>(dtouched + 1);
    int64_t* d_terms = o.take<int64_t>(n * k);
    int64_t* c_index = o.take<int64_t>(ctouched);
    int64_t* c_offs = o.take<int64_t>(ctouched + 1);
    int64_t* c_terms = o.take<int64_t>(static_cast<int64_t>(cflat.size()));
    const size_t block_bytes = sizeof(double) * static_cast<size_t>(ps * ps);
    if (pc * ps > 0)
    {
        PM_HIP(hipMemcpyAsync(B, in->primary_diagonal, block_bytes * pc, hipMemcpyDeviceToDevice, stream));
        PM_HIP(hipMemcpyAsync(g, in->primary_gradient, sizeof(double) * pc * ps, hipMemcpyDeviceToDevice, stream));
    }
    if (n0 > 0)
        PM_HIP(hipMemcpyAsync(pval, in->primary_cross_values, block_bytes * n0, hipMemcpyDeviceToDevice, stream));
    if (ncross > n0)
        hipLaunchKernelGGL(k_fill_f64, dim3(grid_for((ncross - n0) * ps * ps)), dim3(kBlock), 0, stream,
                           pval + n0 * ps * ps, (ncross - n0) * ps * ps, 0.0);
    // the host arrays are consumed by the kernels below: copy them synchronously
    PM_HIP(hipStreamSynchronize(stream));
    auto up = [&](int64_t* dst, const std::vector<int64_t>& v) -> int {
        if (!v.empty()) PM_HIP(hipMemcpy(dst, v.data(), sizeof(int64_t) * v.size(), hipMemcpyHostToDevice));
        return PM_OK;
    };
    PM_TRY(up(prow, crows));
    PM_TRY(up(pcol, ccols));
    PM_TRY(up(d_blocks, dblocks));
    PM_TRY(up(d_offs, doffs));
    PM_TRY(up(d_terms, order));
    PM_TRY(up(c_index, cindex));
    PM_TRY(up(c_offs, coffs));
    PM_TRY(up(c_terms, cflat));
    if (dtouched > 0)
        hipLaunchKernelGGL(k_pt_diag, dim3(grid_for(dtouched * (ps * ps + ps))), dim3(kBlock), 0, stream, dtouched,
                           d_blocks, d_offs, d_terms, static_cast<int>(k), static_cast<int>(rs),
                           static_cast<int>(ps), pb->jacobians, pb->residuals, weights, B, g);
    if (ctouched > 0)
        hipLaunchKernelGGL(k_pt_cross, dim3(grid_for(ctouched * ps * ps)), dim3(kBlock), 0, stream, ctouched,
                           c_index, c_offs, c_terms, static_cast<int>(k), static_cast<int>(rs), static_cast<int>(ps),
                           pb->jacobians, weights, pval);
    PM_HIP(hipGetLastError());
    // the previous outputs (possibly this call's input) are no longer read
    PM_HIP(hipStreamSynchronize(stream));
    w->pt_cross.release();
    w->pt_cross = fresh;
    pm_normal_equations res = *in;
    res.primary_diagonal = B;
    res.primary_gradient = g;
    res.primary_cross_count = ncross;
    res.primary_cross_rows = prow;
    res.primary_cross_cols = pcol;
    res.primary_cross_values = pval;
    *out = res;
    return PM_OK;
}

int pm_normal_equations_assemble(pm_schur_workspace* w, int64_t pc, int64_t ec, int64_t ps, int64_t es,
                                 const pm_residual_batch* rb, pm_normal_equations* out, hipStream_t stream)
{
    if (!w || !rb || !out) return fail(PM_INVALID_ARGUMENT, "pm_normal_equations_assemble: null argument");
    PM_TRY(check_schur_sizes(pc, ec, ps, es));
    if (rb->count < 0 || rb->residual_size <= 0)
        return fail(PM_INVALID_ARGUMENT,
                    "BlockNormalEquations::addResidualBlocks: pointers must be non-null and residual size positive");
    if (rb->count >= (int64_t(1) << 31)) return fail(PM_OVERFLOW, "residual batch must hold fewer than 2^31 residuals");
    if (rb->count > 0 && (ps <= 0 || es <= 0 || pc <= 0 || ec <= 0))
        return fail(PM_INVALID_ARGUMENT, "a residual batch needs primary and eliminated blocks");
    if (rb->count > 0 && (!rb->primary || !rb->eliminated || !rb->primary_jacobians || !rb->eliminated_jacobians ||
                          !rb->residuals))
        return fail(PM_INVALID_ARGUMENT,
                    "BlockNormalEquations::addResidualBlocks: pointers must be non-null and residual size positive");
    if (!rb->weights && (!std::isfinite(rb->huber_delta) || rb->huber_delta < 0.0))
        return fail(PM_INVALID_ARGUMENT, "evaluateHuberLoss: delta must be finite and non-negative");
    PM_TRY(w->bind(stream));
    const int64_t n = rb->count;
    const int64_t sig[6] = {n, pc, ec, ps, es, rb->residual_size};
    hipLaunchKernelGGL(k_schur_status_init, dim3(1), dim3(1), 0, stream, w->dstatus());
    bool reuse = w->ne_builds > 0 && std::equal(sig, sig + 6, w->ne_sig);
    if (reuse && n > 0)
    {
        hipLaunchKernelGGL(k_count_mismatch, dim3(grid_for(n)), dim3(kBlock), 0, stream, rb->primary,
                           w->ne_primary_copy, n, w->dstatus());
        hipLaunchKernelGGL(k_count_mismatch, dim3(grid_for(n)), dim3(kBlock), 0, stream, rb->eliminated,
                           w->ne_eliminated_copy, n, w->dstatus());
        PM_HIP(hipGetLastError());
    }
    // values buffer needs the run count: size it for the worst case (n runs)
    Carve vz;
    vz.take<double>(pc * ps * ps);
    vz.take<double>(ec * es * es);
    vz.take<double>(pc * ps);
    vz.take<double>(ec * es);
    vz.take<double>(n * ps * es);
    vz.take<double>(n);
    PM_TRY(w->ne_values.reserve(vz.off));
    Carve v;
    v.base = w->ne_values.as<char>();
    double* B = v.take<double>(pc * ps * ps);
    double* C = v.take<double>(ec * es * es);
    double* gp = v.take<double>(pc * ps);
    double* ge = v.take<double>(ec * es);
    double* W = v.take<double>(n * ps * es);
    double* weights = v.take<double>(n);
    if (reuse)
    {
        PM_TRY(w->read_status(stream));
        reuse = w->host_status->mismatch == 0;
        hipLaunchKernelGGL(k_schur_status_init, dim3(1), dim3(1), 0, stream, w->dstatus());
    }
    PM_TRY(residual_pass(w, rb, pc, ec, static_cast<int>(ps), static_cast<int>(es), weights, stream, !reuse));
    std::copy(sig, sig + 6, w->ne_sig);
    const int rs = static_cast<int>(rb->residual_size);
    const int64_t runs = w->ne_runs;
    if (pc * ps > 0)
    {
        if (rs == 2 && ps == 6)
            hipLaunchKernelGGL((k_ne_primary_staged<2, 6>),
                               dim3(static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(pc, 4), 65536)))),
                               dim3(kBlock), 0, stream, pc, w->ne_offs_p, w->ne_perm_p, rb->primary_jacobians,
                               rb->residuals, weights, B, gp);
        else
            hipLaunchKernelGGL(rs == 2 ? k_ne_primary<2> : k_ne_primary<0>, dim3(grid_for(pc * (ps * ps + ps))),
                               dim3(kBlock), 0, stream, pc, rs, static_cast<int>(ps), w->ne_offs_p, w->ne_perm_p,
                               rb->primary_jacobians, rb->residuals, weights, B, gp);
    }
    if (ec * es > 0 && rs == 2 && es == 3)
        hipLaunchKernelGGL((k_ne_eliminated_points<2, 3>), dim3(grid_for(ec)), dim3(kBlock), 0, stream, ec,
                           w->ne_offs_e, w->ne_perm_e, rb->eliminated_jacobians, rb->residuals, weights, C, ge);
    else if (ec * es > 0)
        hipLaunchKernelGGL(k_ne_eliminated, dim3(grid_for(ec * (es * es + es))), dim3(kBlock), 0, stream, ec, rs,
                           static_cast<int>(es), w->ne_offs_e, w->ne_perm_e, rb->eliminated_jacobians,
                           rb->residuals, weights, C, ge);
    if (runs > 0 && rs == 2 && ps == 6 && es == 3)
    {
        // single-residual pairs in residual order, the others per run
        hipLaunchKernelGGL((k_ne_cross_single<2, 6, 3>), dim3(grid_for(n)), dim3(kBlock), 0, stream, n,
                           static_cast<const unsigned*>(w->ne_res_cross), rb->primary_jacobians,
                           rb->eliminated_jacobians, weights, W);
        hipLaunchKernelGGL((k_ne_cross_runs<2, 6, 3>), dim3(grid_for(runs)), dim3(kBlock), 0, stream, runs, n,
                           w->ne_run_start, w->ne_run_cross, w->ne_perm_pe, rb->primary_jacobians,
                           rb->eliminated_jacobians, weights, W, true);
    }
    else if (runs > 0)
        hipLaunchKernelGGL(k_ne_cross<>, dim3(grid_for(runs * ps * es)), dim3(kBlock), 0, stream, runs, n, rs,
                           static_cast<int>(ps), static_cast<int>(es), w->ne_run_start, w->ne_run_cross,
                           w->ne_perm_pe, rb->primary_jacobians, rb->eliminated_jacobians, weights, W);
    PM_HIP(hipGetLastError());
    out->primary_count = pc;
    out->eliminated_count = ec;
    out->primary_size = ps;
    out->eliminated_size = es;
    out->primary_diagonal = B;
    out->eliminated_diagonal = C;
    out->primary_gradient = gp;
    out->eliminated_gradient = ge;
    out->cross_count = runs;
    out->cross_primary = w->ne_cross_primary;
    out->cross_eliminated = w->ne_cross_eliminated;
    out->cross_values = W;
    out->primary_cross_count = 0;
    out->primary_cross_rows = nullptr;
    out->primary_cross_cols = nullptr;
    out->primary_cross_values = nullptr;
    return PM_OK;
}

namespace
{
// Seed of the FP64 block-PCG (block_schur_accelerated.cpp:131-193): the
// reduced system, right-hand side and block inverses rounded to FP32 and
// solved by FP32 block-PCG (at most 200 iterations, relative tolerance at
// least 1e-3, absolute at least 1e-7, tolerance max(abs, rel ||r0||)).  A
// converged seed becomes the FP64 solve's initial guess; a seed that does not
// converge or breaks down leaves the FP64 solve's start unchanged.
int mixed_precision_seed(pm_schur_workspace* w, const pm_csr& a, int ps, const pm_schur_options* o, double* dp,
                         pm_schur_report* report, pm_workspace* ws, hipStream_t stream)
{
    const int64_t pd = a.rows, nnz = a.nnz;
    auto al = [](int64_t n) { return static_cast<size_t>((n * 4 + 255) & ~int64_t(255)); };
    PM_TRY(w->seed.reserve(al(nnz) + al(pd) + al(pd * ps) + al(pd)));
    char* base = w->seed.as<char>();
    float* values = reinterpret_cast<float*>(base);
    float* rhs = reinterpret_cast<float*>(base + al(nnz));
    float* inv = reinterpret_cast<float*>(base + al(nnz) + al(pd));
    float* x = reinterpret_cast<float*>(base + al(nnz) + al(pd) + al(pd * ps));
    hipLaunchKernelGGL(k_f64_to_f32, dim3(grid_for(nnz)), dim3(kBlock), 0, stream, w->values, nnz, values);
    hipLaunchKernelGGL(k_f64_to_f32, dim3(grid_for(pd)), dim3(kBlock), 0, stream, w->rhs, pd, rhs);
    hipLaunchKernelGGL(k_f64_to_f32, dim3(grid_for(pd * ps)), dim3(kBlock), 0, stream, w->pinv, pd * ps, inv);
    if (o->use_initial_guess)
        hipLaunchKernelGGL(k_f64_to_f32, dim3(grid_for(pd)), dim3(kBlock), 0, stream, dp, pd, x);
    else PM_HIP(hipMemsetAsync(x, 0, sizeof(float) * pd, stream));
    PM_HIP(hipGetLastError());
    pm_csr af = a;
    af.values = values;
    af.dtype = PM_F32;
    pm_solver_options so{};
    so.max_iterations = std::min(o->max_iterations, 200);
    so.relative_tolerance = std::max(1.0e-3, static_cast<double>(static_cast<float>(o->relative_tolerance)));
    so.absolute_tolerance = std::max(static_cast<double>(1.0e-7f), static_cast<double>(static_cast<float>(o->absolute_tolerance)));
    so.use_jacobi = 1;
    so.require_convergence = 0;
    pm_solver_report sr{};
    const int st = block_pcg_rule(&af, rhs, x, inv, ps, &so, &sr, ws, stream, 0);
    if (st == PM_RUNTIME_ERROR) return PM_OK;  // a breakdown of the seed: the FP64 solve starts as asked
    PM_TRY(st);
    if (sr.converged)
    {
        hipLaunchKernelGGL(k_f32_to_f64, dim3(grid_for(pd)), dim3(kBlock), 0, stream, x, pd, dp);
        PM_HIP(hipGetLastError());
        report->mixed_precision_used = 1;
    }
    return PM_OK;
}
} // namespace

int pm_schur_solve(pm_schur_workspace* w, const pm_normal_equations* eq, double damping, const pm_schur_options* o,
                   double* dp, double* de, pm_schur_report* report, pm_workspace* ws, hipStream_t stream)
{
    if (!w || !o || !report || !ws) return fail(PM_INVALID_ARGUMENT, "solveDampedSchurComplement: null argument");
    std::memset(report, 0, sizeof(*report));
    PM_TRY(check_normal_equations(eq));
    const int64_t pc = eq->primary_count, ec = eq->eliminated_count;
    const int ps = static_cast<int>(eq->primary_size), es = static_cast<int>(eq->eliminated_size);
    const int64_t pd = pc * ps, ed = ec * es;
    if ((pd > 0 && !dp) || (ed > 0 && !de) || (dp && dp == de))
        return fail(PM_INVALID_ARGUMENT, "solveDampedSchurComplement: output vectors must be non-null and distinct");
    if (!std::isfinite(damping) || damping < 0.0 || o->max_iterations <= 0 || !std::isfinite(o->relative_tolerance) ||
        o->relative_tolerance < 0.0 || !std::isfinite(o->absolute_tolerance) || o->absolute_tolerance < 0.0 ||
        !std::isfinite(o->minimum_diagonal) || o->minimum_diagonal <= 0.0 ||
        (o->linear_backend != PM_SCHUR_BLOCK_PCG && o->linear_backend != PM_SCHUR_DENSE_CHOLESKY) ||
        (o->flags & ~(PM_SCHUR_TOLERANCE_SUM | PM_SCHUR_MIXED_PRECISION)) != 0)
        return fail(PM_INVALID_ARGUMENT, "solveDampedSchurComplement: invalid solver options");
    const bool dense = o->linear_backend == PM_SCHUR_DENSE_CHOLESKY;
    PM_TRY(schur_assemble(w, eq, damping, o->minimum_diagonal, report, stream));
    if (dense && report->status == 2)
    {
        // a diagonal block of the reduced system is not positive definite, so
        // neither is the system: the dense backend's failure
        report->status = 8;
        std::snprintf(report->message, sizeof(report->message), "reduced Schur complement is not positive definite");
    }
    if (report->status != 0) return PM_OK;
    if (ed > 0) hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(ed)), dim3(kBlock), 0, stream, de, ed, 0.0);
    if (pd > 0 && !o->use_initial_guess && !dense)
        hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(pd)), dim3(kBlock), 0, stream, dp, pd, 0.0);
    PM_HIP(hipGetLastError());
    if (pd == 0)
    {
        report->converged = 1;
    }
    else
    {
        pm_csr a{};
        a.rows = pd;
        a.cols = pd;
        a.nnz = w->values_count;
        a.row_offsets = w->ro;
        a.col_indices = w->ci;
        a.values = w->values;
        a.dtype = PM_F64;
        a.row_offsets32 = w->ro32;
        a.col_indices32 = w->ci32;
        if (dense)
        {
            PM_TRY(ws->bind(stream));
            DenseCholeskyOutcome dc;
            PM_TRY(dense_cholesky_solve(&a, w->rhs, dp, ws, &dc, stream));
            report->iterations = 1;
            report->initial_residual_norm = dc.rhs_norm;
            report->final_residual_norm = dc.residual_norm;
            if (dc.not_positive_definite)
            {
                report->status = 8;
                std::snprintf(report->message, sizeof(report->message),
                              "reduced Schur complement is not positive definite");
                return PM_OK;
            }
            // block_schur_dense_solver.h:384-388
            const double tol = o->absolute_tolerance + o->relative_tolerance * dc.rhs_norm;
            const double floor = 100.0 * std::numeric_limits<double>::epsilon() * std::max(1.0, dc.rhs_norm);
            report->converged = std::isfinite(dc.residual_norm) && dc.residual_norm <= std::max(tol, floor);
            if (!report->converged)
            {
                report->status = 16;
                std::snprintf(report->message, sizeof(report->message), "dense Cholesky residual exceeds tolerance");
                return PM_OK;
            }
        }
        else
        {
            if (o->flags & PM_SCHUR_MIXED_PRECISION) PM_TRY(mixed_precision_seed(w, a, ps, o, dp, report, ws, stream));
            pm_solver_options so{};
            so.max_iterations = o->max_iterations;
            so.relative_tolerance = o->relative_tolerance;
            so.absolute_tolerance = o->absolute_tolerance;
            so.use_jacobi = 1;
            so.require_convergence = 0;
            pm_solver_report sr{};
            PM_TRY(block_pcg_rule(&a, w->rhs, dp, w->pinv, ps, &so, &sr, ws, stream,
                                  (o->flags & PM_SCHUR_TOLERANCE_SUM) ? 1 : 0));
            report->converged = sr.converged;
            report->iterations = sr.iterations;
            report->initial_residual_norm = sr.initial_residual;
            report->final_residual_norm = sr.final_residual;
        }
    }
    if (!report->converged)
    {
        report->status = 4;
        std::snprintf(report->message, sizeof(report->message), "PCG iteration limit reached");
        return PM_OK;
    }
    if (ec > 0 && pd > 0 && eq->cross_count > 0)
    {
        const int64_t nc = eq->cross_count;
        hipLaunchKernelGGL((ps == 6 && es == 3 ? k_cross_point<6, 3> : k_cross_point<>), dim3(grid_for(nc * es)),
                           dim3(kBlock), 0, stream, nc, ps, es,
                           eq->cross_primary, eq->cross_values, dp, w->crv);
        const auto point_kernel = es == 3                  ? k_back_substitute_point<3, 3>
                                  : es <= kSmallSchurBlock ? k_back_substitute_point<kSmallSchurBlock>
                                                           : k_back_substitute_point<kMaxSchurBlock>;
        hipLaunchKernelGGL(point_kernel, dim3(grid_for(ec)), dim3(kBlock), 0, stream, ec, es, w->adj_offs, w->adj,
                           static_cast<const double*>(w->crv), w->Cinv, eq->eliminated_gradient, de);
    }
    else if (ec > 0)
    {
        const auto back_kernel =
            es <= kSmallSchurBlock ? k_back_substitute<kSmallSchurBlock> : k_back_substitute<kMaxSchurBlock>;
        hipLaunchKernelGGL(back_kernel, dim3(grid_for(ec)), dim3(kBlock), 0, stream, ec, ps, es,
                           pd > 0 ? w->adj_offs : nullptr, w->adj, eq->cross_primary, eq->cross_values, w->Cinv,
                           eq->eliminated_gradient, dp, de);
    }
    PM_HIP(hipGetLastError());
    std::snprintf(report->message, sizeof(report->message), "converged");
    return PM_OK;
}

int pm_schur_assemble(pm_schur_workspace* w, const pm_normal_equations* eq, double damping, double minimum_diagonal,
                      pm_schur_report* report, hipStream_t stream)
{
    if (!w || !report) return fail(PM_INVALID_ARGUMENT, "pm_schur_assemble: null argument");
    std::memset(report, 0, sizeof(*report));
    PM_TRY(check_normal_equations(eq));
    if (!std::isfinite(damping) || damping < 0.0 || !std::isfinite(minimum_diagonal) || minimum_diagonal <= 0.0)
        return fail(PM_INVALID_ARGUMENT, "solveDampedSchurComplement: invalid solver options");
    if (eq->primary_count * eq->primary_size == 0)
        return fail(PM_INVALID_ARGUMENT, "pm_schur_assemble: the reduced system is empty");
    PM_TRY(schur_assemble(w, eq, damping, minimum_diagonal, report, stream));
    if (report->status == 0) std::snprintf(report->message, sizeof(report->message), "assembled");
    return PM_OK;
}

int pm_schur_system(const pm_schur_workspace* w, pm_csr* reduced, const double** rhs, const double** preconditioner)
{
    if (!w || !reduced) return fail(PM_INVALID_ARGUMENT, "pm_schur_system: null argument");
    if (w->pattern_builds == 0 || !w->values) return fail(PM_LOGIC_ERROR, "pm_schur_system: no assembled system");
    std::memset(reduced, 0, sizeof(*reduced));
    reduced->rows = w->sig[0] * w->sig[2];
    reduced->cols = reduced->rows;
    reduced->nnz = w->values_count;
    reduced->row_offsets = w->ro;
    reduced->col_indices = w->ci;
    reduced->values = w->values;
    reduced->dtype = PM_F64;
    reduced->row_offsets32 = w->ro32;
    reduced->col_indices32 = w->ci32;
    if (rhs) *rhs = w->rhs;
    if (preconditioner) *preconditioner = w->pinv;
    return PM_OK;
}

} // extern "C"
