// Learned factorized sparse approximate inverse (no counterpart in the
// reference): a 3-layer message-passing network over the CSR graph of an SPD
// matrix A emits a lower-triangular G on the lower pattern of A with a
// positive diagonal, so M^-1 = G^T G is SPD by construction, and is trained
// on the GPU to minimise the probe loss
//     L = (1/P) sum_p ||(I - G A G^T) z_p||^2 / ||z_p||^2.
//
// Network (hidden width 64, BF16 weights and activations, FP32 accumulation):
//   H0 = relu(X W_in + b_in)                      X: 32 node features
//   M_l = S H_{l-1}                               S_ij = a_ij / sqrt(a_ii a_jj), i != j
//   H_l = relu([H_{l-1}, M_l] W_l + b_l)          l = 1..3
//   off-diagonal (i > j): g_ij = (w_off . (H_i * H_j) + b_off + alpha S_ij) / sqrt(a_jj)
//   diagonal:             g_ii = softplus(w_diag . H_i + b_diag) / sqrt(a_ii)
// At initialisation w_off = b_off = alpha = w_diag = 0 and softplus(b_diag) = 1,
// so G = D^-1/2 and M^-1 is the Jacobi preconditioner.
//
// The dense layers run on v_mfma_f32_16x16x32_bf16 (forward X W, backward
// dP W^T and the weight gradients X^T dP); graph aggregation, the decoder and
// the probe loss are row-gather kernels over CSR with no atomics, and every
// reduction over rows is per-workgroup partials summed in a fixed order, so a
// training step is deterministic.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "pm_common.hpp"
#include "sort.hpp"
#include "spmv_kernels.hpp"
#include "tuning.hpp"
#include "workspace.hpp"

namespace pm
{

constexpr int kHid = 64;     // hidden width
constexpr int kFeat = 32;    // node input features (padded to one MFMA K step)
constexpr int kLayers = 3;
constexpr int kMaxProbes = 8;

// parameter layout (FP32 masters)
constexpr int kOffWin = 0;
constexpr int kOffBin = kOffWin + kFeat * kHid;
constexpr int kOffWl = kOffBin + kHid;                       // kLayers x (2*kHid x kHid)
constexpr int kOffBl = kOffWl + kLayers * 2 * kHid * kHid;   // kLayers x kHid
constexpr int kOffWoff = kOffBl + kLayers * kHid;
constexpr int kOffBoff = kOffWoff + kHid;
constexpr int kOffAlpha = kOffBoff + 1;
constexpr int kOffWdiag = kOffAlpha + 1;
constexpr int kOffBdiag = kOffWdiag + kHid;
constexpr int kParams = kOffBdiag + 1;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf(const __bf16 v) { return static_cast<float>(v); }
__device__ __forceinline__ __bf16 to_bf(float v) { return static_cast<__bf16>(v); }

__device__ __forceinline__ int64_t gid() { return static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; }
__device__ __forceinline__ int64_t gstr() { return static_cast<int64_t>(gridDim.x) * blockDim.x; }

inline int grid_of(int64_t n, int cap = 8192)
{
    return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, kBlock), cap)));
}

// ---- graph preparation ------------------------------------------------------------

// Per row: 1/sqrt(diagonal), lower-pattern length, normalised couplings of the
// row's off-diagonal entries, node features.
template <typename S>
__global__ __launch_bounds__(kBlock) void k_graph_rows(int64_t n, const int32_t* __restrict__ ro,
                                                       const int32_t* __restrict__ ci, const S* __restrict__ v,
                                                       float* __restrict__ dinv, unsigned* __restrict__ lower_count,
                                                       unsigned* __restrict__ upper_count, unsigned* __restrict__ bad,
                                                       unsigned* __restrict__ unsorted)
{
    for (int64_t i = gid(); i < n; i += gstr())
    {
        double d = 0.0;
        unsigned lc = 0, uc = 0;
        int32_t prev = -1;
        bool increasing = true;
        for (int32_t k = ro[i]; k < ro[i + 1]; ++k)
        {
            if (ci[k] == i) d += static_cast<double>(v[k]);
            if (ci[k] <= i) ++lc;
            if (ci[k] >= i) ++uc;
            if (ci[k] <= prev) increasing = false;
            prev = ci[k];
        }
        const bool ok = d > 0.0 && isfinite(d);
        if (!ok) atomicMin(bad, static_cast<unsigned>(i < 0xffffffffll ? i : 0xfffffffe));
        if (!increasing) atomicOr(unsorted, 1u);
        dinv[i] = ok ? static_cast<float>(1.0 / sqrt(d)) : 0.0f;
        lower_count[i] = lc;
        upper_count[i] = uc;
    }
}

// G^T's pattern straight from A's when A's pattern is symmetric: row i of G^T
// is A's row i from the diagonal on (its last uc entries when the row is
// sorted), and entry (i, j) of G^T is entry (j, i) of G, found by bisection in
// G's row j.  Eight lanes share a row.  A missing match raises *asym and the
// caller builds G^T by sorting instead; writes stay inside row i's uc slots
// whatever the input, so an unsorted row is caught by the flags, not by a
// stray store.
constexpr int kUpperLanes = 8;

__global__ __launch_bounds__(kBlock) void k_graph_upper_sym(int64_t n, const int32_t* __restrict__ ro,
                                                            const int32_t* __restrict__ ci,
                                                            const int32_t* __restrict__ uro,
                                                            const int32_t* __restrict__ lro,
                                                            const int32_t* __restrict__ lci,
                                                            int32_t* __restrict__ uci, int32_t* __restrict__ uperm,
                                                            unsigned* __restrict__ asym)
{
    const int sub = static_cast<int>(gid() % kUpperLanes);
    for (int64_t i = gid() / kUpperLanes; i < n; i += gstr() / kUpperLanes)
    {
        const int32_t uc = uro[i + 1] - uro[i];
        const int32_t first = ro[i + 1] - uc;
        bool missing = false;
        for (int32_t k = first + sub; k < ro[i + 1]; k += kUpperLanes)
        {
            const int32_t j = ci[k];
            int32_t lo = lro[j], hi = lro[j + 1];  // G's row j: columns <= j, increasing
            while (lo < hi)
            {
                const int32_t mid = lo + (hi - lo) / 2;
                if (lci[mid] < i) lo = mid + 1;
                else hi = mid;
            }
            const bool found = lo < lro[j + 1] && lci[lo] == i;
            missing |= !found;
            const int32_t out = uro[i] + (k - first);
            uci[out] = j;
            uperm[out] = found ? lo : 0;
        }
        if (missing) atomicOr(asym, 1u);
    }
}

template <typename S>
__global__ __launch_bounds__(kBlock) void k_graph_features(int64_t n, const int32_t* __restrict__ ro,
                                                           const int32_t* __restrict__ ci, const S* __restrict__ v,
                                                           const float* __restrict__ dinv, float* __restrict__ s,
                                                           __bf16* __restrict__ x)
{
    for (int64_t i = gid(); i < n; i += gstr())
    {
        float sum = 0.f, sabs = 0.f, smax = -1e30f, smin = 1e30f, amax = 0.f;
        int deg = 0;
        double d = 0.0;
        for (int32_t k = ro[i]; k < ro[i + 1]; ++k)
        {
            const int32_t j = ci[k];
            if (j == i)
            {
                s[k] = 0.f;
                d += static_cast<double>(v[k]);
                continue;
            }
            const float sij = static_cast<float>(v[k]) * dinv[i] * dinv[j];
            s[k] = sij;
            sum += sij;
            sabs += fabsf(sij);
            smax = fmaxf(smax, sij);
            smin = fminf(smin, sij);
            amax = fmaxf(amax, fabsf(sij));
            ++deg;
        }
        if (deg == 0)
        {
            smax = 0.f;
            smin = 0.f;
        }
        float f[kFeat];
#pragma unroll
        for (int c = 0; c < kFeat; ++c) f[c] = 0.f;
        f[0] = 1.f;
        f[1] = static_cast<float>(deg) / 8.f;
        f[2] = sum;
        f[3] = sabs;
        f[4] = smax;
        f[5] = smin;
        f[6] = amax;
        f[7] = sabs > 0.f ? sum / sabs : 0.f;
        f[8] = static_cast<float>(log(d > 0.0 ? d : 1.0)) / 8.f;
        f[9] = sabs * sabs;
#pragma unroll
        for (int c = 0; c < kFeat; ++c) x[i * kFeat + c] = to_bf(f[c]);
    }
}

// Lower pattern (col <= row) in row order, with the source entry of A.
__global__ void k_graph_flags(unsigned* flags)
{
    flags[0] = 0xffffffffu;  // first row with a bad diagonal
    flags[1] = 0u;           // a row with unsorted or repeated columns
    flags[2] = 0u;           // an upper entry without its lower mirror
}

__global__ __launch_bounds__(kBlock) void k_graph_lower(int64_t n, const int32_t* __restrict__ ro,
                                                        const int32_t* __restrict__ ci,
                                                        const unsigned* __restrict__ lro, int32_t* __restrict__ lro32,
                                                        int32_t* __restrict__ lci, int32_t* __restrict__ lsrc,
                                                        int32_t* __restrict__ ldiag, unsigned long long* __restrict__ keys,
                                                        unsigned* __restrict__ vals, int rb)
{
    for (int64_t i = gid(); i < n; i += gstr())
    {
        unsigned out = lro[i];
        lro32[i] = static_cast<int32_t>(out);
        int32_t diag = -1;
        for (int32_t k = ro[i]; k < ro[i + 1]; ++k)
        {
            const int32_t j = ci[k];
            if (j > i) continue;
            lci[out] = j;
            lsrc[out] = k;
            if (j == i) diag = static_cast<int32_t>(out);
            if (keys)
            {
                keys[out] = (static_cast<unsigned long long>(j) << rb) | static_cast<unsigned long long>(i);
                vals[out] = out;
            }
            ++out;
        }
        ldiag[i] = diag;
        if (i == n - 1) lro32[n] = static_cast<int32_t>(out);
    }
}

// Transposed lower pattern (upper CSR) from the (col, row)-sorted entries.
__global__ __launch_bounds__(kBlock) void k_graph_upper(int64_t nl, const unsigned long long* __restrict__ keys,
                                                        const unsigned* __restrict__ perm, int rb,
                                                        int32_t* __restrict__ uci, int32_t* __restrict__ uperm)
{
    const unsigned long long mask = (1ull << rb) - 1ull;
    for (int64_t k = gid(); k < nl; k += gstr())
    {
        uci[k] = static_cast<int32_t>(keys[k] & mask);
        uperm[k] = static_cast<int32_t>(perm[k]);
    }
}

__global__ __launch_bounds__(kBlock) void k_row_starts(int64_t n, int64_t nl, const unsigned long long* __restrict__ keys,
                                                       int rb, int32_t* __restrict__ uro)
{
    for (int64_t r = gid(); r <= n; r += gstr())
    {
        int64_t lo = 0, hi = nl;
        while (lo < hi)
        {
            const int64_t mid = (lo + hi) / 2;
            if (static_cast<int64_t>(keys[mid] >> rb) < r) lo = mid + 1;
            else hi = mid;
        }
        uro[r] = static_cast<int32_t>(lo);
    }
}

// ---- BF16 MFMA GEMM ---------------------------------------------------------------
//
// Y[n x N] = A[n x K] * B[K x N] stored FP32 (the backward input gradient
// dP W^T), A row-major (lda = K) and B supplied transposed (Bt: N x K
// row-major), FP32 accumulation.  One wavefront owns 16 rows and all N
// columns; per K step of 32 a lane loads 8 consecutive BF16 of A (row l%16)
// and of Bt (column l%16 of the tile), the operand layout of
// v_mfma_f32_16x16x32_bf16, whose result lane l holds rows (l/16)*4 + i of
// column l%16.  A may be the concatenation of two row-major blocks [A0 | A1]
// (each K/2 wide).
template <int K, int N>
__global__ __launch_bounds__(kBlock) void k_gemm(int64_t n, const __bf16* __restrict__ a0,
                                                 const __bf16* __restrict__ a1, const __bf16* __restrict__ bt,
                                                 float* __restrict__ yout)
{
    constexpr int KH = (K > 64) ? K / 2 : K;  // width of each A block
    const int lane = threadIdx.x % kWave;
    const int64_t waves = static_cast<int64_t>(gridDim.x) * (kBlock / kWave);
    for (int64_t t = static_cast<int64_t>(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave; t * 16 < n;
         t += waves)
    {
        const int64_t row0 = t * 16;
        floatx4 acc[N / 16];
#pragma unroll
        for (int j = 0; j < N / 16; ++j) acc[j] = floatx4{0.f, 0.f, 0.f, 0.f};
        const int64_t arow = row0 + lane % 16;
        const bool valid = arow < n;
#pragma unroll
        for (int k0 = 0; k0 < K; k0 += 32)
        {
            const int kk = k0 + (lane / 16) * 8;
            bf16x8 av;
            if (valid)
            {
                const __bf16* src = (K > 64 && kk >= KH) ? a1 + arow * KH + (kk - KH)
                                                          : a0 + arow * KH + kk;
                av = *reinterpret_cast<const bf16x8*>(src);
            }
            else
            {
#pragma unroll
                for (int q = 0; q < 8; ++q) av[q] = to_bf(0.f);
            }
#pragma unroll
            for (int j = 0; j < N / 16; ++j)
            {
                const bf16x8 bv = *reinterpret_cast<const bf16x8*>(bt + (j * 16 + lane % 16) * K + kk);
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[j], 0, 0, 0);
            }
        }
#pragma unroll
        for (int j = 0; j < N / 16; ++j)
        {
            const int col = j * 16 + lane % 16;
#pragma unroll
            for (int i = 0; i < 4; ++i)
            {
                const int64_t row = row0 + (lane / 16) * 4 + i;
                if (row >= n) continue;
                yout[row * N + col] = acc[j][i];
            }
        }
    }
}

// One network layer with the graph aggregation fused in:
//     H_out = relu([A, S A] W + b)   (AGG, K = 128, A = H_in of width 64)
//     H_out = relu(A W + b)          (!AGG, K = 32, A = X)
// One wavefront per 16-row tile, all 64 output columns; each workgroup owns a
// contiguous, XCD-local range of tiles, so the rows a tile gathers from
// (neighbours in the matrix ordering) are mostly in the same XCD's L2.  W^T
// is staged once per workgroup in LDS.  Each lane gathers the M = S A
// channels its own MFMA operand needs (row l%16, channels (l/16)*8 and
// 32 + (l/16)*8), accumulating in FP32 in CSR order and rounding to BF16
// exactly as a separate aggregation would, so the result equals aggregation
// followed by a plain GEMM.  STORE_M keeps M for the backward pass.  The
// 16 x 64 BF16 output tile is staged in LDS and written as 2 KB of contiguous
// 16-byte stores.
template <int K, bool AGG, bool STORE_M>
__global__ __launch_bounds__(kBlock) void k_layer(int64_t n, const int32_t* __restrict__ ro,
                                                  const int32_t* __restrict__ ci, const float* __restrict__ s,
                                                  const __bf16* __restrict__ a, const __bf16* __restrict__ bt,
                                                  const float* __restrict__ bias, __bf16* __restrict__ hout,
                                                  __bf16* __restrict__ mout)
{
    static_assert((AGG && K == 2 * kHid) || (!AGG && K == kFeat), "layer shape");
    constexpr int KS = K / 32;                 // MFMA K steps
    constexpr int KA = AGG ? kHid : kFeat;     // width of a row of A
    __shared__ __attribute__((aligned(16))) __bf16 wlds[kHid * K];
    __shared__ __attribute__((aligned(16))) __bf16 stage[kBlock / kWave][16 * kHid];
    for (int e = threadIdx.x * 8; e < kHid * K; e += kBlock * 8)
        *reinterpret_cast<bf16x8*>(wlds + e) = *reinterpret_cast<const bf16x8*>(bt + e);
    __syncthreads();
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    const int quad = lane / 16;
    float bcol[kHid / 16];
#pragma unroll
    for (int j = 0; j < kHid / 16; ++j) bcol[j] = bias[j * 16 + lane % 16];
    __bf16* st = stage[wave];
    int64_t tb, te;
    block_range(ceil_div(n, int64_t(16)), 1, tb, te);
    for (int64_t t = tb + wave; t < te; t += kBlock / kWave)
    {
        const int64_t row0 = t * 16;
        const int64_t row = row0 + lane % 16;
        const bool valid = row < n;
        bf16x8 av[KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
#pragma unroll
            for (int q = 0; q < 8; ++q) av[ks][q] = to_bf(0.f);
        if (valid)
        {
            av[0] = *reinterpret_cast<const bf16x8*>(a + row * KA + quad * 8);
            if (AGG) av[1] = *reinterpret_cast<const bf16x8*>(a + row * KA + 32 + quad * 8);
        }
        if (AGG)
        {
            float m1[8], m2[8];
#pragma unroll
            for (int q = 0; q < 8; ++q)
            {
                m1[q] = 0.f;
                m2[q] = 0.f;
            }
            if (valid)
            {
                const int32_t k1 = ro[row + 1];
                constexpr int U = 4;
                // batches of U entries, all gathers of a batch in flight at
                // once; entries past the row end and the diagonal are skipped
                for (int32_t k = ro[row]; k < k1; k += U)
                {
                    int32_t jj[U];
                    float sv[U];
                    bf16x8 g1[U], g2[U];
#pragma unroll
                    for (int u = 0; u < U; ++u)
                    {
                        const bool in = k + u < k1;
                        jj[u] = in ? ci[k + u] : static_cast<int32_t>(row);
                        sv[u] = in ? s[k + u] : 0.f;
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u)
                    {
                        g1[u] = *reinterpret_cast<const bf16x8*>(a + static_cast<int64_t>(jj[u]) * kHid + quad * 8);
                        g2[u] = *reinterpret_cast<const bf16x8*>(a + static_cast<int64_t>(jj[u]) * kHid + 32 + quad * 8);
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u)
                    {
                        if (jj[u] == row) continue;
#pragma unroll
                        for (int q = 0; q < 8; ++q)
                        {
                            m1[q] = fmaf(sv[u], bf(g1[u][q]), m1[q]);
                            m2[q] = fmaf(sv[u], bf(g2[u][q]), m2[q]);
                        }
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < 8; ++q)
            {
                av[KS - 2][q] = to_bf(m1[q]);
                av[KS - 1][q] = to_bf(m2[q]);
            }
            if (STORE_M && valid)
            {
                *reinterpret_cast<bf16x8*>(mout + row * kHid + quad * 8) = av[KS - 2];
                *reinterpret_cast<bf16x8*>(mout + row * kHid + 32 + quad * 8) = av[KS - 1];
            }
        }
        floatx4 acc[kHid / 16];
#pragma unroll
        for (int j = 0; j < kHid / 16; ++j) acc[j] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
#pragma unroll
            for (int j = 0; j < kHid / 16; ++j)
            {
                const bf16x8 bv = *reinterpret_cast<const bf16x8*>(wlds + (j * 16 + lane % 16) * K + ks * 32 + quad * 8);
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[ks], bv, acc[j], 0, 0, 0);
            }
#pragma unroll
        for (int j = 0; j < kHid / 16; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                st[(quad * 4 + i) * kHid + j * 16 + lane % 16] = to_bf(fmaxf(acc[j][i] + bcol[j], 0.f));
        wave_lds_sync();
        // lane -> row lane/4 of the tile, 16 channels from (lane%4)*16
        const int srow = lane / 4, scol = (lane % 4) * 16;
        if (row0 + srow < n)
        {
            const bf16x8 o0 = *reinterpret_cast<const bf16x8*>(st + srow * kHid + scol);
            const bf16x8 o1 = *reinterpret_cast<const bf16x8*>(st + srow * kHid + scol + 8);
            *reinterpret_cast<bf16x8*>(hout + (row0 + srow) * kHid + scol) = o0;
            *reinterpret_cast<bf16x8*>(hout + (row0 + srow) * kHid + scol + 8) = o1;
        }
        wave_lds_sync();
    }
}

// Weight gradient dW[K x 64] = A^T dP over all rows, A = [A0 | A1] as above,
// dP (n x 64, BF16).  Wavefront (row strip, 16-wide K tile): per 32 rows a
// lane supplies A^T[k][rows] and dP[rows][c] (strided BF16 reads that stay in
// L1/L2), accumulating a 16 x 64 tile.  Each workgroup writes its partial
// tile set to part[blockIdx][K x 64]; k_reduce_parts sums them in order.
template <int K>
__global__ __launch_bounds__(kBlock) void k_gemm_wgrad(int64_t n, const __bf16* __restrict__ a0,
                                                       const __bf16* __restrict__ a1, const __bf16* __restrict__ dp,
                                                       float* __restrict__ part)
{
    constexpr int KH = (K > 64) ? K / 2 : K;
    constexpr int KT = K / 16;  // K tiles; the 4 waves of a workgroup cover them round-robin
    __shared__ float tile[K * kHid];
    for (int i = threadIdx.x; i < K * kHid; i += kBlock) tile[i] = 0.f;
    __syncthreads();
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    const int64_t strips = ceil_div(n, 32);
    for (int kt = wave; kt < KT; kt += kBlock / kWave)
    {
        floatx4 acc[kHid / 16];
#pragma unroll
        for (int j = 0; j < kHid / 16; ++j) acc[j] = floatx4{0.f, 0.f, 0.f, 0.f};
        const int krow = kt * 16 + lane % 16;  // output row (input feature)
        const __bf16* abase = krow < KH ? a0 : a1;
        const int kcol = krow < KH ? krow : krow - KH;
        for (int64_t s = blockIdx.x; s < strips; s += gridDim.x)
        {
            const int64_t r0 = s * 32 + (lane / 16) * 8;
            bf16x8 av;
#pragma unroll
            for (int q = 0; q < 8; ++q) av[q] = r0 + q < n ? abase[(r0 + q) * KH + kcol] : to_bf(0.f);
#pragma unroll
            for (int j = 0; j < kHid / 16; ++j)
            {
                bf16x8 bv;
#pragma unroll
                for (int q = 0; q < 8; ++q) bv[q] = r0 + q < n ? dp[(r0 + q) * kHid + j * 16 + lane % 16] : to_bf(0.f);
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[j], 0, 0, 0);
            }
        }
#pragma unroll
        for (int j = 0; j < kHid / 16; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i) tile[(kt * 16 + (lane / 16) * 4 + i) * kHid + j * 16 + lane % 16] = acc[j][i];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < K * kHid; i += kBlock) part[static_cast<int64_t>(blockIdx.x) * K * kHid + i] = tile[i];
}

// out[i] = sum over b of part[b][i] in block order.
__global__ __launch_bounds__(kBlock) void k_reduce_parts(const float* __restrict__ part, int blocks, int width,
                                                         float* __restrict__ out)
{
    for (int64_t i = gid(); i < width; i += gstr())
    {
        float s = 0.f;
        for (int b = 0; b < blocks; ++b) s += part[static_cast<int64_t>(b) * width + i];
        out[i] = s;
    }
}

// ---- graph kernels ----------------------------------------------------------------

// dH[i] += sum_j S_ij dM[j] (S symmetric), FP32 in and out.
__global__ __launch_bounds__(kBlock) void k_aggregate_back(int64_t n, const int32_t* __restrict__ ro,
                                                           const int32_t* __restrict__ ci, const float* __restrict__ s,
                                                           const float* __restrict__ dm, float* __restrict__ dh)
{
    const int lane = threadIdx.x % kWave;
    const int64_t waves = static_cast<int64_t>(gridDim.x) * (kBlock / kWave);
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave; i < n; i += waves)
    {
        float acc = dh[i * kHid + lane];
        for (int32_t k = ro[i]; k < ro[i + 1]; ++k)
        {
            const int32_t j = ci[k];
            if (j != i) acc = fmaf(s[k], dm[static_cast<int64_t>(j) * kHid + lane], acc);
        }
        dh[i * kHid + lane] = acc;
    }
}

__device__ __forceinline__ float softplus(float x) { return x > 20.f ? x : log1pf(expf(x)); }
__device__ __forceinline__ float sigmoid(float x) { return 1.f / (1.f + expf(-x)); }

// Decoder: G on the lower pattern, also the raw outputs for the backward
// pass.  Eight lanes per row, each holding 8 of the 64 channels of H_i (one
// 16-byte load); an entry's 64-term dot product is 8 FP32 terms per lane in
// channel order, then a 3-step butterfly over the row's 8 lanes.
template <typename G>
__global__ __launch_bounds__(kBlock) void k_decode(int64_t n, const int32_t* __restrict__ lro,
                                                   const int32_t* __restrict__ lci, const int32_t* __restrict__ lsrc,
                                                   const float* __restrict__ s, const float* __restrict__ dinv,
                                                   const __bf16* __restrict__ h, const float* __restrict__ params,
                                                   float* __restrict__ out, G* __restrict__ g)
{
    constexpr int kLanesPerRow = 8;
    constexpr int kRowsPerWave = kWave / kLanesPerRow;
    const int lane = threadIdx.x % kWave;
    const int sub = lane % kLanesPerRow;
    float w_off[8], w_diag[8];
#pragma unroll
    for (int q = 0; q < 8; ++q)
    {
        w_off[q] = params[kOffWoff + sub * 8 + q];
        w_diag[q] = params[kOffWdiag + sub * 8 + q];
    }
    const float b_off = params[kOffBoff], alpha = params[kOffAlpha], b_diag = params[kOffBdiag];
    // contiguous XCD-local row range per workgroup; 32 rows per pass
    constexpr int kRowsPerPass = (kBlock / kWave) * kRowsPerWave;
    int64_t rb, re;
    block_range(n, kRowsPerPass, rb, re);
    for (int64_t i = rb + (threadIdx.x / kWave) * kRowsPerWave + lane / kLanesPerRow; i < re; i += kRowsPerPass)
    {
        const bf16x8 hv = *reinterpret_cast<const bf16x8*>(h + i * kHid + sub * 8);
        float hi[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) hi[q] = bf(hv[q]);
        const int32_t k1 = lro[i + 1];
        // batches of U entries: every index and neighbour load of a batch is
        // in flight at once (long rows, such as a reduced Schur system's ~70
        // lower entries, would otherwise be one dependent chain per entry)
        constexpr int U = 4;
        for (int32_t k0 = lro[i]; k0 < k1; k0 += U)
        {
            int32_t jj[U];
#pragma unroll
            for (int u = 0; u < U; ++u) jj[u] = k0 + u < k1 ? lci[k0 + u] : static_cast<int32_t>(i);
            bf16x8 hj[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                hj[u] = *reinterpret_cast<const bf16x8*>(h + static_cast<int64_t>(jj[u]) * kHid + sub * 8);
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
                const int32_t k = k0 + u;
                if (k >= k1) break;
                const int32_t j = jj[u];
                float part = 0.f;
                if (j == i)
                {
#pragma unroll
                    for (int q = 0; q < 8; ++q) part = fmaf(w_diag[q], hi[q], part);
                }
                else
                {
#pragma unroll
                    for (int q = 0; q < 8; ++q) part = fmaf(w_off[q] * hi[q], bf(hj[u][q]), part);
                }
#pragma unroll
                for (int off = kLanesPerRow / 2; off > 0; off >>= 1) part += __shfl_xor(part, off, kLanesPerRow);
                if (sub == 0)
                {
                    float o;
                    if (j == i)
                    {
                        o = part + b_diag;
                        g[k] = static_cast<G>(softplus(o) * dinv[i]);
                    }
                    else
                    {
                        o = part + b_off + alpha * s[lsrc[k]];
                        g[k] = static_cast<G>(o * dinv[j]);
                    }
                    out[k] = o;
                }
            }
        }
    }
}

// G^T values in upper-CSR order.
template <typename G>
__global__ __launch_bounds__(kBlock) void k_gather_upper(int64_t nl, const int32_t* __restrict__ uperm,
                                                         const G* __restrict__ g, G* __restrict__ gt)
{
    for (int64_t k = gid(); k < nl; k += gstr()) gt[k] = g[uperm[k]];
}

// ---- probe loss ---------------------------------------------------------------------

__device__ __forceinline__ unsigned long long mix64(unsigned long long x)
{
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

// Standard normal probe values z[i][p] from (seed, i, p) (Box-Muller).
__global__ __launch_bounds__(kBlock) void k_probes(int64_t n, int probes, unsigned long long seed, float* __restrict__ z)
{
    for (int64_t t = gid(); t < n * probes; t += gstr())
    {
        const unsigned long long h1 = mix64(seed * 0x9e3779b97f4a7c15ull + static_cast<unsigned long long>(t) * 2 + 1);
        const unsigned long long h2 = mix64(h1 ^ 0xd1b54a32d192ed03ull);
        const double u1 = (static_cast<double>(h1 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        const double u2 = static_cast<double>(h2 >> 11) * (1.0 / 9007199254740992.0);
        z[t] = static_cast<float>(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
    }
}

// Y[i][p] = sum_k val_k X[col_k][p], 8 lanes per row (one per probe).
template <typename V>
__global__ __launch_bounds__(kBlock) void k_spmm8(int64_t n, int probes, const int32_t* __restrict__ ro,
                                                  const int32_t* __restrict__ ci, const V* __restrict__ val,
                                                  const float* __restrict__ x, float* __restrict__ y)
{
    const int64_t groups = static_cast<int64_t>(gridDim.x) * (kBlock / kMaxProbes);
    const int p = threadIdx.x % kMaxProbes;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * (kBlock / kMaxProbes) + threadIdx.x / kMaxProbes; i < n;
         i += groups)
    {
        if (p >= probes) continue;
        float acc = 0.f;
        for (int32_t k = ro[i]; k < ro[i + 1]; ++k)
            acc = fmaf(static_cast<float>(val[k]), x[static_cast<int64_t>(ci[k]) * probes + p], acc);
        y[i * probes + p] = acc;
    }
}

// r = z - w; per-workgroup partials of ||r_p||^2 and ||z_p||^2.
__global__ __launch_bounds__(kBlock) void k_residual_probes(int64_t n, int probes, const float* __restrict__ z,
                                                            const float* __restrict__ w, float* __restrict__ r,
                                                            double* __restrict__ part)
{
    __shared__ double red[2 * kMaxProbes][kBlock / kWave];
    double rr[kMaxProbes], zz[kMaxProbes];
#pragma unroll
    for (int p = 0; p < kMaxProbes; ++p) rr[p] = zz[p] = 0.0;
    for (int64_t i = gid(); i < n; i += gstr())
        for (int p = 0; p < probes; ++p)
        {
            const float zi = z[i * probes + p];
            const float ri = zi - w[i * probes + p];
            r[i * probes + p] = ri;
            rr[p] += static_cast<double>(ri) * ri;
            zz[p] += static_cast<double>(zi) * zi;
        }
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
#pragma unroll
    for (int p = 0; p < kMaxProbes; ++p)
    {
        const double a = wave_sum(rr[p]), b = wave_sum(zz[p]);
        if (lane == 0)
        {
            red[p][wave] = a;
            red[kMaxProbes + p][wave] = b;
        }
    }
    __syncthreads();
    if (threadIdx.x < 2 * kMaxProbes)
    {
        double s = 0.0;
        for (int w2 = 0; w2 < kBlock / kWave; ++w2) s += red[threadIdx.x][w2];
        part[static_cast<int64_t>(blockIdx.x) * 2 * kMaxProbes + threadIdx.x] = s;
    }
}

// Per-system choice between the learned factor and Jacobi.  Per row, over the
// probes: e[2i] = sum_p (z - G A G^T z)_i^2 (w holds G A G^T z) and, once the
// Jacobi product is in v (v = A D^-1/2 z), e[2i+1] = sum_p (z - D^-1/2 v)_i^2.
template <bool JACOBI>
__global__ __launch_bounds__(kBlock) void k_select_rows(int64_t n, int probes, const float* __restrict__ z,
                                                        const float* __restrict__ w, const float* __restrict__ dinv,
                                                        double* __restrict__ e)
{
    for (int64_t i = gid(); i < n; i += gstr())
    {
        double acc = 0.0;
        const float d = JACOBI ? dinv[i] : 1.f;
        for (int p = 0; p < probes; ++p)
        {
            const float wi = JACOBI ? d * w[i * probes + p] : w[i * probes + p];
            const double ri = static_cast<double>(z[i * probes + p] - wi);
            acc = __builtin_fma(ri, ri, acc);
        }
        e[2 * i + (JACOBI ? 1 : 0)] = acc;
    }
}

// y[i][p] = dinv[i] x[i][p]
__global__ __launch_bounds__(kBlock) void k_scale_probe_rows(int64_t n, int probes, const float* __restrict__ dinv,
                                                             const float* __restrict__ x, float* __restrict__ y)
{
    for (int64_t t = gid(); t < n * probes; t += gstr()) y[t] = dinv[t / probes] * x[t];
}

// One wavefront per system: the two probe losses summed over its rows in a
// fixed order; Jacobi is chosen when the learned factor does not lower it.
__global__ __launch_bounds__(kBlock) void k_select_systems(int64_t systems, const int64_t* __restrict__ range,
                                                           const double* __restrict__ e, double* __restrict__ loss,
                                                           uint8_t* __restrict__ jacobi)
{
    const int lane = threadIdx.x % kWave;
    const int64_t waves = static_cast<int64_t>(gridDim.x) * (kBlock / kWave);
    for (int64_t s = static_cast<int64_t>(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave; s < systems;
         s += waves)
    {
        double l = 0.0, j = 0.0;
        for (int64_t i = range[s] + lane; i < range[systems + s]; i += kWave)
        {
            l += e[2 * i];
            j += e[2 * i + 1];
        }
        l = wave_sum(l);
        j = wave_sum(j);
        if (lane == 0)
        {
            loss[2 * s] = l;
            loss[2 * s + 1] = j;
            jacobi[s] = l < j ? 0 : 1;
        }
    }
}

// loss and per-probe scale c_p = 1 / (P ||z_p||^2); stat[0] = loss.
__global__ void k_loss(const double* __restrict__ part, int blocks, int probes, double* __restrict__ stat)
{
    if (threadIdx.x != 0) return;
    double loss = 0.0;
    for (int p = 0; p < probes; ++p)
    {
        double rr = 0.0, zz = 0.0;
        for (int b = 0; b < blocks; ++b)
        {
            rr += part[static_cast<int64_t>(b) * 2 * kMaxProbes + p];
            zz += part[static_cast<int64_t>(b) * 2 * kMaxProbes + kMaxProbes + p];
        }
        const double c = zz > 0.0 ? 1.0 / (probes * zz) : 0.0;
        stat[1 + p] = c;
        loss += rr * c;
    }
    stat[0] = loss;
}

// dL/dw = -2 c_p r (in place).
__global__ __launch_bounds__(kBlock) void k_scale_probes(int64_t n, int probes, const double* __restrict__ stat,
                                                         float* __restrict__ r)
{
    for (int64_t t = gid(); t < n * probes; t += gstr())
        r[t] = static_cast<float>(-2.0 * stat[1 + t % probes]) * r[t];
}

// dL/dG on the lower pattern: sum_p dw[i][p] v[j][p] + z[i][p] du[j][p]; then
// through the parametrisation to dL/d(out).
__global__ __launch_bounds__(kBlock) void k_grad_out(int64_t n, int probes, const int32_t* __restrict__ lro,
                                                     const int32_t* __restrict__ lci, const float* __restrict__ dinv,
                                                     const float* __restrict__ out, const float* __restrict__ dw,
                                                     const float* __restrict__ v, const float* __restrict__ z,
                                                     const float* __restrict__ du, float* __restrict__ dout)
{
    for (int64_t i = gid(); i < n; i += gstr())
        for (int32_t k = lro[i]; k < lro[i + 1]; ++k)
        {
            const int32_t j = lci[k];
            float dg = 0.f;
            for (int p = 0; p < probes; ++p)
                dg += dw[i * probes + p] * v[static_cast<int64_t>(j) * probes + p] +
                      z[i * probes + p] * du[static_cast<int64_t>(j) * probes + p];
            dout[k] = j == i ? dg * dinv[i] * sigmoid(out[k]) : dg * dinv[j];
        }
}

// Decoder backward: dH3 (FP32, one wavefront per row) and per-workgroup
// partials of the decoder parameter gradients
// [w_off(64) | b_off | alpha | w_diag(64) | b_diag] (131 values).
constexpr int kDecParams = 2 * kHid + 3;
__global__ __launch_bounds__(kBlock) void k_decode_back(int64_t n, const int32_t* __restrict__ lro,
                                                        const int32_t* __restrict__ lci, const int32_t* __restrict__ lsrc,
                                                        const int32_t* __restrict__ uro, const int32_t* __restrict__ uci,
                                                        const int32_t* __restrict__ uperm, const float* __restrict__ s,
                                                        const __bf16* __restrict__ h, const float* __restrict__ params,
                                                        const float* __restrict__ dout, float* __restrict__ dh,
                                                        float* __restrict__ part)
{
    __shared__ float red[kBlock / kWave][kDecParams];
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int64_t waves = static_cast<int64_t>(gridDim.x) * (kBlock / kWave);
    const float w_off = params[kOffWoff + lane], w_diag = params[kOffWdiag + lane];
    float gw_off = 0.f, gw_diag = 0.f, gb_off = 0.f, galpha = 0.f, gb_diag = 0.f;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * (kBlock / kWave) + wave; i < n; i += waves)
    {
        const float hi = bf(h[i * kHid + lane]);
        float acc = 0.f;
        for (int32_t k = lro[i]; k < lro[i + 1]; ++k)
        {
            const int32_t j = lci[k];
            const float d = dout[k];
            if (j == i)
            {
                acc += d * w_diag;
                gw_diag += d * hi;
                gb_diag += d;
            }
            else
            {
                const float hj = bf(h[static_cast<int64_t>(j) * kHid + lane]);
                acc += d * w_off * hj;
                gw_off += d * hi * hj;
                gb_off += d;
                galpha += d * s[lsrc[k]];
            }
        }
        // entries (m, i) with m > i: row i of the upper pattern
        for (int32_t k = uro[i]; k < uro[i + 1]; ++k)
        {
            const int32_t m = uci[k];
            if (m == i) continue;
            acc += dout[uperm[k]] * w_off * bf(h[static_cast<int64_t>(m) * kHid + lane]);
        }
        dh[i * kHid + lane] = acc;
    }
    red[wave][lane] = gw_off;
    red[wave][kHid + 2 + lane] = gw_diag;
    if (lane == 0)
    {
        red[wave][kHid] = gb_off;
        red[wave][kHid + 1] = galpha;
        red[wave][2 * kHid + 2] = gb_diag;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < kDecParams; q += kBlock)
    {
        float t = 0.f;
        for (int w2 = 0; w2 < kBlock / kWave; ++w2) t += red[w2][q];
        part[static_cast<int64_t>(blockIdx.x) * kDecParams + q] = t;
    }
}

// dP = dH * (H > 0) as BF16 for the MFMA backward, and per-workgroup bias
// gradient partials (64 values).
__global__ __launch_bounds__(kBlock) void k_relu_back(int64_t n, const float* __restrict__ dh,
                                                      const __bf16* __restrict__ h, __bf16* __restrict__ dp,
                                                      float* __restrict__ part)
{
    __shared__ float red[kBlock / kWave][kHid];
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int64_t waves = static_cast<int64_t>(gridDim.x) * (kBlock / kWave);
    float gb = 0.f;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * (kBlock / kWave) + wave; i < n; i += waves)
    {
        const float d = bf(h[i * kHid + lane]) > 0.f ? dh[i * kHid + lane] : 0.f;
        dp[i * kHid + lane] = to_bf(d);
        gb += d;
    }
    red[wave][lane] = gb;
    __syncthreads();
    if (threadIdx.x < kHid)
    {
        float t = 0.f;
        for (int w2 = 0; w2 < kBlock / kWave; ++w2) t += red[w2][threadIdx.x];
        part[static_cast<int64_t>(blockIdx.x) * kHid + threadIdx.x] = t;
    }
}

// Split the [dH | dM] input gradient (n x 128 FP32) into dH (overwrite) and dM.
__global__ __launch_bounds__(kBlock) void k_split_grad(int64_t n, const float* __restrict__ y, float* __restrict__ dh,
                                                       float* __restrict__ dm)
{
    for (int64_t t = gid(); t < n * kHid; t += gstr())
    {
        const int64_t i = t / kHid, c = t % kHid;
        dh[t] = y[i * 2 * kHid + c];
        dm[t] = y[i * 2 * kHid + kHid + c];
    }
}

// Adam on the FP32 masters, then the BF16 weight copies (W and W^T).
__global__ __launch_bounds__(kBlock) void k_adam(int count, float* __restrict__ p, float* __restrict__ m,
                                                 float* __restrict__ v, const float* __restrict__ g, float lr,
                                                 float b1, float b2, float eps, float c1, float c2)
{
    for (int i = static_cast<int>(gid()); i < count; i += static_cast<int>(gstr()))
    {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    }
}

__global__ __launch_bounds__(kBlock) void k_weights_bf16(const float* __restrict__ p, __bf16* __restrict__ w,
                                                         __bf16* __restrict__ wt)
{
    // W_in: kFeat x kHid -> w[0..), W_in^T: kHid x kFeat
    // W_l : 2kHid x kHid  -> w[kFeat*kHid + l*2kHid*kHid ..), W_l^T: kHid x 2kHid
    const int total = kFeat * kHid + kLayers * 2 * kHid * kHid;
    for (int t = static_cast<int>(gid()); t < total; t += static_cast<int>(gstr()))
    {
        int rows, base, local;
        if (t < kFeat * kHid)
        {
            rows = kFeat;
            base = 0;
            local = t;
        }
        else
        {
            rows = 2 * kHid;
            const int q = t - kFeat * kHid;
            base = kFeat * kHid + (q / (2 * kHid * kHid)) * 2 * kHid * kHid;
            local = q % (2 * kHid * kHid);
        }
        const int r = local / kHid, c = local % kHid;
        const float val = p[(t < kFeat * kHid ? kOffWin : kOffWl) + (t < kFeat * kHid ? t : t - kFeat * kHid)];
        w[base + r * kHid + c] = to_bf(val);
        wt[base + c * rows + r] = to_bf(val);
    }
}

} // namespace pm

using namespace pm;

struct pm_fsai_net
{
    DeviceBuffer params;  // kParams FP32 masters
    DeviceBuffer adam_m, adam_v, grad;
    DeviceBuffer w, wt;   // BF16 weights and transposes
    int64_t steps = 0;
};

struct pm_fsai_graph
{
    hipStream_t stream = nullptr;
    int64_t n = 0, nnz = 0, nl = 0;
    int32_t dtype = PM_F32;
    const int32_t* ro = nullptr;  // A (int32 mirrors, not owned)
    const int32_t* ci = nullptr;
    const void* values = nullptr;
    DeviceBuffer buf;      // graph arrays
    DeviceBuffer act;      // activations and gradients
    DeviceBuffer parts;    // per-workgroup partials
    DeviceBuffer tmp;      // graph-build counts and scan scratch
    DeviceBuffer select;   // per-system ranges and losses of pm_fsai_batch_select
    int device = 0;
    float* dinv = nullptr;
    float* s = nullptr;
    __bf16* x = nullptr;
    int32_t *lro = nullptr, *lci = nullptr, *lsrc = nullptr, *ldiag = nullptr;
    int32_t *uro = nullptr, *uci = nullptr, *uperm = nullptr;
    int64_t *lro64 = nullptr, *lci64 = nullptr, *uro64 = nullptr, *uci64 = nullptr;
    void* g = nullptr;    // G values (A's dtype), lower CSR order
    void* gt = nullptr;   // G^T values, upper CSR order
    float* out = nullptr; // decoder outputs
};

namespace pm
{

__global__ __launch_bounds__(kBlock) void k_widen(int64_t n, const int32_t* __restrict__ a, int64_t* __restrict__ b)
{
    for (int64_t i = gid(); i < n; i += gstr()) b[i] = a[i];
}

// activations layout: H[0..3] (n x 64 BF16), M[1..3] (n x 64 BF16), dP (n x 64
// BF16), dH, dM (n x 64 FP32), Y (n x 128 FP32), probe arrays (n x 8 FP32 x 6),
// dout (nl FP32), stat (16 doubles)
struct Act
{
    __bf16* h[kLayers + 1];
    __bf16* m[kLayers + 1];
    __bf16* dp;
    float *dh, *dm, *y;
    float *z, *u, *v, *w, *r, *q;
    float* dout;
    double* stat;
};

Act carve_act(pm_fsai_graph* gr, size_t* bytes)
{
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t n = static_cast<size_t>(std::max<int64_t>(gr->n, 1));
    const size_t nl = static_cast<size_t>(std::max<int64_t>(gr->nl, 1));
    const size_t hb = al(n * kHid * sizeof(__bf16)), hf = al(n * kHid * sizeof(float));
    const size_t pb = al(n * kMaxProbes * sizeof(float));
    size_t off = 0;
    Act a{};
    char* base = gr->act.as<char>();
    auto take = [&](size_t b) {
        char* p = base ? base + off : nullptr;
        off += b;
        return p;
    };
    for (int l = 0; l <= kLayers; ++l) a.h[l] = reinterpret_cast<__bf16*>(take(hb));
    for (int l = 0; l <= kLayers; ++l) a.m[l] = reinterpret_cast<__bf16*>(take(hb));
    a.dp = reinterpret_cast<__bf16*>(take(hb));
    a.dh = reinterpret_cast<float*>(take(hf));
    a.dm = reinterpret_cast<float*>(take(hf));
    a.y = reinterpret_cast<float*>(take(2 * hf));
    a.z = reinterpret_cast<float*>(take(pb));
    a.u = reinterpret_cast<float*>(take(pb));
    a.v = reinterpret_cast<float*>(take(pb));
    a.w = reinterpret_cast<float*>(take(pb));
    a.r = reinterpret_cast<float*>(take(pb));
    a.q = reinterpret_cast<float*>(take(pb));
    a.dout = reinterpret_cast<float*>(take(al(nl * sizeof(float))));
    a.stat = reinterpret_cast<double*>(take(256));
    *bytes = off;
    return a;
}

constexpr int kPartBlocks = 512;  // workgroups of the partial-sum kernels

int graph_build(pm_fsai_graph* gr, const pm_csr* a, hipStream_t stream)
{
    const int64_t n = a->rows;
    gr->n = n;
    gr->nnz = a->nnz;
    gr->dtype = a->dtype;
    gr->ro = a->row_offsets32;
    gr->ci = a->col_indices32;
    gr->values = a->values;
    gr->stream = stream;
    // temporaries: per-row counts, their scans and the scan scratch; the
    // key/value arrays of the sorted transpose only when A's pattern is not
    // symmetric
    DeviceBuffer sort_tmp;
    struct Release
    {
        DeviceBuffer& b;
        ~Release() { b.release(); }
    } rel{sort_tmp};
    const int64_t cap = std::max<int64_t>(a->nnz, 1);
    const size_t tbytes = sizeof(unsigned) * (n + 2) * 3 + sizeof(unsigned) * scan_scratch_words(n + 2) + 4096;
    PM_TRY(gr->tmp.reserve(tbytes));
    unsigned* counts = gr->tmp.as<unsigned>();
    unsigned* lro_u = counts + (n + 2);
    unsigned* ucounts = lro_u + (n + 2);
    unsigned* scratch = ucounts + (n + 2);
    unsigned* flags = scratch + scan_scratch_words(n + 2) + 8;  // [bad row, unsorted, asym]

    // graph buffer: dinv, s, x, lro, ldiag (n), lci, lsrc (nl<=nnz), uro, uci, uperm, 64-bit mirrors, g, gt, out
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t es = a->dtype == PM_F64 ? 8 : 4;
    const size_t gbytes = al(n * 4) + al(cap * 4) + al(n * kFeat * 2) + al((n + 1) * 4) * 2 + al(n * 4) +
                          al(cap * 4) * 4 + al((n + 1) * 8) * 2 + al(cap * 8) * 2 + al(cap * es) * 2 + al(cap * 4);
    PM_TRY(gr->buf.reserve(gbytes));
    char* p = gr->buf.as<char>();
    auto take = [&](size_t b) {
        char* q = p;
        p += al(b);
        return q;
    };
    gr->dinv = reinterpret_cast<float*>(take(n * 4));
    gr->s = reinterpret_cast<float*>(take(cap * 4));
    gr->x = reinterpret_cast<__bf16*>(take(n * kFeat * 2));
    gr->lro = reinterpret_cast<int32_t*>(take((n + 1) * 4));
    gr->uro = reinterpret_cast<int32_t*>(take((n + 1) * 4));
    gr->ldiag = reinterpret_cast<int32_t*>(take(n * 4));
    gr->lci = reinterpret_cast<int32_t*>(take(cap * 4));
    gr->lsrc = reinterpret_cast<int32_t*>(take(cap * 4));
    gr->uci = reinterpret_cast<int32_t*>(take(cap * 4));
    gr->uperm = reinterpret_cast<int32_t*>(take(cap * 4));
    gr->lro64 = reinterpret_cast<int64_t*>(take((n + 1) * 8));
    gr->uro64 = reinterpret_cast<int64_t*>(take((n + 1) * 8));
    gr->lci64 = reinterpret_cast<int64_t*>(take(cap * 8));
    gr->uci64 = reinterpret_cast<int64_t*>(take(cap * 8));
    gr->g = take(cap * es);
    gr->gt = take(cap * es);
    gr->out = reinterpret_cast<float*>(take(cap * 4));

    hipLaunchKernelGGL(k_graph_flags, dim3(1), dim3(1), 0, stream, flags);
    const int g = grid_of(n);
    if (a->dtype == PM_F32)
        hipLaunchKernelGGL(k_graph_rows<float>, dim3(g), dim3(kBlock), 0, stream, n, gr->ro, gr->ci,
                           static_cast<const float*>(a->values), gr->dinv, counts, ucounts, flags, flags + 1);
    else
        hipLaunchKernelGGL(k_graph_rows<double>, dim3(g), dim3(kBlock), 0, stream, n, gr->ro, gr->ci,
                           static_cast<const double*>(a->values), gr->dinv, counts, ucounts, flags, flags + 1);
    PM_HIP(hipGetLastError());
    PM_TRY(scan_u32(counts, lro_u, n + 1, scratch, nullptr, stream));
    PM_TRY(scan_u32(ucounts, reinterpret_cast<unsigned*>(gr->uro), n + 1, scratch, nullptr, stream));
    if (a->dtype == PM_F32)
        hipLaunchKernelGGL(k_graph_features<float>, dim3(g), dim3(kBlock), 0, stream, n, gr->ro, gr->ci,
                           static_cast<const float*>(a->values), gr->dinv, gr->s, gr->x);
    else
        hipLaunchKernelGGL(k_graph_features<double>, dim3(g), dim3(kBlock), 0, stream, n, gr->ro, gr->ci,
                           static_cast<const double*>(a->values), gr->dinv, gr->s, gr->x);
    const int rb = std::max(1, bits_for(n));
    hipLaunchKernelGGL(k_graph_lower, dim3(g), dim3(kBlock), 0, stream, n, gr->ro, gr->ci, lro_u, gr->lro, gr->lci,
                       gr->lsrc, gr->ldiag, static_cast<unsigned long long*>(nullptr), static_cast<unsigned*>(nullptr),
                       rb);
    hipLaunchKernelGGL(k_graph_upper_sym, dim3(grid_of(n * kUpperLanes)), dim3(kBlock), 0, stream, n, gr->ro, gr->ci,
                       gr->uro, gr->lro, gr->lci, gr->uci, gr->uperm, flags + 2);
    PM_HIP(hipGetLastError());
    struct
    {
        unsigned flags[3];
        unsigned lower_total, upper_total;
    } h{};
    PM_HIP(hipMemcpyAsync(h.flags, flags, sizeof(h.flags), hipMemcpyDeviceToHost, stream));
    PM_HIP(hipMemcpyAsync(&h.lower_total, lro_u + n, sizeof(unsigned), hipMemcpyDeviceToHost, stream));
    PM_HIP(hipMemcpyAsync(&h.upper_total, gr->uro + n, sizeof(unsigned), hipMemcpyDeviceToHost, stream));
    PM_HIP(hipStreamSynchronize(stream));
    if (h.flags[0] != 0xffffffffu)
        return fail(PM_RUNTIME_ERROR, "learned preconditioner: row %u has a missing or non-positive diagonal",
                    h.flags[0]);
    gr->nl = h.lower_total;
    if (h.flags[1] != 0 || h.flags[2] != 0 || h.upper_total != h.lower_total)
    {
        // A's pattern is not symmetric (or its rows are not sorted): G^T from
        // the (col, row)-sorted entries of G
        const int64_t hist = radix_hist_words(cap);
        const size_t sbytes = sizeof(unsigned long long) * cap * 2 + sizeof(unsigned) * cap * 2 +
                              sizeof(unsigned) * (hist + scan_scratch_words(std::max(hist, cap + n + 2))) + 4096;
        PM_TRY(sort_tmp.reserve(sbytes));
        unsigned long long* keys_a = sort_tmp.as<unsigned long long>();
        unsigned long long* keys_b = keys_a + cap;
        unsigned* vals_a = reinterpret_cast<unsigned*>(keys_b + cap);
        unsigned* vals_b = vals_a + cap;
        unsigned* hist_buf = vals_b + cap;
        unsigned* sort_scratch = hist_buf + hist;
        hipLaunchKernelGGL(k_graph_lower, dim3(g), dim3(kBlock), 0, stream, n, gr->ro, gr->ci, lro_u, gr->lro,
                           gr->lci, gr->lsrc, gr->ldiag, keys_a, vals_a, rb);
        PM_HIP(hipGetLastError());
        bool in_b = false;
        PM_TRY(radix_sort_pairs(keys_a, vals_a, keys_b, vals_b, gr->nl, 2 * rb, hist_buf, sort_scratch, stream,
                                &in_b));
        const unsigned long long* sk = in_b ? keys_b : keys_a;
        const unsigned* sv = in_b ? vals_b : vals_a;
        hipLaunchKernelGGL(k_graph_upper, dim3(grid_of(gr->nl)), dim3(kBlock), 0, stream, gr->nl, sk, sv, rb,
                           gr->uci, gr->uperm);
        hipLaunchKernelGGL(k_row_starts, dim3(grid_of(n + 1)), dim3(kBlock), 0, stream, n, gr->nl, sk, rb, gr->uro);
    }
    hipLaunchKernelGGL(k_widen, dim3(grid_of(n + 1)), dim3(kBlock), 0, stream, n + 1, gr->lro, gr->lro64);
    hipLaunchKernelGGL(k_widen, dim3(grid_of(n + 1)), dim3(kBlock), 0, stream, n + 1, gr->uro, gr->uro64);
    hipLaunchKernelGGL(k_widen, dim3(grid_of(gr->nl)), dim3(kBlock), 0, stream, gr->nl, gr->lci, gr->lci64);
    hipLaunchKernelGGL(k_widen, dim3(grid_of(gr->nl)), dim3(kBlock), 0, stream, gr->nl, gr->uci, gr->uci64);
    PM_HIP(hipGetLastError());
    size_t abytes = 0;
    carve_act(gr, &abytes);
    PM_TRY(gr->act.reserve(abytes));
    PM_TRY(gr->parts.reserve(sizeof(float) * kPartBlocks * (2 * kHid * kHid + kDecParams + kHid) + 4096));
    PM_HIP(hipStreamSynchronize(stream));
    return PM_OK;
}

// keep_m stores every layer's aggregation M for the backward pass.
int net_forward(pm_fsai_net* net, pm_fsai_graph* gr, const Act& a, hipStream_t stream, bool keep_m)
{
    const int64_t n = gr->n;
    const int gw = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(ceil_div(n, 16), 4), 2048)));
    const int gr_rows = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 32), 16384)));
    const float* P = net->params.as<float>();
    const __bf16* wt = net->wt.as<__bf16>();
    hipLaunchKernelGGL((k_layer<kFeat, false, false>), dim3(gw), dim3(kBlock), 0, stream, n, nullptr, nullptr, nullptr,
                       gr->x, wt, P + kOffBin, a.h[0], nullptr);
    for (int l = 1; l <= kLayers; ++l)
    {
        const __bf16* wtl = wt + kFeat * kHid + (l - 1) * 2 * kHid * kHid;
        if (keep_m)
            hipLaunchKernelGGL((k_layer<2 * kHid, true, true>), dim3(gw), dim3(kBlock), 0, stream, n, gr->ro, gr->ci,
                               gr->s, a.h[l - 1], wtl, P + kOffBl + (l - 1) * kHid, a.h[l], a.m[l]);
        else
            hipLaunchKernelGGL((k_layer<2 * kHid, true, false>), dim3(gw), dim3(kBlock), 0, stream, n, gr->ro, gr->ci,
                               gr->s, a.h[l - 1], wtl, P + kOffBl + (l - 1) * kHid, a.h[l], nullptr);
    }
    if (gr->dtype == PM_F32)
        hipLaunchKernelGGL(k_decode<float>, dim3(gr_rows), dim3(kBlock), 0, stream, n, gr->lro, gr->lci, gr->lsrc,
                           gr->s, gr->dinv, a.h[kLayers], P, gr->out, static_cast<float*>(gr->g));
    else
        hipLaunchKernelGGL(k_decode<double>, dim3(gr_rows), dim3(kBlock), 0, stream, n, gr->lro, gr->lci, gr->lsrc,
                           gr->s, gr->dinv, a.h[kLayers], P, gr->out, static_cast<double*>(gr->g));
    const int gl = grid_of(gr->nl);
    if (gr->dtype == PM_F32)
        hipLaunchKernelGGL(k_gather_upper<float>, dim3(gl), dim3(kBlock), 0, stream, gr->nl, gr->uperm,
                           static_cast<const float*>(gr->g), static_cast<float*>(gr->gt));
    else
        hipLaunchKernelGGL(k_gather_upper<double>, dim3(gl), dim3(kBlock), 0, stream, gr->nl, gr->uperm,
                           static_cast<const double*>(gr->g), static_cast<double*>(gr->gt));
    PM_HIP(hipGetLastError());
    return PM_OK;
}

// Loss and full parameter gradient into net->grad.
int net_loss_grad(pm_fsai_net* net, pm_fsai_graph* gr, const Act& a, int probes, unsigned long long seed,
                  hipStream_t stream)
{
    const int64_t n = gr->n;
    const int gp = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, kBlock / kMaxProbes), 16384)));
    const int gn = grid_of(n * probes);
    hipLaunchKernelGGL(k_probes, dim3(gn), dim3(kBlock), 0, stream, n, probes, seed, a.z);
    auto spmm = [&](const int32_t* ro, const int32_t* ci, const void* val, bool graph_dtype_g, const float* x, float* y) {
        if (graph_dtype_g ? gr->dtype == PM_F32 : gr->dtype == PM_F32)
            hipLaunchKernelGGL(k_spmm8<float>, dim3(gp), dim3(kBlock), 0, stream, n, probes, ro, ci,
                               static_cast<const float*>(val), x, y);
        else
            hipLaunchKernelGGL(k_spmm8<double>, dim3(gp), dim3(kBlock), 0, stream, n, probes, ro, ci,
                               static_cast<const double*>(val), x, y);
    };
    spmm(gr->uro, gr->uci, gr->gt, true, a.z, a.u);      // u = G^T z
    spmm(gr->ro, gr->ci, gr->values, false, a.u, a.v);   // v = A u
    spmm(gr->lro, gr->lci, gr->g, true, a.v, a.w);       // w = G v
    double* dpart = reinterpret_cast<double*>(gr->parts.as<char>());
    hipLaunchKernelGGL(k_residual_probes, dim3(kPartBlocks), dim3(kBlock), 0, stream, n, probes, a.z, a.w, a.r, dpart);
    hipLaunchKernelGGL(k_loss, dim3(1), dim3(64), 0, stream, dpart, kPartBlocks, probes, a.stat);
    hipLaunchKernelGGL(k_scale_probes, dim3(gn), dim3(kBlock), 0, stream, n, probes, a.stat, a.r);  // r := dL/dw
    spmm(gr->uro, gr->uci, gr->gt, true, a.r, a.q);      // q = G^T dL/dw = dL/dv
    spmm(gr->ro, gr->ci, gr->values, false, a.q, a.u);   // u := A q = dL/du
    hipLaunchKernelGGL(k_grad_out, dim3(grid_of(n)), dim3(kBlock), 0, stream, n, probes, gr->lro, gr->lci, gr->dinv,
                       gr->out, a.r, a.v, a.z, a.u, a.dout);
    PM_HIP(hipGetLastError());

    float* grad = net->grad.as<float>();
    float* part = gr->parts.as<float>();
    const float* P = net->params.as<float>();
    const __bf16* w = net->w.as<__bf16>();
    const int gr_rows = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 4), 16384)));
    const int gw = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(ceil_div(n, 16), 4), 8192)));
    hipLaunchKernelGGL(k_decode_back, dim3(kPartBlocks), dim3(kBlock), 0, stream, n, gr->lro, gr->lci, gr->lsrc,
                       gr->uro, gr->uci, gr->uperm, gr->s, a.h[kLayers], P, a.dout, a.dh, part);
    hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(kBlock), 0, stream, part, kPartBlocks, kDecParams,
                       grad + kOffWoff);
    for (int l = kLayers; l >= 1; --l)
    {
        hipLaunchKernelGGL(k_relu_back, dim3(kPartBlocks), dim3(kBlock), 0, stream, n, a.dh, a.h[l], a.dp, part);
        hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(kBlock), 0, stream, part, kPartBlocks, kHid,
                           grad + kOffBl + (l - 1) * kHid);
        hipLaunchKernelGGL(k_gemm_wgrad<2 * kHid>, dim3(kPartBlocks), dim3(kBlock), 0, stream, n, a.h[l - 1], a.m[l],
                           a.dp, part);
        hipLaunchKernelGGL(k_reduce_parts, dim3(grid_of(2 * kHid * kHid)), dim3(kBlock), 0, stream, part,
                           kPartBlocks, 2 * kHid * kHid, grad + kOffWl + (l - 1) * 2 * kHid * kHid);
        // [dH, dM] = dP W_l^T: the GEMM's B^T operand is W_l itself (2kHid x kHid row-major)
        const __bf16* wl = w + kFeat * kHid + (l - 1) * 2 * kHid * kHid;
        hipLaunchKernelGGL((k_gemm<kHid, 2 * kHid>), dim3(gw), dim3(kBlock), 0, stream, n, a.dp, nullptr, wl, a.y);
        hipLaunchKernelGGL(k_split_grad, dim3(grid_of(n * kHid)), dim3(kBlock), 0, stream, n, a.y, a.dh, a.dm);
        hipLaunchKernelGGL(k_aggregate_back, dim3(gr_rows), dim3(kBlock), 0, stream, n, gr->ro, gr->ci, gr->s, a.dm,
                           a.dh);
    }
    hipLaunchKernelGGL(k_relu_back, dim3(kPartBlocks), dim3(kBlock), 0, stream, n, a.dh, a.h[0], a.dp, part);
    hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(kBlock), 0, stream, part, kPartBlocks, kHid, grad + kOffBin);
    hipLaunchKernelGGL(k_gemm_wgrad<kFeat>, dim3(kPartBlocks), dim3(kBlock), 0, stream, n, gr->x, nullptr, a.dp, part);
    hipLaunchKernelGGL(k_reduce_parts, dim3(grid_of(kFeat * kHid)), dim3(kBlock), 0, stream, part, kPartBlocks,
                       kFeat * kHid, grad + kOffWin);
    PM_HIP(hipGetLastError());
    return PM_OK;
}

} // namespace pm

extern "C" {

int32_t pm_fsai_parameter_count(void);
int pm_fsai_net_set_parameters(pm_fsai_net* net, const float* host, int64_t count);
int pm_fsai_net_destroy(pm_fsai_net* net);

// Hidden layers: N(0, 2 / fan_in) (He); encoder bias 0.1 so ReLUs start
// active; decoder zero except softplus(b_diag) = 1, which makes G = D^-1/2.
int pm_fsai_initial_parameters(uint64_t seed, float* p, int64_t count)
{
    if (!p || count != kParams)
        return fail(PM_INVALID_ARGUMENT, "pm_fsai_initial_parameters: expected %d parameters", kParams);
    unsigned long long state = seed * 0x9e3779b97f4a7c15ull + 0x2545f4914f6cdd1dull;
    auto next = [&]() {
        state += 0x9e3779b97f4a7c15ull;
        unsigned long long z = state;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        return (static_cast<double>(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    auto normal = [&]() { return std::sqrt(-2.0 * std::log(next())) * std::cos(6.283185307179586 * next()); };
    for (int i = 0; i < kParams; ++i) p[i] = 0.f;
    for (int i = 0; i < kFeat * kHid; ++i) p[kOffWin + i] = static_cast<float>(normal() * std::sqrt(2.0 / kFeat));
    for (int i = 0; i < kHid; ++i) p[kOffBin + i] = 0.1f;
    for (int i = 0; i < kLayers * 2 * kHid * kHid; ++i)
        p[kOffWl + i] = static_cast<float>(normal() * std::sqrt(2.0 / (2 * kHid)));
    for (int i = 0; i < kLayers * kHid; ++i) p[kOffBl + i] = 0.1f;
    p[kOffBdiag] = static_cast<float>(std::log(std::exp(1.0) - 1.0));
    return PM_OK;
}

int pm_fsai_net_create(uint64_t seed, pm_fsai_net** out)
{
    if (!out) return fail(PM_INVALID_ARGUMENT, "pm_fsai_net_create: null output");
    auto* net = new (std::nothrow) pm_fsai_net();
    if (!net) return fail(PM_RUNTIME_ERROR, "pm_fsai_net_create: out of host memory");
    int st = net->params.reserve(sizeof(float) * kParams);
    if (st == PM_OK) st = net->adam_m.reserve(sizeof(float) * kParams);
    if (st == PM_OK) st = net->adam_v.reserve(sizeof(float) * kParams);
    if (st == PM_OK) st = net->grad.reserve(sizeof(float) * kParams);
    if (st == PM_OK) st = net->w.reserve(sizeof(__bf16) * (kFeat * kHid + kLayers * 2 * kHid * kHid));
    if (st == PM_OK) st = net->wt.reserve(sizeof(__bf16) * (kFeat * kHid + kLayers * 2 * kHid * kHid));
    if (st == PM_OK)
    {
        std::vector<float> init(kParams);
        st = pm_fsai_initial_parameters(seed, init.data(), kParams);
        if (st == PM_OK) st = pm_fsai_net_set_parameters(net, init.data(), kParams);
    }
    if (st != PM_OK)
    {
        pm_fsai_net_destroy(net);
        return st;
    }
    *out = net;
    return PM_OK;
}

int pm_fsai_net_destroy(pm_fsai_net* net)
{
    if (!net) return PM_OK;
    (void)hipDeviceSynchronize();
    net->params.release();
    net->adam_m.release();
    net->adam_v.release();
    net->grad.release();
    net->w.release();
    net->wt.release();
    delete net;
    return PM_OK;
}

int32_t pm_fsai_parameter_count(void) { return kParams; }

// Host parameters in; Adam moments reset.
int pm_fsai_net_set_parameters(pm_fsai_net* net, const float* host, int64_t count)
{
    if (!net || !host || count != kParams)
        return fail(PM_INVALID_ARGUMENT, "pm_fsai_net_set_parameters: expected %d parameters", kParams);
    PM_HIP(hipMemcpy(net->params.ptr, host, sizeof(float) * kParams, hipMemcpyHostToDevice));
    PM_HIP(hipMemset(net->adam_m.ptr, 0, sizeof(float) * kParams));
    PM_HIP(hipMemset(net->adam_v.ptr, 0, sizeof(float) * kParams));
    hipLaunchKernelGGL(k_weights_bf16, dim3(64), dim3(kBlock), 0, nullptr, net->params.as<float>(), net->w.as<__bf16>(),
                       net->wt.as<__bf16>());
    PM_HIP(hipGetLastError());
    PM_HIP(hipDeviceSynchronize());
    net->steps = 0;
    return PM_OK;
}

int pm_fsai_net_get_parameters(const pm_fsai_net* net, float* host, int64_t count)
{
    if (!net || !host || count != kParams)
        return fail(PM_INVALID_ARGUMENT, "pm_fsai_net_get_parameters: expected %d parameters", kParams);
    PM_HIP(hipDeviceSynchronize());
    PM_HIP(hipMemcpy(host, net->params.ptr, sizeof(float) * kParams, hipMemcpyDeviceToHost));
    return PM_OK;
}

namespace
{

// Device storage of destroyed graphs, kept for the next graph on the same
// device: a caller that builds a graph per system (a new sparsity pattern for
// every solve) would otherwise pay hipMalloc and hipFree -- whose device-wide
// synchronisation costs more than a small system's whole graph build -- for
// each of a graph's four buffers.  At most kPooledGraphs sets of at most
// kPooledGraphBytes each are kept, until pm_fsai_graph_pool_release or the
// end of the process; PLAMATRIX_AMD_GRAPH_POOL=0 turns the pool off.
constexpr size_t kPooledGraphs = 4;
constexpr size_t kPooledGraphBytes = size_t(1) << 30;

bool pool_enabled()
{
    static const bool on = [] {
        const char* env = std::getenv("PLAMATRIX_AMD_GRAPH_POOL");
        return env == nullptr || env[0] != '0';
    }();
    return on;
}

struct GraphStorage
{
    int device = 0;
    pm::DeviceBuffer buf, act, parts, tmp;
};

std::mutex g_graph_pool_mutex;
std::vector<GraphStorage> g_graph_pool;

void take_pooled_storage(pm_fsai_graph* gr)
{
    std::lock_guard<std::mutex> lock(g_graph_pool_mutex);
    for (size_t i = 0; i < g_graph_pool.size(); ++i)
        if (g_graph_pool[i].device == gr->device)
        {
            gr->buf = g_graph_pool[i].buf;
            gr->act = g_graph_pool[i].act;
            gr->parts = g_graph_pool[i].parts;
            gr->tmp = g_graph_pool[i].tmp;
            g_graph_pool.erase(g_graph_pool.begin() + static_cast<std::ptrdiff_t>(i));
            return;
        }
}

void release_storage(pm_fsai_graph* gr)
{
    const size_t total = gr->buf.bytes + gr->act.bytes + gr->parts.bytes + gr->tmp.bytes;
    if (pool_enabled())
    {
        std::lock_guard<std::mutex> lock(g_graph_pool_mutex);
        if (g_graph_pool.size() < kPooledGraphs && total <= kPooledGraphBytes)
        {
            g_graph_pool.push_back(GraphStorage{gr->device, gr->buf, gr->act, gr->parts, gr->tmp});
            gr->buf = gr->act = gr->parts = gr->tmp = pm::DeviceBuffer{};
            return;
        }
    }
    gr->buf.release();
    gr->act.release();
    gr->parts.release();
    gr->tmp.release();
}

} // namespace

int pm_fsai_graph_create(const pm_csr* a, pm_fsai_graph** out, hipStream_t stream)
{
    PM_TRY(check_csr(a, "learned preconditioner"));
    if (!out) return fail(PM_INVALID_ARGUMENT, "pm_fsai_graph_create: null output");
    if (a->rows != a->cols) return fail(PM_INVALID_ARGUMENT, "learned preconditioner: matrix must be square");
    if (!a->row_offsets32 || !a->col_indices32)
        return fail(PM_INVALID_ARGUMENT, "learned preconditioner: the matrix needs its 32-bit index mirrors");
    if (a->rows < 1) return fail(PM_INVALID_ARGUMENT, "learned preconditioner: empty matrix");
    auto* gr = new (std::nothrow) pm_fsai_graph();
    if (!gr) return fail(PM_RUNTIME_ERROR, "pm_fsai_graph_create: out of host memory");
    if (hipGetDevice(&gr->device) != hipSuccess)
    {
        (void)hipGetLastError();
        delete gr;
        return fail(PM_RUNTIME_ERROR, "pm_fsai_graph_create: no current HIP device");
    }
    take_pooled_storage(gr);
    const int st = graph_build(gr, a, stream);
    if (st != PM_OK)
    {
        (void)hipStreamSynchronize(stream);
        release_storage(gr);
        delete gr;
        return st;
    }
    *out = gr;
    return PM_OK;
}

int pm_fsai_graph_destroy(pm_fsai_graph* gr)
{
    if (!gr) return PM_OK;
    // forward, training and factored solves may have used this graph's
    // storage on any stream; the next graph may reuse it at once, so every
    // stream of the device must have drained first
    int current = 0;
    const bool switched = hipGetDevice(&current) == hipSuccess && current != gr->device &&
                          hipSetDevice(gr->device) == hipSuccess;
    (void)hipDeviceSynchronize();
    release_storage(gr);
    if (switched) (void)hipSetDevice(current);
    delete gr;
    return PM_OK;
}

// Frees the storage that destroyed graphs left for reuse on the current device.
int pm_fsai_graph_pool_release(void)
{
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess)
    {
        (void)hipGetLastError();
        return fail(PM_RUNTIME_ERROR, "pm_fsai_graph_pool_release: no current HIP device");
    }
    std::vector<GraphStorage> mine;
    {
        std::lock_guard<std::mutex> lock(g_graph_pool_mutex);
        for (size_t i = 0; i < g_graph_pool.size();)
            if (g_graph_pool[i].device == device)
            {
                mine.push_back(g_graph_pool[i]);
                g_graph_pool.erase(g_graph_pool.begin() + static_cast<std::ptrdiff_t>(i));
            }
            else ++i;
    }
    for (GraphStorage& st : mine)
    {
        st.buf.release();
        st.act.release();
        st.parts.release();
        st.tmp.release();
    }
    return PM_OK;
}

// Bytes of device storage the pool holds on the current device.
int pm_fsai_graph_pool_bytes(int64_t* bytes)
{
    if (!bytes) return fail(PM_INVALID_ARGUMENT, "pm_fsai_graph_pool_bytes: null output");
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess)
    {
        (void)hipGetLastError();
        return fail(PM_RUNTIME_ERROR, "pm_fsai_graph_pool_bytes: no current HIP device");
    }
    std::lock_guard<std::mutex> lock(g_graph_pool_mutex);
    int64_t total = 0;
    for (const GraphStorage& st : g_graph_pool)
        if (st.device == device)
            total += static_cast<int64_t>(st.buf.bytes + st.act.bytes + st.parts.bytes + st.tmp.bytes);
    *bytes = total;
    return PM_OK;
}

// G (lower, row order) and G^T (upper) as CSR descriptors in A's dtype.
int pm_fsai_factors(const pm_fsai_graph* gr, pm_csr* g, pm_csr* gt)
{
    if (!gr || !g || !gt) return fail(PM_INVALID_ARGUMENT, "pm_fsai_factors: null argument");
    *g = pm_csr{gr->n, gr->n, gr->nl, gr->lro64, gr->lci64, gr->g, gr->dtype, gr->lro, gr->lci};
    *gt = pm_csr{gr->n, gr->n, gr->nl, gr->uro64, gr->uci64, gr->gt, gr->dtype, gr->uro, gr->uci};
    return PM_OK;
}

int pm_fsai_batch_select(pm_fsai_graph* gr, int64_t systems, const int64_t* row_begin, const int64_t* row_end,
                         int32_t probes, uint64_t seed, uint8_t* jacobi_systems, double* losses, hipStream_t stream)
{
    if (!gr || systems < 0 || (systems > 0 && (!row_begin || !row_end || !jacobi_systems)))
        return fail(PM_INVALID_ARGUMENT, "pm_fsai_batch_select: null argument");
    if (probes < 1 || probes > kMaxProbes) return fail(PM_INVALID_ARGUMENT, "probes must be in [1, %d]", kMaxProbes);
    if (systems == 0) return PM_OK;
    const int64_t n = gr->n;
    std::vector<int64_t> range(static_cast<size_t>(2 * systems));
    for (int64_t s = 0; s < systems; ++s)
    {
        if (row_begin[s] < 0 || row_end[s] < row_begin[s] || row_end[s] > n)
            return fail(PM_INVALID_ARGUMENT, "pm_fsai_batch_select: system %lld row range [%lld, %lld) is outside "
                        "the %lld-row graph", static_cast<long long>(s), static_cast<long long>(row_begin[s]),
                        static_cast<long long>(row_end[s]), static_cast<long long>(n));
        range[static_cast<size_t>(s)] = row_begin[s];
        range[static_cast<size_t>(systems + s)] = row_end[s];
    }
    PM_TRY(gr->select.reserve(sizeof(int64_t) * 2 * systems + sizeof(double) * 2 * systems + 256));
    int64_t* drange = gr->select.as<int64_t>();
    double* dloss = reinterpret_cast<double*>(drange + 2 * systems);
    PM_HIP(hipMemcpyAsync(drange, range.data(), sizeof(int64_t) * 2 * systems, hipMemcpyHostToDevice, stream));
    size_t bytes = 0;
    const Act a = carve_act(gr, &bytes);
    double* e = reinterpret_cast<double*>(a.q);  // n x 8 floats hold the n x 2 doubles
    const int gp = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, kBlock / kMaxProbes), 16384)));
    const int gn = grid_of(n * probes);
    auto spmm = [&](const int32_t* ro, const int32_t* ci, const void* val, const float* x, float* y) {
        if (gr->dtype == PM_F32)
            hipLaunchKernelGGL(k_spmm8<float>, dim3(gp), dim3(kBlock), 0, stream, n, probes, ro, ci,
                               static_cast<const float*>(val), x, y);
        else
            hipLaunchKernelGGL(k_spmm8<double>, dim3(gp), dim3(kBlock), 0, stream, n, probes, ro, ci,
                               static_cast<const double*>(val), x, y);
    };
    hipLaunchKernelGGL(k_probes, dim3(gn), dim3(kBlock), 0, stream, n, probes, static_cast<unsigned long long>(seed),
                       a.z);
    spmm(gr->uro, gr->uci, gr->gt, a.z, a.u);     // u = G^T z
    spmm(gr->ro, gr->ci, gr->values, a.u, a.v);   // v = A u
    spmm(gr->lro, gr->lci, gr->g, a.v, a.w);      // w = G v
    hipLaunchKernelGGL(k_select_rows<false>, dim3(grid_of(n)), dim3(kBlock), 0, stream, n, probes, a.z, a.w,
                       gr->dinv, e);
    hipLaunchKernelGGL(k_scale_probe_rows, dim3(gn), dim3(kBlock), 0, stream, n, probes, gr->dinv, a.z, a.u);
    spmm(gr->ro, gr->ci, gr->values, a.u, a.v);   // v = A D^-1/2 z
    hipLaunchKernelGGL(k_select_rows<true>, dim3(grid_of(n)), dim3(kBlock), 0, stream, n, probes, a.z, a.v,
                       gr->dinv, e);
    const int gs = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(systems, kBlock / kWave), 1024)));
    hipLaunchKernelGGL(k_select_systems, dim3(gs), dim3(kBlock), 0, stream, systems, drange, e, dloss,
                       jacobi_systems);
    PM_HIP(hipGetLastError());
    if (losses)
    {
        PM_HIP(hipMemcpyAsync(losses, dloss, sizeof(double) * 2 * systems, hipMemcpyDeviceToHost, stream));
        PM_HIP(hipStreamSynchronize(stream));
    }
    else
    {
        // the host range array goes out of scope
        PM_HIP(hipStreamSynchronize(stream));
    }
    return PM_OK;
}

int pm_fsai_forward(pm_fsai_net* net, pm_fsai_graph* gr, hipStream_t stream)
{
    if (!net || !gr) return fail(PM_INVALID_ARGUMENT, "pm_fsai_forward: null argument");
    size_t bytes = 0;
    const Act a = carve_act(gr, &bytes);
    return net_forward(net, gr, a, stream, false);
}

int pm_fsai_loss_grad(pm_fsai_net* net, pm_fsai_graph* gr, int32_t probes, uint64_t seed, double* loss,
                      float* host_grad, hipStream_t stream)
{
    if (!net || !gr) return fail(PM_INVALID_ARGUMENT, "pm_fsai_loss_grad: null argument");
    if (probes < 1 || probes > kMaxProbes) return fail(PM_INVALID_ARGUMENT, "probes must be in [1, %d]", kMaxProbes);
    size_t bytes = 0;
    const Act a = carve_act(gr, &bytes);
    PM_TRY(net_forward(net, gr, a, stream, true));
    PM_TRY(net_loss_grad(net, gr, a, probes, seed, stream));
    if (loss) PM_HIP(hipMemcpyAsync(loss, a.stat, sizeof(double), hipMemcpyDeviceToHost, stream));
    if (host_grad)
        PM_HIP(hipMemcpyAsync(host_grad, net->grad.ptr, sizeof(float) * kParams, hipMemcpyDeviceToHost, stream));
    PM_HIP(hipStreamSynchronize(stream));
    return PM_OK;
}

// One Adam step on the probe loss of this graph (8 probes at most).
int pm_fsai_train_step(pm_fsai_net* net, pm_fsai_graph* gr, int32_t probes, uint64_t seed, float lr, double* loss,
                       hipStream_t stream)
{
    if (!net || !gr) return fail(PM_INVALID_ARGUMENT, "pm_fsai_train_step: null argument");
    if (probes < 1 || probes > kMaxProbes) return fail(PM_INVALID_ARGUMENT, "probes must be in [1, %d]", kMaxProbes);
    if (!(lr > 0.f) || !std::isfinite(lr)) return fail(PM_INVALID_ARGUMENT, "learning rate must be positive");
    size_t bytes = 0;
    const Act a = carve_act(gr, &bytes);
    PM_TRY(net_forward(net, gr, a, stream, true));
    PM_TRY(net_loss_grad(net, gr, a, probes, seed, stream));
    net->steps += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f / (1.f - std::pow(b1, static_cast<float>(net->steps)));
    const float c2 = 1.f / (1.f - std::pow(b2, static_cast<float>(net->steps)));
    hipLaunchKernelGGL(k_adam, dim3(grid_of(kParams)), dim3(kBlock), 0, stream, kParams, net->params.as<float>(),
                       net->adam_m.as<float>(), net->adam_v.as<float>(), net->grad.as<float>(), lr, b1, b2, 1e-8f, c1,
                       c2);
    hipLaunchKernelGGL(k_weights_bf16, dim3(64), dim3(kBlock), 0, stream, net->params.as<float>(), net->w.as<__bf16>(),
                       net->wt.as<__bf16>());
    PM_HIP(hipGetLastError());
    if (loss)
    {
        PM_HIP(hipMemcpyAsync(loss, a.stat, sizeof(double), hipMemcpyDeviceToHost, stream));
        PM_HIP(hipStreamSynchronize(stream));
    }
    return PM_OK;
}

} // extern "C"
