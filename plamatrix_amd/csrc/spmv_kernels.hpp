// CSR SpMV family for gfx950.
//
// Wave-walk rows: a wavefront owns 64 consecutive rows (lane i <-> row i).
// It loads the 65 row offsets with one coalesced load and stages the chunk's
// contiguous values and column indices in LDS with back-to-back coalesced
// loads.  Then every lane walks its own row left to right out of LDS.  For
// banded and stencil matrices, entry j of 64 consecutive rows points at
// nearly consecutive columns, so each x gather of the wave touches a few
// cache lines instead of one line per lane; that gather cost, not HBM, is
// what bounds a lane-parallel CSR SpMV on this part.  Chunks of more than
// kMaxWindows * kStreamCap entries (long rows) fall back to 8 lanes per row
// with a fixed butterfly.  Workgroups own contiguous chunk ranges and are
// renumbered so each XCD walks a contiguous eighth of the matrix, keeping
// the gathered x entries in its own L2.
//
// Reductions across workgroups never use atomics or device-scope fences (on a
// multi-XCD part such a fence writes back the whole L2): each workgroup
// publishes its FP64 partial sums, and every workgroup of the *next* kernel
// re-reduces all of them in the same fixed order.  The result is identical in
// every workgroup and from run to run.
#pragma once

#include <cmath>

#include "pm_common.hpp"
#include "workspace.hpp"

namespace pm
{

constexpr int kChunkRows = 64;               // rows per wavefront chunk
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kStreamCap = 512;              // entries staged per wave and window (LDS)
constexpr int kMaxWindows = 4;               // windows per chunk before the long-row fallback
constexpr int kWalkUnroll = 8;               // independent gathers per lane and step
constexpr int kSolverGrid = 1024;            // workgroups of the solver kernels (partials per reduction)

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

// Load of a streamed-once array element, optionally non-temporal (the line is
// not kept in the caches).
template <typename T>
__device__ __forceinline__ T stream_load(const T* p, bool nt)
{
    return nt ? __builtin_nontemporal_load(p) : *p;
}

// Contiguous range [begin, end) of `units` owned by this workgroup
// (XCD-aware), rounded to multiples of `step`.
__device__ __forceinline__ void block_range(int64_t units, int64_t step, int64_t& begin, int64_t& end)
{
    const int64_t nb = gridDim.x;
    const int64_t lb = xcd_logical_block(blockIdx.x, nb);
    const int64_t per = ceil_div(ceil_div(units, nb), step) * step;
    begin = lb * per;
    end = begin + per < units ? begin + per : units;
}

// Chunks are dealt to wavefronts round-robin over the whole grid: wave w of
// workgroup b takes chunks b*W + w, b*W + w + grid*W, ...  Neighbouring
// chunks run concurrently on different CUs, which keeps the HBM pages and the
// gathered x lines of one region hot at the same time.
#define PM_FOR_CHUNKS(c, chunks, W)                                                                  \
    for (int64_t c = static_cast<int64_t>(blockIdx.x) * (W) + threadIdx.x / kWave; c < (chunks);       \
         c += static_cast<int64_t>(gridDim.x) * (W))

constexpr int kSpmvGrid = 2048;  // workgroups of the standalone SpMV (8 per CU)

__host__ __device__ inline int64_t chunk_count(int64_t rows) { return ceil_div(rows > 0 ? rows : 1, kChunkRows); }

// Grid of a kernel that publishes partial sums: at most kSolverGrid workgroups.
inline int solver_grid_chunks(int64_t rows)
{
    const int64_t g = ceil_div(chunk_count(rows), kWavesPerBlock);
    return static_cast<int>(g < kSolverGrid ? g : kSolverGrid);
}

inline int solver_grid_units(int64_t units)
{
    const int64_t g = ceil_div(units > 0 ? units : 1, 4 * kBlock);
    return static_cast<int>(g < kSolverGrid ? g : kSolverGrid);
}

constexpr int kStageLoads = kStreamCap / kWave;  // staging loads per lane and window

// Entry sources of chunk_row_sums: value(k) loads the matrix entry, gather(c)
// the vector entry of column c, madd(a, g, acc) accumulates a * g in the
// accumulation type.
template <typename S>
struct SpmvSource
{
    const S* __restrict__ v;
    const S* __restrict__ x;
    using V = S;
    using G = S;
    __device__ __forceinline__ S value(int64_t k, bool nt) const { return stream_load(v + k, nt); }
    template <typename I>
    __device__ __forceinline__ S gather(I c) const
    {
        return x[c];
    }
    __device__ __forceinline__ S madd(S a, S g, S acc) const { return fma_s(a, g, acc); }
};

// Entries of A times x accumulated in FP64 (x of type X).
template <typename S, typename X>
struct Fp64Source
{
    const S* __restrict__ v;
    const X* __restrict__ x;
    using V = S;
    using G = X;
    __device__ __forceinline__ S value(int64_t k, bool nt) const { return stream_load(v + k, nt); }
    template <typename I>
    __device__ __forceinline__ X gather(I c) const
    {
        return x[c];
    }
    __device__ __forceinline__ double madd(S a, X g, double acc) const
    {
        return __builtin_fma(static_cast<double>(a), static_cast<double>(g), acc);
    }
};

// Gathered entries of p_new = z + beta p_old, formed on the fly.
template <typename S>
struct DirectionSource
{
    const S* __restrict__ v;
    const S* __restrict__ z;
    const S* __restrict__ p_old;
    S beta;
    using V = S;
    using G = S;
    __device__ __forceinline__ S value(int64_t k, bool nt) const { return stream_load(v + k, nt); }
    template <typename I>
    __device__ __forceinline__ S gather(I c) const
    {
        return fma_s(beta, p_old[c], z[c]);
    }
    __device__ __forceinline__ S madd(S a, S g, S acc) const { return fma_s(a, g, acc); }
};

// Gathered entries of r_new = r - alpha q, formed on the fly (neg_alpha = -alpha).
template <typename S>
struct StepSource
{
    const S* __restrict__ v;
    const S* __restrict__ r;
    const S* __restrict__ q;
    S neg_alpha;
    using V = S;
    using G = S;
    __device__ __forceinline__ S value(int64_t k, bool nt) const { return stream_load(v + k, nt); }
    template <typename I>
    __device__ __forceinline__ S gather(I c) const
    {
        return fma_s(neg_alpha, q[c], r[c]);
    }
    __device__ __forceinline__ S madd(S a, S g, S acc) const { return fma_s(a, g, acc); }
};

// Value of lane `lane_id` in every lane (v_readlane: no LDS traffic).
template <typename T>
__device__ __forceinline__ T lane_value(T x, int lane_id)
{
    if constexpr (sizeof(T) == 8)
    {
        const unsigned long long u = static_cast<unsigned long long>(x);
        const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(u & 0xffffffffull), lane_id));
        const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(u >> 32), lane_id));
        return static_cast<T>((static_cast<unsigned long long>(hi) << 32) | lo);
    }
    else
    {
        return static_cast<T>(__builtin_amdgcn_readlane(static_cast<int>(x), lane_id));
    }
}

// Chunks whose entries do not fit one LDS window: up to kMaxWindows windows
// walked like the single-window path, else L lanes per row (L = 8 by
// default; 16, 32 or 64 give each lane group longer contiguous runs of a
// long row) with a fixed butterfly.  Inlined on purpose: an out-of-line call
// gives the kernel a stack (scratch) and ~70 VGPRs, which halves the
// throughput of a wavefront that loops over many chunks.
template <typename Acc, typename I, typename Src, bool NT, int L = 8>
__device__ __forceinline__ Acc chunk_row_sums_long(int64_t base, int64_t total, int64_t my_s, int64_t my_e,
                                                const I* __restrict__ ci, typename Src::V* lv, I* lc,
                                                const Src src)
{
    const int lane = threadIdx.x % kWave;
    Acc acc = Acc(0);
    if (total <= static_cast<int64_t>(kMaxWindows) * kStreamCap)
    {
        const int s = static_cast<int>(my_s - base);
        const int e = static_cast<int>(my_e - base);
        const int n = static_cast<int>(total);
        const auto* vb = src.v + base;
        const I* cb = ci + base;
        for (int w0 = 0; w0 < n; w0 += kStreamCap)
        {
#pragma unroll
            for (int j = 0; j < kStageLoads; ++j)
            {
                const int k = lane + j * kWave;
                if (w0 + k < n)
                {
                    lv[k] = stream_load(vb + w0 + k, NT);
                    lc[k] = stream_load(cb + w0 + k, NT);
                }
            }
            wave_lds_sync();
            const int lo = (s > w0 ? s : w0) - w0;
            const int hi = (e < w0 + kStreamCap ? e : w0 + kStreamCap) - w0;
            for (int k = lo; k < hi; ++k) acc = src.madd(lv[k], src.gather(lc[k]), acc);
            wave_lds_sync();
        }
        return acc;
    }
    static_assert(L >= 2 && L <= kWave && (L & (L - 1)) == 0, "lanes per row");
    constexpr int G = kWave / L;  // rows per pass
    const int group = lane / L;
    const int sub = lane % L;
    for (int pass = 0; pass < L; ++pass)
    {
        const int row_lane = pass * G + group;
        const int64_t s = __shfl(my_s, row_lane, kWave);
        const int64_t e = __shfl(my_e, row_lane, kWave);
        Acc part = Acc(0);
        for (int64_t k = s + sub; k < e; k += L) part = src.madd(stream_load(src.v + k, NT), src.gather(ci[k]), part);
        part = group_sum<L>(part);
        const Acc got = __shfl(part, (lane % G) * L, kWave);
        if (lane / G == pass) acc = got;
    }
    return acc;
}

// Row sums of the 64-row chunk starting at row0 (lane i -> row row0 + i).
// lv / lc are this wavefront's kStreamCap-entry LDS staging arrays.  s_out /
// e_out receive the lane's entry range; rows at or beyond `rows` get an empty
// range.  Each row is accumulated left to right.
template <typename Acc, typename I, typename Src, bool NT = false, bool DMA = false, int L = 8>
__device__ __forceinline__ Acc chunk_row_sums(int64_t row0, int64_t rows, const I* __restrict__ ro,
                                              const I* __restrict__ ci, typename Src::V* lv, I* lc,
                                              const Src& src, int64_t& s_out, int64_t& e_out)
{
    const int lane = threadIdx.x % kWave;
    const int64_t last = row0 + kChunkRows < rows ? row0 + kChunkRows : rows;
    const int64_t r = row0 + lane;
    const I s_i = ro[r < last ? r : last];
    const I e_i = ro[r + 1 < last ? r + 1 : last];
    const int64_t base = static_cast<int64_t>(lane_value(s_i, 0));
    const int64_t total = static_cast<int64_t>(lane_value(e_i, kWave - 1)) - base;
    s_out = static_cast<int64_t>(s_i);
    e_out = static_cast<int64_t>(e_i);
    if (total > kStreamCap) return chunk_row_sums_long<Acc, I, Src, NT, L>(base, total, s_out, e_out, ci, lv, lc, src);
    const int n = static_cast<int>(total);
    const int s = static_cast<int>(s_out - base);
    const int e = static_cast<int>(e_out - base);
    const auto* vb = src.v + base;
    const I* cb = ci + base;
    if constexpr (DMA && sizeof(typename Src::V) == 4 && sizeof(I) == 4)
    {
        // 4-byte values and indices go straight from memory into LDS
        // (global_load_lds), without a round trip through VGPRs
#pragma unroll
        for (int j = 0; j < kStageLoads; ++j)
        {
            const int k = lane + j * kWave;
            if (k < n)
            {
                __builtin_amdgcn_global_load_lds(vb + k, lv + j * kWave, 4, 0, 0);
                __builtin_amdgcn_global_load_lds(cb + k, lc + j * kWave, 4, 0, 0);
            }
        }
        __builtin_amdgcn_s_waitcnt(0);
    }
    else
    {
#pragma unroll
        for (int j = 0; j < kStageLoads; ++j)
        {
            const int k = lane + j * kWave;
            if (k < n)
            {
                lv[k] = stream_load(vb + k, NT);
                lc[k] = stream_load(cb + k, NT);
            }
        }
    }
    wave_lds_sync();
    Acc acc = Acc(0);
    for (int k = s; k < e; k += kWalkUnroll)
    {
        typename Src::V a[kWalkUnroll];
        typename Src::G g[kWalkUnroll];
#pragma unroll
        for (int j = 0; j < kWalkUnroll; ++j)
        {
            if (k + j < e)
            {
                a[j] = lv[k + j];
                g[j] = src.gather(lc[k + j]);
            }
        }
#pragma unroll
        for (int j = 0; j < kWalkUnroll; ++j)
        {
            if (k + j < e) acc = src.madd(a[j], g[j], acc);
        }
    }
    wave_lds_sync();
    return acc;
}

// Every workgroup sums `count` (at most CAP) published partials (NV values
// each) in the same order; all threads receive the totals.  Value k's slots
// start at base[k].
template <int NV, int CAP = kMaxPartials>
__device__ __forceinline__ void reduce_partials_at(const double* const (&base)[NV], int count, double (&out)[NV])
{
    static_assert(CAP % kBlock == 0 && CAP <= kMaxPartials, "partial capacity");
    __shared__ double lds[NV * 16];
    __shared__ double result[NV];
    // every load is issued before the first add (kBlock threads, at most
    // CAP / kBlock values each); the adds run in slot order
    constexpr int kPer = CAP / kBlock;
    double vals[NV][kPer];
#pragma unroll
    for (int k = 0; k < NV; ++k)
    {
#pragma unroll
        for (int j = 0; j < kPer; ++j)
        {
            const int b = threadIdx.x + j * kBlock;
            vals[k][j] = b < count ? base[k][b] : 0.0;
        }
    }
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k)
    {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < kPer; ++j) s += vals[k][j];
        acc[k] = s;
    }
    block_sum<NV>(acc, lds);
    if (threadIdx.x == 0)
    {
#pragma unroll
        for (int k = 0; k < NV; ++k) result[k] = acc[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = result[k];
    __syncthreads();
}

// reduce_partials_at over the layout publish_partials writes: value k's slots
// at partials + k * kMaxPartials.
template <int NV>
__device__ __forceinline__ void reduce_partials(const double* __restrict__ partials, int count, double (&out)[NV])
{
    const double* base[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) base[k] = partials + k * kMaxPartials;
    reduce_partials_at<NV>(base, count, out);
}

// Publish this workgroup's NV partial sums at slot blockIdx.x.
template <int NV>
__device__ __forceinline__ void publish_partials(double (&v)[NV], double* __restrict__ partials)
{
    __shared__ double lds[NV * 16];
    block_sum<NV>(v, lds);
    if (threadIdx.x == 0)
    {
#pragma unroll
        for (int k = 0; k < NV; ++k) partials[k * kMaxPartials + blockIdx.x] = v[k];
    }
}

// SpMM (spmm, sparse_ops_cpu.cpp:244-268) with B (cols x k) and Y (rows x k)
// column-major, in tiles of kSpmmTile columns: k_spmm_pack copies the tile's
// columns of B into a row-major cols x kSpmmTile block (zero past column k),
// so an entry's operands are one contiguous 32- or 64-byte load; k_spmm_tile
// stages each 64-row chunk's values and column indices in LDS once and walks
// them for all columns of the tile.  Every (row, column) sum runs left to
// right with the SpMV's FMA order, so a column of Y equals the SpMV of that
// column of B bit for bit.  Chunks too long for one LDS window take the SpMV
// chunk routine once per column.
constexpr int kSpmmTile = 8;

template <typename S>
__global__ __launch_bounds__(kBlock) void k_spmm_pack(int64_t cols, int64_t k, int64_t j0, const S* __restrict__ b,
                                                      S* __restrict__ bt)
{
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < cols;
         i += static_cast<int64_t>(gridDim.x) * kBlock)
    {
        S row[kSpmmTile];
#pragma unroll
        for (int t = 0; t < kSpmmTile; ++t) row[t] = j0 + t < k ? b[(j0 + t) * cols + i] : S(0);
#pragma unroll
        for (int t = 0; t < kSpmmTile; ++t) bt[i * kSpmmTile + t] = row[t];
    }
}

// Column t of the packed block as an SpMV operand.
template <typename S>
struct PackedSource
{
    const S* __restrict__ v;
    const S* __restrict__ bt;
    int t;
    using V = S;
    using G = S;
    __device__ __forceinline__ S value(int64_t k, bool nt) const { return stream_load(v + k, nt); }
    template <typename I>
    __device__ __forceinline__ S gather(I c) const
    {
        return bt[static_cast<int64_t>(c) * kSpmmTile + t];
    }
    __device__ __forceinline__ S madd(S a, S g, S acc) const { return fma_s(a, g, acc); }
};

template <typename S, typename I>
__global__ __launch_bounds__(kBlock) void k_spmm_tile(int64_t rows, int64_t kt, const I* __restrict__ ro,
                                                      const I* __restrict__ ci, const S* __restrict__ v,
                                                      const S* __restrict__ bt, S* __restrict__ y)
{
    __shared__ S lv_all[kWavesPerBlock][kStreamCap];
    __shared__ I lc_all[kWavesPerBlock][kStreamCap];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    S* lv = lv_all[wave];
    I* lc = lc_all[wave];
    PM_FOR_CHUNKS(c, chunk_count(rows), kWavesPerBlock)
    {
        const int64_t row0 = c * kChunkRows;
        const int64_t last = row0 + kChunkRows < rows ? row0 + kChunkRows : rows;
        const int64_t r = row0 + lane;
        const I s_i = ro[r < last ? r : last];
        const I e_i = ro[r + 1 < last ? r + 1 : last];
        const int64_t base = static_cast<int64_t>(lane_value(s_i, 0));
        const int64_t total = static_cast<int64_t>(lane_value(e_i, kWave - 1)) - base;
        if (total > kStreamCap)
        {
            for (int t = 0; t < kt; ++t)
            {
                int64_t s, e;
                const S acc = chunk_row_sums<S>(row0, rows, ro, ci, lv, lc, PackedSource<S>{v, bt, t}, s, e);
                if (r < rows) y[t * rows + r] = acc;
            }
            continue;
        }
        const int n = static_cast<int>(total);
        if constexpr (sizeof(S) == 4 && sizeof(I) == 4)
        {
#pragma unroll
            for (int j = 0; j < kStageLoads; ++j)
            {
                const int kk = lane + j * kWave;
                if (kk < n)
                {
                    __builtin_amdgcn_global_load_lds(v + base + kk, lv + j * kWave, 4, 0, 0);
                    __builtin_amdgcn_global_load_lds(ci + base + kk, lc + j * kWave, 4, 0, 0);
                }
            }
            __builtin_amdgcn_s_waitcnt(0);
        }
        else
        {
#pragma unroll
            for (int j = 0; j < kStageLoads; ++j)
            {
                const int kk = lane + j * kWave;
                if (kk < n)
                {
                    lv[kk] = v[base + kk];
                    lc[kk] = ci[base + kk];
                }
            }
        }
        wave_lds_sync();
        const int s = static_cast<int>(static_cast<int64_t>(s_i) - base);
        const int e = static_cast<int>(static_cast<int64_t>(e_i) - base);
        // an entry's tile operands: kSpmmTile contiguous values, 16-byte loads
        using Vec = S __attribute__((ext_vector_type(16 / sizeof(S))));
        constexpr int kVecs = kSpmmTile * sizeof(S) / 16;
        S acc[kSpmmTile];
#pragma unroll
        for (int t = 0; t < kSpmmTile; ++t) acc[t] = S(0);
        for (int q = s; q < e; ++q)
        {
            const S a = lv[q];
            const Vec* g = reinterpret_cast<const Vec*>(bt + static_cast<int64_t>(lc[q]) * kSpmmTile);
            Vec gv[kVecs];
#pragma unroll
            for (int u = 0; u < kVecs; ++u) gv[u] = g[u];
#pragma unroll
            for (int u = 0; u < kVecs; ++u)
#pragma unroll
                for (int w = 0; w < 16 / static_cast<int>(sizeof(S)); ++w)
                    acc[u * (16 / sizeof(S)) + w] = fma_s(a, gv[u][w], acc[u * (16 / sizeof(S)) + w]);
        }
        if (r < rows)
        {
#pragma unroll
            for (int t = 0; t < kSpmmTile; ++t)
                if (t < kt) y[t * rows + r] = acc[t];
        }
        wave_lds_sync();
    }
}

// y = A x (standalone; any grid size, no reductions).  Streams the matrix
// once from HBM: 4-byte values and indices are staged by global_load_lds,
// which measured 27.8 us against 31.1 us for VGPR staging on the 128^3
// 7-point matrix (the solver's cache-resident iterations measured the other
// way round and keep VGPR staging).
template <typename S, typename I>
__global__ __launch_bounds__(kBlock) void k_spmv(int64_t rows, const I* __restrict__ ro, const I* __restrict__ ci,
                                                 const S* __restrict__ v, const S* __restrict__ x,
                                                 S* __restrict__ y)
{
    __shared__ S lv_all[kWavesPerBlock][kStreamCap];
    __shared__ I lc_all[kWavesPerBlock][kStreamCap];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    PM_FOR_CHUNKS(c, chunk_count(rows), kWavesPerBlock)
    {
        const int64_t row0 = c * kChunkRows;
        int64_t s, e;
        const S acc = chunk_row_sums<S, I, SpmvSource<S>, true, true>(row0, rows, ro, ci, lv_all[wave],
                                                                      lc_all[wave], SpmvSource<S>{v, x}, s, e);
        if (row0 + lane < rows) y[row0 + lane] = acc;
    }
}

// Lanes-per-row SpMV (y = A x) kept for matrices whose rows are long on
// average, selectable through the autotuner (lanes > 0).
template <typename S, typename I, int L>
__global__ __launch_bounds__(kBlock) void k_spmv_lanes(int64_t rows, const I* __restrict__ ro,
                                                       const I* __restrict__ ci, const S* __restrict__ v,
                                                       const S* __restrict__ x, S* __restrict__ y)
{
    constexpr int64_t gpb = kBlock / L;
    const int group = threadIdx.x / L;
    const int lane = threadIdx.x % L;
    int64_t begin, end;
    block_range(rows, gpb, begin, end);
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
//           in iterative_solver_cpu.cpp:127-160 does); first bad row recorded;
//   ZERO_P: p0[row] = 0 (the first direction update then starts from z).
template <typename S, typename I, bool JACOBI, bool ZERO_P>
__global__ __launch_bounds__(kBlock) void k_residual(int64_t rows, const I* __restrict__ ro,
                                                     const I* __restrict__ ci, const S* __restrict__ v,
                                                     const S* __restrict__ b, const S* __restrict__ x,
                                                     S* __restrict__ r, S* __restrict__ inv, S* __restrict__ p0,
                                                     SolverState* st)
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
        const double acc =
            chunk_row_sums<double>(row0, rows, ro, ci, lv_all[wave], lc_all[wave], Fp64Source<S, S>{v, x}, s, e);
        const int64_t row = row0 + lane;
        if (row < rows)
        {
            r[row] = static_cast<S>(static_cast<double>(b[row]) - acc);
            if (ZERO_P) p0[row] = S(0);
            if (JACOBI)
            {
                double diag = 0.0, scale = 0.0;
                bool found = false;
                for (int64_t k = s; k < e; ++k)
                {
                    const double a = static_cast<double>(v[k]);
                    scale = fmax(scale, fabs(a));
                    if (static_cast<int64_t>(ci[k]) == row)
                    {
                        diag += a;
                        found = true;
                    }
                }
                const double threshold = static_cast<double>(scalar_epsilon<S>()) * scale * 16.0;
                const double reciprocal = 1.0 / diag;
                const bool ok = found && isfinite(diag) && diag > threshold && isfinite(reciprocal);
                inv[row] = ok ? static_cast<S>(reciprocal) : static_cast<S>(NAN);
                if (!ok) atomicMin(&st->bad_row, static_cast<unsigned long long>(row));
            }
        }
    }
}

// FP64 true residual partial sums: sum (b - A x)^2 and sum b^2, finished by
// the last workgroup (used once per call, outside the iteration).
template <typename S, typename V, typename I>
__global__ __launch_bounds__(kBlock) void k_relres(int64_t rows, const I* __restrict__ ro, const I* __restrict__ ci,
                                                   const S* __restrict__ v, const V* __restrict__ b,
                                                   const V* __restrict__ x, double* partials, unsigned* counter,
                                                   double* out)
{
    __shared__ S lv_all[kWavesPerBlock][kStreamCap];
    __shared__ I lc_all[kWavesPerBlock][kStreamCap];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    double rr = 0.0, bb = 0.0;
    PM_FOR_CHUNKS(c, chunk_count(rows), kWavesPerBlock)
    {
        const int64_t row0 = c * kChunkRows;
        int64_t s, e;
        const double acc =
            chunk_row_sums<double>(row0, rows, ro, ci, lv_all[wave], lc_all[wave], Fp64Source<S, V>{v, x}, s, e);
        const int64_t row = row0 + lane;
        if (row < rows)
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

// alpha from rho and p.q (alphaKernel, iterative_solver_kernels.cuh:169-202);
// the step is applied in S, so it must also be finite in S.
template <typename S>
__device__ __forceinline__ bool step_length(double rho, double den, double& alpha)
{
    alpha = 0.0;
    if (rho == 0.0 && den == 0.0) return true;
    if (isfinite(rho) && isfinite(den) && rho > 0.0 && den > 0.0)
    {
        alpha = rho / den;
        if (isfinite(alpha) && isfinite(static_cast<S>(alpha))) return true;
    }
    alpha = 0.0;
    return false;
}

// Dispatch helper: calls F<L>::run(...) for the runtime lanes-per-row value.
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

// Lanes-per-row for the lanes-per-row kernel from the mean row length.
inline int heuristic_lanes(int64_t rows, int64_t nnz)
{
    if (rows <= 0) return 1;
    const double mean = static_cast<double>(nnz) / static_cast<double>(rows);
    int lanes = 1;
    while (lanes < 64 && lanes < mean) lanes *= 2;
    return lanes;
}

} // namespace pm
