// Device COO -> CSR for gfx950.
//
//   1. pack     key = (row << cb) | col  (cb = bits of cols-1), payload = the
//               triplet's position; out-of-range coordinates and (optionally)
//               non-finite values are recorded as the lowest offending
//               position, like the reference's deferred CooStatus
//               (coo_to_csr.cu:116-161, 300-365).
//   2. sort     stable LSD radix sort over only the rb+cb key bits, 8 bits per
//               pass.  Within a workgroup tile the rank of each key among equal
//               digits is computed with 64-lane ballots (wave multisplit) and
//               a per-digit running offset over the waves in order, so equal
//               keys keep their input order: the same permutation as the
//               reference's std::stable_sort (sparse_ops_cpu.cpp:58-68).
//   3. runs     head flags + exclusive scan give each distinct (row, col) its
//               output slot.
//   4. fold     one thread per run sums its duplicates left to right starting
//               from the first value (sparse_ops_cpu.cpp:75-91), so the result
//               is bit-identical to the CPU conversion.
//   5. offsets  row_offsets[r] = first run whose row >= r (binary search), no
//               atomics, deterministic.
#include <algorithm>
#include <cmath>

#include "pm_common.hpp"
#include "sort.hpp"
#include "workspace.hpp"

namespace pm
{

constexpr int kScanTile = 2048;    // elements per workgroup in scans
constexpr int kSortItems = 8;      // keys per thread per radix tile
constexpr int kSortTile = kBlock * kSortItems;
constexpr unsigned long long kInvalidKey = ~0ull;

// ---- exclusive scan of u32 ---------------------------------------------------

__device__ __forceinline__ unsigned block_exclusive_scan(unsigned v, unsigned* lds, unsigned& total)
{
    // inclusive wave scan
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    unsigned x = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1)
    {
        const unsigned y = __shfl_up(x, off, kWave);
        if (lane >= off) x += y;
    }
    if (lane == kWave - 1) lds[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        unsigned run = 0;
        for (int w = 0; w < static_cast<int>(blockDim.x / kWave); ++w)
        {
            const unsigned t = lds[w];
            lds[w] = run;
            run += t;
        }
        lds[16] = run;
    }
    __syncthreads();
    total = lds[16];
    const unsigned result = x - v + lds[wave];
    __syncthreads();
    return result;
}

__global__ __launch_bounds__(kBlock) void k_tile_sums(const unsigned* in, int64_t n, unsigned* sums)
{
    __shared__ unsigned lds[32];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + threadIdx.x * (kScanTile / kBlock);
    unsigned s = 0;
    for (int i = 0; i < kScanTile / kBlock; ++i)
        if (base + i < n) s += in[base + i];
    unsigned total;
    block_exclusive_scan(s, lds, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Exclusive scan of each tile, offset by the scanned tile sums (null = 0).
__global__ __launch_bounds__(kBlock) void k_tile_scan(const unsigned* in, unsigned* out, int64_t n,
                                                      const unsigned* tile_offsets, unsigned* grand_total)
{
    __shared__ unsigned lds[32];
    constexpr int per = kScanTile / kBlock;
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + threadIdx.x * per;
    unsigned vals[per];
    unsigned s = 0;
#pragma unroll
    for (int i = 0; i < per; ++i)
    {
        vals[i] = base + i < n ? in[base + i] : 0u;
        s += vals[i];
    }
    unsigned total;
    unsigned run = block_exclusive_scan(s, lds, total) + (tile_offsets ? tile_offsets[blockIdx.x] : 0u);
#pragma unroll
    for (int i = 0; i < per; ++i)
    {
        if (base + i < n) out[base + i] = run;
        run += vals[i];
    }
    if (grand_total && blockIdx.x == gridDim.x - 1 && threadIdx.x == kBlock - 1) *grand_total = run;
}

// Scratch needed by scan_u32 for n elements (u32 words).
int64_t scan_scratch_words(int64_t n)
{
    int64_t words = 0;
    while (n > kScanTile)
    {
        n = ceil_div(n, kScanTile);
        words += 2 * n + 64;
    }
    return words + 64;
}

// out[i] = sum(in[0..i)); *grand_total (device) = sum of all when non-null.
int scan_u32(const unsigned* in, unsigned* out, int64_t n, unsigned* scratch, unsigned* grand_total, hipStream_t stream)
{
    if (n <= 0) return PM_OK;
    const int64_t tiles = ceil_div(n, kScanTile);
    if (tiles == 1)
    {
        hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(kBlock), 0, stream, in, out, n, nullptr, grand_total);
        PM_HIP(hipGetLastError());
        return PM_OK;
    }
    unsigned* sums = scratch;
    unsigned* scanned = scratch + tiles + 32;
    unsigned* rest = scanned + tiles + 32;
    hipLaunchKernelGGL(k_tile_sums, dim3(tiles), dim3(kBlock), 0, stream, in, n, sums);
    PM_HIP(hipGetLastError());
    PM_TRY(scan_u32(sums, scanned, tiles, rest, nullptr, stream));
    hipLaunchKernelGGL(k_tile_scan, dim3(tiles), dim3(kBlock), 0, stream, in, out, n, scanned, grand_total);
    PM_HIP(hipGetLastError());
    return PM_OK;
}

// ---- conversion kernels --------------------------------------------------------

__global__ void k_coo_status_init(CooStatus* st, long long expected)
{
    st->invalid_source = ~0ull;
    st->invalid_value = ~0ull;
    st->nnz = 0;
    st->expected_nnz = expected;
}

__global__ void k_coo_status_init_expected(CooStatus* st, long long expected) { st->expected_nnz = expected; }

template <typename S>
__global__ __launch_bounds__(kBlock) void k_coo_pack(int64_t count, int64_t rows, int64_t cols, int cb,
                                                     const int64_t* __restrict__ r, const int64_t* __restrict__ c,
                                                     const S* __restrict__ v, int check_finite,
                                                     unsigned long long* __restrict__ keys,
                                                     unsigned* __restrict__ perm, CooStatus* st)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
    {
        perm[i] = static_cast<unsigned>(i);
        const int64_t row = r[i], col = c[i];
        if (check_finite && !isfinite(static_cast<double>(v[i])))
            atomicMin(&st->invalid_value, static_cast<unsigned long long>(i));
        if (row < 0 || row >= rows || col < 0 || col >= cols)
        {
            keys[i] = kInvalidKey;
            atomicMin(&st->invalid_source, static_cast<unsigned long long>(i));
        }
        else
        {
            keys[i] = (static_cast<unsigned long long>(row) << cb) | static_cast<unsigned long long>(col);
        }
    }
}

// Per-tile digit histogram, stored digit-major: hist[d * tiles + tile].
__global__ __launch_bounds__(kBlock) void k_radix_hist(const unsigned long long* __restrict__ keys, int64_t count,
                                                       int shift, unsigned* __restrict__ hist, int64_t tiles)
{
    __shared__ unsigned h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kSortTile;
    for (int i = 0; i < kSortItems; ++i)
    {
        const int64_t idx = base + i * kBlock + threadIdx.x;
        if (idx < count) atomicAdd(&h[(keys[idx] >> shift) & 0xFF], 1u);
    }
    __syncthreads();
    hist[static_cast<int64_t>(threadIdx.x) * tiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one radix pass.  Tile elements are visited in input order
// (round, wave, lane); ranks among equal digits come from ballots inside the
// wave and a running per-digit offset across waves and rounds.
__global__ __launch_bounds__(kBlock) void k_radix_scatter(const unsigned long long* __restrict__ keys_in,
                                                          const unsigned* __restrict__ vals_in, int64_t count,
                                                          int shift, const unsigned* __restrict__ offsets,
                                                          int64_t tiles, unsigned long long* __restrict__ keys_out,
                                                          unsigned* __restrict__ vals_out)
{
    __shared__ unsigned base_offset[256];
    __shared__ unsigned running[256];
    __shared__ unsigned wave_hist[kBlock / kWave][256];
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    base_offset[threadIdx.x] = offsets[static_cast<int64_t>(threadIdx.x) * tiles + blockIdx.x];
    running[threadIdx.x] = 0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; ++w) wave_hist[w][threadIdx.x] = 0;
    __syncthreads();
    const unsigned long long lanes_below = (1ull << lane) - 1ull;
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kSortTile;
    for (int round = 0; round < kSortItems; ++round)
    {
        const int64_t idx = base + round * kBlock + threadIdx.x;
        const bool active = idx < count;
        unsigned long long key = 0;
        unsigned val = 0;
        unsigned digit = 0;
        if (active)
        {
            key = keys_in[idx];
            val = vals_in[idx];
            digit = static_cast<unsigned>((key >> shift) & 0xFF);
        }
        unsigned long long peers = __ballot(active);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit)
        {
            const unsigned long long ones = __ballot(active && ((digit >> bit) & 1u));
            peers &= ((digit >> bit) & 1u) ? ones : ~ones;
        }
        const unsigned rank = static_cast<unsigned>(__popcll(peers & lanes_below));
        if (active && rank == 0) wave_hist[wave][digit] = static_cast<unsigned>(__popcll(peers));
        __syncthreads();
        {
            // digit-wise exclusive offsets over the waves, in wave order
            const int d = threadIdx.x;
            unsigned run = running[d];
#pragma unroll
            for (int w = 0; w < kBlock / kWave; ++w)
            {
                const unsigned t = wave_hist[w][d];
                wave_hist[w][d] = run;
                run += t;
            }
            running[d] = run;
        }
        __syncthreads();
        if (active)
        {
            const unsigned pos = base_offset[digit] + wave_hist[wave][digit] + rank;
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
#pragma unroll
        for (int w = 0; w < kBlock / kWave; ++w) wave_hist[w][threadIdx.x] = 0;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void k_mark_heads(const unsigned long long* __restrict__ keys, int64_t count,
                                                       unsigned* __restrict__ heads)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
    {
        const unsigned long long k = keys[i];
        heads[i] = (k != kInvalidKey && (i == 0 || k != keys[i - 1])) ? 1u : 0u;
    }
}

__global__ void k_store_nnz(const unsigned* total, CooStatus* st) { st->nnz = static_cast<long long>(*total); }

template <typename S>
__global__ __launch_bounds__(kBlock) void k_fold_runs(int64_t count, const unsigned long long* __restrict__ keys,
                                                      const unsigned* __restrict__ perm,
                                                      const unsigned* __restrict__ heads,
                                                      const unsigned* __restrict__ run_id, const S* __restrict__ v,
                                                      int cb, int check_finite, int64_t capacity,
                                                      int64_t* __restrict__ col_out, S* __restrict__ val_out,
                                                      unsigned long long* __restrict__ run_keys, CooStatus* st)
{
    const unsigned long long col_mask = cb >= 64 ? ~0ull : ((1ull << cb) - 1ull);
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
    {
        if (!heads[i]) continue;
        const unsigned long long key = keys[i];
        S sum = v[perm[i]];
        if (check_finite && !isfinite(static_cast<double>(sum)))
            atomicMin(&st->invalid_value, static_cast<unsigned long long>(perm[i]));
        for (int64_t j = i + 1; j < count && keys[j] == key; ++j)
        {
            sum += v[perm[j]];
            if (check_finite && !isfinite(static_cast<double>(sum)))
                atomicMin(&st->invalid_value, static_cast<unsigned long long>(perm[j]));
        }
        const int64_t run = run_id[i];
        if (run < capacity)
        {
            col_out[run] = static_cast<int64_t>(key & col_mask);
            val_out[run] = sum;
            run_keys[run] = key;
        }
    }
}

// row_offsets[r] = number of runs with row < r  (r in [0, rows]).
__global__ __launch_bounds__(kBlock) void k_row_offsets(int64_t rows, const unsigned long long* __restrict__ run_keys,
                                                        const CooStatus* st, int64_t capacity, int cb,
                                                        int64_t* __restrict__ row_offsets)
{
    int64_t nnz = st->nnz;
    if (nnz > capacity) nnz = capacity;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r <= rows; r += stride)
    {
        int64_t lo = 0, hi = nnz;
        while (lo < hi)
        {
            const int64_t mid = (lo + hi) / 2;
            if (static_cast<int64_t>(run_keys[mid] >> cb) < r) lo = mid + 1;
            else hi = mid;
        }
        row_offsets[r] = lo;
    }
}

int radix_sort_pairs(unsigned long long* keys_a, unsigned* vals_a, unsigned long long* keys_b, unsigned* vals_b,
                     int64_t count, int key_bits, unsigned* hist, unsigned* scratch, hipStream_t stream, bool* in_b)
{
    *in_b = false;
    if (count <= 0) return PM_OK;
    const int passes = (key_bits + 7) / 8;
    const int64_t tiles = ceil_div(count, kSortTile);
    unsigned long long* kin = keys_a;
    unsigned long long* kout = keys_b;
    unsigned* vin = vals_a;
    unsigned* vout = vals_b;
    for (int pass = 0; pass < passes; ++pass)
    {
        const int shift = pass * 8;
        hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(kBlock), 0, stream, kin, count, shift, hist, tiles);
        PM_HIP(hipGetLastError());
        PM_TRY(scan_u32(hist, hist, 256 * tiles, scratch, nullptr, stream));
        hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(kBlock), 0, stream, kin, vin, count, shift, hist,
                           tiles, kout, vout);
        PM_HIP(hipGetLastError());
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    *in_b = (passes % 2) == 1;
    return PM_OK;
}

int64_t radix_hist_words(int64_t count) { return 256 * ceil_div(count > 0 ? count : 1, kSortTile); }

int bits_for(int64_t n)
{
    int b = 0;
    while (b < 63 && (int64_t(1) << b) < n) ++b;
    return b;
}

struct CooBuffers
{
    unsigned long long* keys_a;
    unsigned long long* keys_b;
    unsigned* vals_a;
    unsigned* vals_b;
    unsigned* hist;
    unsigned* heads;
    unsigned* run_id;
    unsigned* scratch;
    unsigned* total;
    CooStatus* status;
    unsigned long long* sorted_keys;
    unsigned* sorted_perm;
};

CooBuffers coo_layout(pm_workspace* ws, int64_t count, size_t* bytes_out)
{
    auto align = [](size_t b) { return (b + 255) & ~size_t(255); };
    const int64_t tiles = ceil_div(count > 0 ? count : 1, kSortTile);
    const int64_t hist_n = 256 * tiles;
    const int64_t big = std::max<int64_t>(hist_n, count > 0 ? count : 1);
    size_t off = 0;
    size_t o_keys_a = off; off += align(sizeof(unsigned long long) * (count + 1));
    size_t o_keys_b = off; off += align(sizeof(unsigned long long) * (count + 1));
    size_t o_vals_a = off; off += align(sizeof(unsigned) * (count + 1));
    size_t o_vals_b = off; off += align(sizeof(unsigned) * (count + 1));
    size_t o_hist = off; off += align(sizeof(unsigned) * hist_n);
    size_t o_heads = off; off += align(sizeof(unsigned) * (count + 1));
    size_t o_run = off; off += align(sizeof(unsigned) * (count + 1));
    size_t o_scr = off; off += align(sizeof(unsigned) * scan_scratch_words(big));
    size_t o_total = off; off += 256;
    size_t o_status = off; off += 256;
    *bytes_out = off;
    CooBuffers b{};
    if (!ws) return b;
    char* base = ws->sort.as<char>();
    b.keys_a = reinterpret_cast<unsigned long long*>(base + o_keys_a);
    b.keys_b = reinterpret_cast<unsigned long long*>(base + o_keys_b);
    b.vals_a = reinterpret_cast<unsigned*>(base + o_vals_a);
    b.vals_b = reinterpret_cast<unsigned*>(base + o_vals_b);
    b.hist = reinterpret_cast<unsigned*>(base + o_hist);
    b.heads = reinterpret_cast<unsigned*>(base + o_heads);
    b.run_id = reinterpret_cast<unsigned*>(base + o_run);
    b.scratch = reinterpret_cast<unsigned*>(base + o_scr);
    b.total = reinterpret_cast<unsigned*>(base + o_total);
    b.status = reinterpret_cast<CooStatus*>(base + o_status);
    return b;
}

int prepare_coo(int64_t rows, int64_t cols, int64_t count, const int64_t* r, const int64_t* c, const void* v,
                int32_t dtype, int32_t flags, pm_workspace* ws, hipStream_t stream)
{
    if (!ws) return fail(PM_INVALID_ARGUMENT, "cooToCsr: null workspace");
    if (rows < 0 || cols < 0) return fail(PM_INVALID_ARGUMENT, "cooToCsr dimensions must be non-negative");
    if (count < 0) return fail(PM_INVALID_ARGUMENT, "cooToCsr row, column, and value arrays must have equal lengths");
    if (dtype != PM_F32 && dtype != PM_F64) return fail(PM_INVALID_ARGUMENT, "cooToCsr: unknown dtype %d", dtype);
    if (count > 0 && (!r || !c || !v)) return fail(PM_INVALID_ARGUMENT, "cooToCsr: null triplet arrays");
    if (count >= (int64_t(1) << 31)) return fail(PM_OVERFLOW, "cooToCsr triplet count exceeds 2^31 - 1");
    const int rb = bits_for(rows), cb = bits_for(cols);
    if (rb + cb > 63) return fail(PM_OVERFLOW, "cooToCsr: rows x cols needs more than 63 key bits");
    PM_TRY(ws->bind(stream));
    PM_TRY(ws->ensure_events());
    if (ws->coo_pending_status)
        return fail(PM_LOGIC_ERROR, "workspace status must be checked before another COO conversion");
    size_t bytes = 0;
    coo_layout(nullptr, count, &bytes);
    PM_TRY(ws->sort.reserve(bytes));
    CooBuffers b = coo_layout(ws, count, &bytes);
    hipLaunchKernelGGL(k_coo_status_init, dim3(1), dim3(1), 0, stream, b.status, -1ll);
    ws->coo_rows = rows;
    ws->coo_cols = cols;
    ws->coo_count = count;
    ws->coo_dtype = dtype;
    ws->coo_flags = flags;
    ws->coo_key_bits = cb;
    ws->coo_values = v;
    if (count > 0)
    {
        const int grid = static_cast<int>(std::min<int64_t>(ceil_div(count, kBlock), 8192));
        if (dtype == PM_F32)
            hipLaunchKernelGGL(k_coo_pack<float>, dim3(grid), dim3(kBlock), 0, stream, count, rows, cols, cb, r, c,
                               static_cast<const float*>(v), flags & PM_COO_CHECK_FINITE, b.keys_a, b.vals_a,
                               b.status);
        else
            hipLaunchKernelGGL(k_coo_pack<double>, dim3(grid), dim3(kBlock), 0, stream, count, rows, cols, cb, r, c,
                               static_cast<const double*>(v), flags & PM_COO_CHECK_FINITE, b.keys_a, b.vals_a,
                               b.status);
        PM_HIP(hipGetLastError());
        bool in_b = false;
        PM_TRY(radix_sort_pairs(b.keys_a, b.vals_a, b.keys_b, b.vals_b, count, rb + cb, b.hist, b.scratch, stream,
                                &in_b));
        unsigned long long* kin = in_b ? b.keys_b : b.keys_a;
        // kin/vin now hold the sorted keys and permutation
        hipLaunchKernelGGL(k_mark_heads, dim3(grid), dim3(kBlock), 0, stream, kin, count, b.heads);
        PM_HIP(hipGetLastError());
        PM_TRY(scan_u32(b.heads, b.run_id, count, b.scratch, b.total, stream));
        hipLaunchKernelGGL(k_store_nnz, dim3(1), dim3(1), 0, stream, b.total, b.status);
        PM_HIP(hipGetLastError());
        // remember where the sorted data lives
        ws->coo_prepared = true;
        ws->coo_key_bits = cb | (kin == b.keys_a ? 0 : 0x100);
    }
    else
    {
        ws->coo_prepared = true;
        ws->coo_key_bits = cb;
    }
    return PM_OK;
}

int coo_status_error(const CooStatus& s, const char* name)
{
    if (s.invalid_source != ~0ull)
        return fail(PM_OUT_OF_RANGE, "%s: coordinate out of range at source offset %llu", name, s.invalid_source);
    if (s.invalid_value != ~0ull)
        return fail(PM_INVALID_ARGUMENT, "%s: non-finite value at source offset %llu", name, s.invalid_value);
    if (s.expected_nnz >= 0 && s.nnz != s.expected_nnz)
        return fail(PM_RUNTIME_ERROR, "%s: output nnz is %lld, but duplicate-combined nnz is %lld", name,
                    s.expected_nnz, s.nnz);
    return PM_OK;
}

int finish_coo(int64_t* row_offsets, int64_t* col_out, void* val_out, int64_t expected, pm_workspace* ws,
               hipStream_t stream)
{
    if (!ws || !ws->coo_prepared) return fail(PM_LOGIC_ERROR, "cooToCsr: no prepared conversion in this workspace");
    PM_TRY(ws->bind(stream));
    if (expected < 0) return fail(PM_INVALID_ARGUMENT, "cooToCsr: negative output nnz");
    if (!row_offsets || (expected > 0 && (!col_out || !val_out)))
        return fail(PM_INVALID_ARGUMENT, "cooToCsr: null output arrays");
    const int64_t count = ws->coo_count;
    size_t bytes = 0;
    CooBuffers b = coo_layout(ws, count, &bytes);
    const int cb = ws->coo_key_bits & 0xFF;
    const bool sorted_in_b = (ws->coo_key_bits & 0x100) != 0;
    const unsigned long long* keys = sorted_in_b ? b.keys_b : b.keys_a;
    const unsigned* perm = sorted_in_b ? b.vals_b : b.vals_a;
    unsigned long long* run_keys = sorted_in_b ? b.keys_a : b.keys_b;  // the free key buffer
    // record the caller's nnz for the deferred mismatch check
    hipLaunchKernelGGL(k_coo_status_init_expected, dim3(1), dim3(1), 0, stream, b.status, static_cast<long long>(expected));
    if (count > 0)
    {
        const int grid = static_cast<int>(std::min<int64_t>(ceil_div(count, kBlock), 8192));
        const int check = ws->coo_flags & PM_COO_CHECK_FINITE;
        if (ws->coo_dtype == PM_F32)
            hipLaunchKernelGGL(k_fold_runs<float>, dim3(grid), dim3(kBlock), 0, stream, count, keys, perm, b.heads,
                               b.run_id, static_cast<const float*>(ws->coo_values), cb, check, expected, col_out,
                               static_cast<float*>(val_out), run_keys, b.status);
        else
            hipLaunchKernelGGL(k_fold_runs<double>, dim3(grid), dim3(kBlock), 0, stream, count, keys, perm, b.heads,
                               b.run_id, static_cast<const double*>(ws->coo_values), cb, check, expected, col_out,
                               static_cast<double*>(val_out), run_keys, b.status);
        PM_HIP(hipGetLastError());
    }
    const int rgrid = static_cast<int>(std::min<int64_t>(ceil_div(ws->coo_rows + 1, kBlock), 8192));
    hipLaunchKernelGGL(k_row_offsets, dim3(rgrid), dim3(kBlock), 0, stream, ws->coo_rows, run_keys, b.status, expected,
                       cb, row_offsets);
    PM_HIP(hipGetLastError());
    PM_HIP(hipEventRecord(ws->done_event, stream));
    ws->coo_pending_status = true;
    ws->coo_prepared = false;
    return PM_OK;
}

} // namespace pm

using namespace pm;

extern "C" {

int pm_coo_to_csr_prepare(int64_t rows, int64_t cols, int64_t count, const int64_t* r, const int64_t* c,
                          const void* v, int32_t dtype, int32_t flags, pm_workspace* ws, hipStream_t stream)
{
    return prepare_coo(rows, cols, count, r, c, v, dtype, flags, ws, stream);
}

int pm_coo_to_csr_nnz(pm_workspace* ws, int64_t* nnz)
{
    if (!ws || !ws->coo_prepared) return fail(PM_LOGIC_ERROR, "cooToCsr: no prepared conversion in this workspace");
    if (!nnz) return fail(PM_INVALID_ARGUMENT, "cooToCsr: null nnz output");
    size_t bytes = 0;
    CooBuffers b = coo_layout(ws, ws->coo_count, &bytes);
    PM_HIP(hipMemcpyAsync(ws->host_coo, b.status, sizeof(CooStatus), hipMemcpyDeviceToHost, ws->stream));
    PM_HIP(hipStreamSynchronize(ws->stream));
    CooStatus s = *ws->host_coo;
    s.expected_nnz = -1;
    const int status = coo_status_error(s, "cooToCsr");
    if (status != PM_OK)
    {
        ws->coo_prepared = false;
        return status;
    }
    *nnz = s.nnz;
    return PM_OK;
}

int pm_coo_to_csr_finish(int64_t* row_offsets, int64_t* col_indices, void* values, int64_t expected_nnz,
                         pm_workspace* ws, hipStream_t stream)
{
    return finish_coo(row_offsets, col_indices, values, expected_nnz, ws, stream);
}

int pm_coo_to_csr_async(int64_t rows, int64_t cols, int64_t count, const int64_t* r, const int64_t* c, const void* v,
                        int32_t dtype, int32_t flags, int64_t* row_offsets, int64_t* out_cols, void* out_vals,
                        int64_t expected_nnz, pm_workspace* ws, hipStream_t stream)
{
    PM_TRY(prepare_coo(rows, cols, count, r, c, v, dtype, flags, ws, stream));
    return finish_coo(row_offsets, out_cols, out_vals, expected_nnz, ws, stream);
}

int pm_workspace_check_status(pm_workspace* ws)
{
    if (!ws) return fail(PM_INVALID_ARGUMENT, "checkStatus: null workspace");
    if (!ws->coo_pending_status) return PM_OK;
    const hipError_t q = hipEventQuery(ws->done_event);
    if (q == hipErrorNotReady)
        return fail(PM_LOGIC_ERROR, "SparseOpsWorkspace::checkStatus requires stream synchronization first");
    if (q != hipSuccess) return check_hip(q, "hipEventQuery", __FILE__, __LINE__);
    size_t bytes = 0;
    CooBuffers b = coo_layout(ws, ws->coo_count, &bytes);
    PM_HIP(hipMemcpy(ws->host_coo, b.status, sizeof(CooStatus), hipMemcpyDeviceToHost));
    ws->coo_pending_status = false;
    return coo_status_error(*ws->host_coo, "cooToCsrAsync");
}

} // extern "C"
