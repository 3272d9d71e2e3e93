// Detection and construction of the diagonal layout (see dia.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <type_traits>

#include "dia.hpp"

namespace pm
{

namespace
{

constexpr int kEmpty = INT_MIN;

// table[0..kMaxDia): distinct offsets (kEmpty where unused); table[kMaxDia]:
// flags (1: a row's columns are not strictly increasing, 2: more than kMaxDia
// diagonals, 4: with UNIT, a row whose Jacobi inverse is not exactly 1).  Each
// workgroup collects its rows' offsets in LDS first, then merges them into the
// global table with at most kMaxDia compare-and-swaps.
__global__ void k_dia_reset(int* table)
{
    const int t = threadIdx.x;
    if (t < kMaxDia) table[t] = kEmpty;
    if (t == kMaxDia) table[t] = 0;
}

__device__ __forceinline__ bool dia_insert(int* slots, int off, bool shared_mem)
{
    for (int s = 0; s < kMaxDia; ++s)
    {
        int cur = shared_mem ? __atomic_load_n(&slots[s], __ATOMIC_RELAXED)
                             : __hip_atomic_load(&slots[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == off) return true;
        if (cur == kEmpty)
        {
            cur = atomicCAS(&slots[s], kEmpty, off);
            if (cur == kEmpty || cur == off) return true;
        }
    }
    return false;
}

// The Jacobi inverse of a row as the solver's initialisation forms it (sum of
// the diagonal entries, rejected when missing, non-finite or not above
// 16 eps(S) max|a_row|, reciprocal in FP64 rounded to S): UNIT asks whether it
// is exactly 1 in every row, i.e. whether z = D^-1 r is r itself.
template <typename S, bool UNIT, typename I>
__global__ __launch_bounds__(kBlock) void k_dia_detect(int64_t rows, const I* __restrict__ ro,
                                                      const I* __restrict__ ci, const S* __restrict__ v,
                                                      int* table)
{
    __shared__ int local[kMaxDia];
    __shared__ int flags;
    if (threadIdx.x < kMaxDia) local[threadIdx.x] = kEmpty;
    if (threadIdx.x == 0) flags = 0;
    __syncthreads();
    int my_flags = 0;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < rows; r += stride)
    {
        int64_t prev = -1;
        const int64_t end = static_cast<int64_t>(ro[r + 1]);
        double diag = 0.0, scale = 0.0;
        bool found = false;
        for (int64_t k = static_cast<int64_t>(ro[r]); k < end; ++k)
        {
            const int64_t c = static_cast<int64_t>(ci[k]);
            if (c <= prev) my_flags |= 1;
            prev = c;
            if (!dia_insert(local, static_cast<int>(c - r), true)) my_flags |= 2;
            if constexpr (UNIT)
            {
                const double a = static_cast<double>(v[k]);
                scale = fmax(scale, fabs(a));
                if (c == r)
                {
                    diag += a;
                    found = true;
                }
            }
        }
        if constexpr (UNIT)
        {
            const double threshold = static_cast<double>(scalar_epsilon<S>()) * scale * 16.0;
            const double reciprocal = 1.0 / diag;
            const bool ok = found && isfinite(diag) && diag > threshold && isfinite(reciprocal);
            if (!ok || static_cast<S>(reciprocal) != S(1)) my_flags |= 4;
        }
    }
    if (my_flags) atomicOr(&flags, my_flags);
    __syncthreads();
    if (threadIdx.x < kMaxDia)
    {
        const int off = local[threadIdx.x];
        if (off != kEmpty && !dia_insert(table, off, false)) atomicOr(&table[kMaxDia], 2);
    }
    if (threadIdx.x == 0 && flags) atomicOr(&table[kMaxDia], flags);
}

template <typename S, typename I>
__global__ __launch_bounds__(kBlock) void k_dia_fill(int64_t rows, const I* __restrict__ ro,
                                                    const I* __restrict__ ci, const S* __restrict__ v,
                                                    DiaArgs dia, S* __restrict__ dv, uint16_t* __restrict__ mask)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < rows; r += stride)
    {
        unsigned m = 0;
        int d = 0;
        const int64_t end = static_cast<int64_t>(ro[r + 1]);
        for (int64_t k = static_cast<int64_t>(ro[r]); k < end; ++k)
        {
            const int off = static_cast<int>(static_cast<int64_t>(ci[k]) - r);
            while (d < dia.ndiag && dia.off[d] < off) ++d;  // columns increase, so d only moves forward
            if (d >= dia.ndiag || dia.off[d] != off) continue;  // cannot happen after detection
            dv[static_cast<int64_t>(d) * rows + r] = v[k];
            m |= 1u << d;
        }
        mask[r] = static_cast<uint16_t>(m);
    }
}

// Mirror check of a filled layout: flag 8 when some lower entry (row,
// row + off[d]) is present without its mirror (row + off[d], row) on diagonal
// mirror.of[d], or the other way round, or their values differ in any bit;
// flag 16 when some row's main-diagonal entry is absent or not exactly 1.
struct DiaMirror
{
    int32_t of[kMaxDia];  // index of diagonal -off[d], for every lower diagonal d
};

template <typename S>
__device__ __forceinline__ bool same_bits(S a, S b)
{
    if constexpr (sizeof(S) == 4) return __float_as_uint(a) == __float_as_uint(b);
    else return __double_as_longlong(a) == __double_as_longlong(b);
}

template <typename S>
__global__ __launch_bounds__(kBlock) void k_dia_mirror(int64_t rows, DiaArgs dia, DiaMirror mirror,
                                                      const S* __restrict__ dv, const uint16_t* __restrict__ mask,
                                                      int main_d, int* table)
{
    __shared__ int flags;
    if (threadIdx.x == 0) flags = 0;
    __syncthreads();
    int my_flags = 0;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < rows; r += stride)
    {
        const unsigned m = mask[r];
        for (int d = 0; d < dia.ndiag && dia.off[d] < 0; ++d)
        {
            const bool here = (m >> d) & 1u;
            const int64_t src = r + dia.off[d];
            if (src < 0)
            {
                if (here) my_flags |= 8;
                continue;
            }
            const int e = mirror.of[d];
            const bool there = (mask[src] >> e) & 1u;
            if (here != there) my_flags |= 8;
            else if (here)
            {
                const S a = dv[static_cast<int64_t>(d) * rows + r];
                const S b = dv[static_cast<int64_t>(e) * rows + src];
                if (!same_bits(a, b)) my_flags |= 8;
            }
        }
        if (main_d < 0 || !((m >> main_d) & 1u) || dv[static_cast<int64_t>(main_d) * rows + r] != S(1))
            my_flags |= 16;
    }
    if (my_flags) atomicOr(&flags, my_flags);
    __syncthreads();
    if (threadIdx.x == 0 && flags) atomicOr(&table[kMaxDia], flags);
}

} // namespace

unsigned long long dia_signature(const DiaLayout& d)
{
    unsigned long long sig = 1469598103934665603ull;  // FNV-1a
    auto mix = [&](unsigned long long v) { sig = (sig ^ v) * 1099511628211ull; };
    mix(d.active ? 1 : 0);
    if (!d.active) return sig;
    for (int i = 0; i < d.args.ndiag; ++i)
    {
        mix(static_cast<unsigned>(d.args.off[i]));
        mix(static_cast<unsigned>(d.args.slot[i]));
        mix(static_cast<unsigned>(d.args.shift[i]));
    }
    mix(reinterpret_cast<unsigned long long>(d.values));
    mix(reinterpret_cast<unsigned long long>(d.mask));
    return sig;
}

int dia_build(DeviceBuffer& storage, const pm_csr* a, DiaLayout* out, hipStream_t stream, bool* unit_diagonal,
              bool mirror)
{
    *out = DiaLayout{};
    if (unit_diagonal) *unit_diagonal = false;
    const int64_t n = a->rows;
    const bool compact = a->row_offsets32 && a->col_indices32;
    if (n <= 0 || a->nnz <= 0 || a->rows != a->cols || n >= (int64_t(1) << 31)) return PM_OK;
    if (!compact && (!a->row_offsets || !a->col_indices)) return PM_OK;
    const size_t es = a->dtype == PM_F64 ? 8 : 4;
    const size_t table_bytes = 256;
    PM_TRY(storage.reserve(table_bytes));
    int* table = storage.as<int>();
    hipLaunchKernelGGL(k_dia_reset, dim3(1), dim3(64), 0, stream, table);
    const int grid = static_cast<int>(std::min<int64_t>(ceil_div(n, kBlock), 2048));
    // the layout reads A through its 32-bit mirrors when it has them, else
    // through the 64-bit arrays
    auto detect_with = [&](auto ro, auto ci) {
        using I = std::remove_cv_t<std::remove_pointer_t<decltype(ro)>>;
        if (a->dtype == PM_F64)
        {
            const auto detect = unit_diagonal ? k_dia_detect<double, true, I> : k_dia_detect<double, false, I>;
            hipLaunchKernelGGL(detect, dim3(grid), dim3(kBlock), 0, stream, n, ro, ci,
                               static_cast<const double*>(a->values), table);
        }
        else
        {
            const auto detect = unit_diagonal ? k_dia_detect<float, true, I> : k_dia_detect<float, false, I>;
            hipLaunchKernelGGL(detect, dim3(grid), dim3(kBlock), 0, stream, n, ro, ci,
                               static_cast<const float*>(a->values), table);
        }
    };
    if (compact) detect_with(static_cast<const int32_t*>(a->row_offsets32), static_cast<const int32_t*>(a->col_indices32));
    else detect_with(static_cast<const int64_t*>(a->row_offsets), static_cast<const int64_t*>(a->col_indices));
    PM_HIP(hipGetLastError());
    int host[kMaxDia + 1];
    PM_HIP(hipMemcpyAsync(host, table, sizeof(host), hipMemcpyDeviceToHost, stream));
    PM_HIP(hipStreamSynchronize(stream));
    if (unit_diagonal) *unit_diagonal = (host[kMaxDia] & 4) == 0;
    if ((host[kMaxDia] & 3) != 0) return PM_OK;
    DiaArgs args{};
    for (int s = 0; s < kMaxDia; ++s)
        if (host[s] != kEmpty) args.off[args.ndiag++] = host[s];
    std::sort(args.off, args.off + args.ndiag);
    for (int d = 0; d < args.ndiag; ++d)
    {
        args.slot[d] = d;
        args.shift[d] = 0;
    }
    if (args.ndiag == 0 || static_cast<double>(args.ndiag) * static_cast<double>(n) >
                               kDiaMaxFill * static_cast<double>(a->nnz))
        return PM_OK;
    const size_t values_bytes = (es * static_cast<size_t>(args.ndiag) * static_cast<size_t>(n) + 255) & ~size_t(255);
    const size_t mask_bytes = (sizeof(uint16_t) * static_cast<size_t>(n) + 255) & ~size_t(255);
    PM_TRY(storage.reserve(table_bytes + values_bytes + mask_bytes));
    char* base = storage.as<char>();
    void* dv = base + table_bytes;
    uint16_t* mask = reinterpret_cast<uint16_t*>(base + table_bytes + values_bytes);
    auto fill_with = [&](auto ro, auto ci) {
        using I = std::remove_cv_t<std::remove_pointer_t<decltype(ro)>>;
        if (a->dtype == PM_F64)
            hipLaunchKernelGGL((k_dia_fill<double, I>), dim3(grid), dim3(kBlock), 0, stream, n, ro, ci,
                               static_cast<const double*>(a->values), args, static_cast<double*>(dv), mask);
        else
            hipLaunchKernelGGL((k_dia_fill<float, I>), dim3(grid), dim3(kBlock), 0, stream, n, ro, ci,
                               static_cast<const float*>(a->values), args, static_cast<float*>(dv), mask);
    };
    if (compact) fill_with(static_cast<const int32_t*>(a->row_offsets32), static_cast<const int32_t*>(a->col_indices32));
    else fill_with(static_cast<const int64_t*>(a->row_offsets), static_cast<const int64_t*>(a->col_indices));
    PM_HIP(hipGetLastError());
    out->streams = args.ndiag;
    if (mirror)
    {
        // the mirror of every lower diagonal must be a diagonal of the layout
        DiaMirror mirrors{};
        int main_d = -1, lower = 0;
        bool paired = true;
        for (int d = 0; d < args.ndiag; ++d)
        {
            if (args.off[d] == 0) main_d = d;
            if (args.off[d] >= 0) continue;
            ++lower;
            int e = -1;
            for (int k = 0; k < args.ndiag; ++k)
                if (args.off[k] == -args.off[d]) e = k;
            if (e < 0) paired = false;
            mirrors.of[d] = e;
        }
        if (paired)
        {
            table = reinterpret_cast<int*>(base);  // the storage may have moved
            hipLaunchKernelGGL(k_dia_reset, dim3(1), dim3(64), 0, stream, table);
            if (a->dtype == PM_F64)
                hipLaunchKernelGGL(k_dia_mirror<double>, dim3(grid), dim3(kBlock), 0, stream, n, args, mirrors,
                                   static_cast<const double*>(dv), mask, main_d, table);
            else
                hipLaunchKernelGGL(k_dia_mirror<float>, dim3(grid), dim3(kBlock), 0, stream, n, args, mirrors,
                                   static_cast<const float*>(dv), mask, main_d, table);
            PM_HIP(hipGetLastError());
            int flags = 0;
            PM_HIP(hipMemcpyAsync(&flags, table + kMaxDia, sizeof(flags), hipMemcpyDeviceToHost, stream));
            PM_HIP(hipStreamSynchronize(stream));
            if ((flags & 8) == 0)
            {
                for (int d = 0; d < args.ndiag; ++d)
                    if (args.off[d] < 0)
                    {
                        args.slot[d] = mirrors.of[d];
                        args.shift[d] = args.off[d];
                    }
                out->streams -= lower;
            }
            if ((flags & 16) == 0 && main_d >= 0)
            {
                args.slot[main_d] = -1;
                --out->streams;
            }
        }
    }
    out->active = true;
    out->args = args;
    out->values = dv;
    out->mask = mask;
    return PM_OK;
}

} // namespace pm
