// SpMV, CSR validation, index compaction and the FP64 true residual.
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "spmv_kernels.hpp"
#include "tuning.hpp"

namespace pm
{

int check_csr(const pm_csr* a, const char* op)
{
    if (!a) return fail(PM_INVALID_ARGUMENT, "%s: null CSR descriptor", op);
    if (a->rows < 0 || a->cols < 0 || a->nnz < 0)
        return fail(PM_INVALID_ARGUMENT, "%s: CSR dimensions must be non-negative", op);
    if (a->dtype != PM_F32 && a->dtype != PM_F64) return fail(PM_INVALID_ARGUMENT, "%s: unknown dtype %d", op, a->dtype);
    if (!a->row_offsets && !a->row_offsets32) return fail(PM_INVALID_ARGUMENT, "%s: CSR row offsets are null", op);
    if (a->nnz > 0 && ((!a->col_indices && !a->col_indices32) || !a->values))
        return fail(PM_INVALID_ARGUMENT, "%s: CSR arrays are null", op);
    return PM_OK;
}

bool use_compact(const pm_csr* a) { return a->row_offsets32 != nullptr && a->col_indices32 != nullptr; }

template <typename S>
struct SpmvRun
{
    template <int L>
    struct F
    {
        static int run(const pm_csr* a, const void* x, void* y, hipStream_t stream)
        {
            const int grid = static_cast<int>(std::min<int64_t>(ceil_div(a->rows, kBlock / L), 1 << 20));
            if (use_compact(a))
                hipLaunchKernelGGL((k_spmv_lanes<S, int32_t, L>), dim3(grid), dim3(kBlock), 0, stream, a->rows,
                                   a->row_offsets32, a->col_indices32, static_cast<const S*>(a->values),
                                   static_cast<const S*>(x), static_cast<S*>(y));
            else
                hipLaunchKernelGGL((k_spmv_lanes<S, int64_t, L>), dim3(grid), dim3(kBlock), 0, stream, a->rows,
                                   a->row_offsets, a->col_indices, static_cast<const S*>(a->values),
                                   static_cast<const S*>(x), static_cast<S*>(y));
            PM_HIP(hipGetLastError());
            return PM_OK;
        }
    };
};

template <typename S>
int spmv_stream(const pm_csr* a, const void* x, void* y, hipStream_t stream)
{
    const int grid = static_cast<int>(std::min<int64_t>(ceil_div(chunk_count(a->rows), kWavesPerBlock), kSpmvGrid));
    if (use_compact(a))
        hipLaunchKernelGGL((k_spmv<S, int32_t>), dim3(grid), dim3(kBlock), 0, stream, a->rows, a->row_offsets32,
                           a->col_indices32, static_cast<const S*>(a->values), static_cast<const S*>(x),
                           static_cast<S*>(y));
    else
        hipLaunchKernelGGL((k_spmv<S, int64_t>), dim3(grid), dim3(kBlock), 0, stream, a->rows, a->row_offsets,
                           a->col_indices, static_cast<const S*>(a->values), static_cast<const S*>(x),
                           static_cast<S*>(y));
    PM_HIP(hipGetLastError());
    return PM_OK;
}

// Y = A B in tiles of kSpmmTile columns through a packed row-major copy of
// the tile's columns of B in the workspace scratch.
template <typename S>
int spmm_tiles_into(const pm_csr* a, const S* b, int64_t k, S* y, S* bt, hipStream_t stream)
{
    const int64_t cols = a->cols > 0 ? a->cols : 1;
    const int grid = static_cast<int>(std::min<int64_t>(ceil_div(chunk_count(a->rows), kWavesPerBlock), kSpmvGrid));
    const int pgrid = static_cast<int>(std::min<int64_t>(ceil_div(cols, kBlock), 8192));
    for (int64_t j0 = 0; j0 < k; j0 += kSpmmTile)
    {
        const int64_t kt = k - j0 < kSpmmTile ? k - j0 : kSpmmTile;
        if (a->cols > 0) hipLaunchKernelGGL(k_spmm_pack<S>, dim3(pgrid), dim3(kBlock), 0, stream, a->cols, k, j0, b, bt);
        if (use_compact(a))
            hipLaunchKernelGGL((k_spmm_tile<S, int32_t>), dim3(grid), dim3(kBlock), 0, stream, a->rows, kt,
                               a->row_offsets32, a->col_indices32, static_cast<const S*>(a->values), bt,
                               y + j0 * a->rows);
        else
            hipLaunchKernelGGL((k_spmm_tile<S, int64_t>), dim3(grid), dim3(kBlock), 0, stream, a->rows, kt,
                               a->row_offsets, a->col_indices, static_cast<const S*>(a->values), bt,
                               y + j0 * a->rows);
        PM_HIP(hipGetLastError());
    }
    return PM_OK;
}

template <typename S>
int spmm_tiles(const pm_csr* a, const S* b, int64_t k, S* y, pm_workspace* ws, hipStream_t stream)
{
    const int64_t cols = a->cols > 0 ? a->cols : 1;
    PM_TRY(ws->scratch.reserve(sizeof(S) * static_cast<size_t>(cols) * kSpmmTile + 256));
    return spmm_tiles_into<S>(a, b, k, y, ws->scratch.as<S>(), stream);
}

// ---- allocation-free SpMM --------------------------------------------------
//
// The packed-B tile of the SpMM lives in `scratch`, a file-scope device
// array: each GPU the library runs on holds its own copy (kNoAllocScratchBytes)
// from the first load of the library until the process ends.  Host code never
// uses the array's name as an address (on the host it is only a stand-in
// symbol): its device address is looked up once per device with
// hipGetSymbolAddress and kept in g_scratch_ptr.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mutex;
std::vector<void*> g_scratch_ptr;  // per device, null until looked up

int scratch_address(void** out)
{
    int dev = 0;
    PM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (static_cast<size_t>(dev) >= g_scratch_ptr.size()) g_scratch_ptr.resize(static_cast<size_t>(dev) + 1, nullptr);
    if (!g_scratch_ptr[static_cast<size_t>(dev)])
    {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(PM_DEVICE_ERROR, "spmm (allocation-free): no device address for the scratch array on "
                        "device %d: %s", dev, hipGetErrorString(e));
        g_scratch_ptr[static_cast<size_t>(dev)] = ptr;
    }
    *out = g_scratch_ptr[static_cast<size_t>(dev)];
    return PM_OK;
}

// lanes == 0: wave-stream kernel; lanes >= 1: lanes-per-row kernel.
int spmv_launch(const pm_csr* a, const void* x, void* y, int lanes, hipStream_t stream)
{
    if (a->rows == 0) return PM_OK;
    if (lanes == 0) return a->dtype == PM_F32 ? spmv_stream<float>(a, x, y, stream) : spmv_stream<double>(a, x, y, stream);
    if (a->dtype == PM_F32) return dispatch_lanes<SpmvRun<float>::F>(lanes, a, x, y, stream);
    return dispatch_lanes<SpmvRun<double>::F>(lanes, a, x, y, stream);
}

// ---- structural validation -------------------------------------------------
struct ValidateOut
{
    unsigned long long bad_ends;   // 1 if offsets[0] != 0 or offsets[rows] != nnz
    unsigned long long bad_row;    // first row with invalid offsets
    unsigned long long bad_col;    // first position with column out of range
    unsigned long long bad_value;  // first position with a non-finite value
};

template <typename S>
__global__ void k_validate(int64_t rows, int64_t cols, int64_t nnz, const int64_t* ro, const int64_t* ci, const S* v,
                           ValidateOut* out)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    const int64_t tid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (tid == 0 && (ro[0] != 0 || ro[rows] != nnz)) out->bad_ends = 1;
    for (int64_t row = tid; row < rows; row += stride)
    {
        const int64_t b = ro[row], e = ro[row + 1];
        if (b < 0 || e < b || e > nnz) atomicMin(&out->bad_row, static_cast<unsigned long long>(row));
    }
    for (int64_t p = tid; p < nnz; p += stride)
    {
        const int64_t c = ci[p];
        if (c < 0 || c >= cols) atomicMin(&out->bad_col, static_cast<unsigned long long>(p));
        if (!isfinite(static_cast<double>(v[p]))) atomicMin(&out->bad_value, static_cast<unsigned long long>(p));
    }
}

__global__ void k_validate_init(ValidateOut* out)
{
    out->bad_ends = 0;
    out->bad_row = ~0ull;
    out->bad_col = ~0ull;
    out->bad_value = ~0ull;
}

__global__ void k_compact(int64_t rows, int64_t nnz, const int64_t* ro, const int64_t* ci, int32_t* ro32,
                          int32_t* ci32)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    const int64_t tid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    for (int64_t i = tid; i <= rows; i += stride) ro32[i] = static_cast<int32_t>(ro[i]);
    for (int64_t i = tid; i < nnz; i += stride) ci32[i] = static_cast<int32_t>(ci[i]);
}

template <typename S, typename V>
int relres_stream(const pm_csr* a, const void* b, const void* x, double* partials, unsigned* counter, double* out,
                  hipStream_t stream)
{
    const int grid = solver_grid_chunks(a->rows);
    if (use_compact(a))
        hipLaunchKernelGGL((k_relres<S, V, int32_t>), dim3(grid), dim3(kBlock), 0, stream, a->rows, a->row_offsets32,
                           a->col_indices32, static_cast<const S*>(a->values), static_cast<const V*>(b),
                           static_cast<const V*>(x), partials, counter, out);
    else
        hipLaunchKernelGGL((k_relres<S, V, int64_t>), dim3(grid), dim3(kBlock), 0, stream, a->rows, a->row_offsets,
                           a->col_indices, static_cast<const S*>(a->values), static_cast<const V*>(b),
                           static_cast<const V*>(x), partials, counter, out);
    PM_HIP(hipGetLastError());
    return PM_OK;
}

__global__ void k_zero_counter(unsigned* counter) { *counter = 0u; }

int relres_launch(const pm_csr* a, const void* b, const void* x, int32_t vec_dtype, double* partials,
                  unsigned* counter, double* out, hipStream_t stream)
{
    if (a->dtype == PM_F32 && vec_dtype == PM_F32)
        return relres_stream<float, float>(a, b, x, partials, counter, out, stream);
    if (a->dtype == PM_F32) return relres_stream<float, double>(a, b, x, partials, counter, out, stream);
    if (vec_dtype == PM_F64) return relres_stream<double, double>(a, b, x, partials, counter, out, stream);
    return fail(PM_INVALID_ARGUMENT, "true residual: FP32 vectors with an FP64 matrix are not supported");
}

} // namespace pm

using namespace pm;

extern "C" {

int pm_spmv_lanes(const pm_csr* a, const void* x, void* y, int32_t lanes, pm_workspace* ws, hipStream_t stream)
{
    PM_TRY(check_csr(a, "spmv"));
    if (a->rows > 0 && !y) return fail(PM_INVALID_ARGUMENT, "spmv: null output");
    if (a->cols > 0 && !x) return fail(PM_INVALID_ARGUMENT, "spmv: null input");
    if (x != nullptr && x == y) return fail(PM_INVALID_ARGUMENT, "spmv input and output data must not alias");
    if (ws) PM_TRY(ws->bind(stream));
    return spmv_launch(a, x, y, lanes, stream);
}

int pm_spmv(const pm_csr* a, const void* x, void* y, pm_workspace* ws, hipStream_t stream)
{
    PM_TRY(check_csr(a, "spmv"));
    return pm_spmv_lanes(a, x, y, tuned_lanes(a), ws, stream);
}

int pm_spmm(const pm_csr* a, const void* b, int64_t k, void* y, pm_workspace* ws, hipStream_t stream)
{
    PM_TRY(check_csr(a, "spmm"));
    if (k < 0) return fail(PM_INVALID_ARGUMENT, "spmm: negative column count");
    if (!ws) return fail(PM_INVALID_ARGUMENT, "spmm: null workspace");
    if (a->rows > 0 && k > 0 && !y) return fail(PM_INVALID_ARGUMENT, "spmm: null output");
    if (a->cols > 0 && k > 0 && !b) return fail(PM_INVALID_ARGUMENT, "spmm: null input");
    if (b != nullptr && b == y) return fail(PM_INVALID_ARGUMENT, "spmm input and output data must not alias");
    PM_TRY(ws->bind(stream));
    if (a->rows == 0 || k == 0) return PM_OK;
    return a->dtype == PM_F32 ? spmm_tiles<float>(a, static_cast<const float*>(b), k, static_cast<float*>(y), ws, stream)
                              : spmm_tiles<double>(a, static_cast<const double*>(b), k, static_cast<double*>(y), ws,
                                                   stream);
}

int pm_spmm_noalloc(const pm_csr* a, const void* b, int64_t k, void* y, hipStream_t stream)
{
    PM_TRY(check_csr(a, "spmm (allocation-free)"));
    if (k < 0) return fail(PM_INVALID_ARGUMENT, "spmm (allocation-free): negative column count");
    if (a->rows > 0 && k > 0 && !y) return fail(PM_INVALID_ARGUMENT, "spmm (allocation-free): null output");
    if (a->cols > 0 && k > 0 && !b) return fail(PM_INVALID_ARGUMENT, "spmm (allocation-free): null input");
    if (b != nullptr && b == y)
        return fail(PM_INVALID_ARGUMENT, "spmm (allocation-free): input and output data must not alias");
    const size_t es = a->dtype == PM_F64 ? sizeof(double) : sizeof(float);
    const size_t need = es * static_cast<size_t>(a->cols > 0 ? a->cols : 1) * kSpmmTile;
    if (need > kNoAllocScratchBytes)
        return fail(PM_OVERFLOW, "spmm (allocation-free): %lld columns need %zu bytes of packed tile, above the "
                    "%zu-byte scratch (at most %zu columns for this dtype)", static_cast<long long>(a->cols), need,
                    kNoAllocScratchBytes, kNoAllocScratchBytes / (es * kSpmmTile));
    if (a->rows == 0 || k == 0) return PM_OK;
    void* bt = nullptr;
    PM_TRY(scratch_address(&bt));
    return a->dtype == PM_F32
               ? spmm_tiles_into<float>(a, static_cast<const float*>(b), k, static_cast<float*>(y),
                                        static_cast<float*>(bt), stream)
               : spmm_tiles_into<double>(a, static_cast<const double*>(b), k, static_cast<double*>(y),
                                         static_cast<double*>(bt), stream);
}

int pm_csr_validate(const pm_csr* a, pm_workspace* ws, hipStream_t stream)
{
    PM_TRY(check_csr(a, "validateStructure"));
    if (!a->row_offsets || (a->nnz > 0 && !a->col_indices))
        return fail(PM_INVALID_ARGUMENT, "validateStructure needs the int64 CSR arrays");
    if (!ws) return fail(PM_INVALID_ARGUMENT, "validateStructure: null workspace");
    PM_TRY(ws->bind(stream));
    PM_TRY(ws->scratch.reserve(4096));
    ValidateOut* out = ws->scratch.as<ValidateOut>();
    hipLaunchKernelGGL(k_validate_init, dim3(1), dim3(1), 0, stream, out);
    const int64_t work = a->rows + 1 > a->nnz ? a->rows + 1 : a->nnz;
    const int grid = static_cast<int>(std::min<int64_t>(ceil_div(work, kBlock), 4096));
    if (a->dtype == PM_F32)
        hipLaunchKernelGGL(k_validate<float>, dim3(grid), dim3(kBlock), 0, stream, a->rows, a->cols, a->nnz,
                           a->row_offsets, a->col_indices, static_cast<const float*>(a->values), out);
    else
        hipLaunchKernelGGL(k_validate<double>, dim3(grid), dim3(kBlock), 0, stream, a->rows, a->cols, a->nnz,
                           a->row_offsets, a->col_indices, static_cast<const double*>(a->values), out);
    PM_HIP(hipGetLastError());
    ValidateOut host{};
    PM_HIP(hipMemcpyAsync(&host, out, sizeof(host), hipMemcpyDeviceToHost, stream));
    PM_HIP(hipStreamSynchronize(stream));
    if (host.bad_ends) return fail(PM_INVALID_ARGUMENT, "CSRMatrix row offsets must start at zero and end at nnz");
    if (host.bad_row != ~0ull)
        return fail(PM_INVALID_ARGUMENT, "CSRMatrix invalid row offsets at row %llu", host.bad_row);
    if (host.bad_col != ~0ull || host.bad_value != ~0ull)
    {
        if (host.bad_col <= host.bad_value)
            return fail(PM_INVALID_ARGUMENT, "CSRMatrix column index out of range at position %llu", host.bad_col);
        return fail(PM_INVALID_ARGUMENT, "CSRMatrix value is not finite at position %llu", host.bad_value);
    }
    return PM_OK;
}

int pm_csr_compact_indices(const pm_csr* a, int32_t* ro32, int32_t* ci32, hipStream_t stream)
{
    PM_TRY(check_csr(a, "compact indices"));
    if (!a->row_offsets || (a->nnz > 0 && !a->col_indices))
        return fail(PM_INVALID_ARGUMENT, "compact indices needs the int64 CSR arrays");
    if (a->nnz > INT32_MAX || a->cols > INT32_MAX || a->rows >= INT32_MAX)
        return fail(PM_OVERFLOW, "CSR does not fit 32-bit indices (rows %lld, cols %lld, nnz %lld)",
                    static_cast<long long>(a->rows), static_cast<long long>(a->cols), static_cast<long long>(a->nnz));
    if (!ro32 || (a->nnz > 0 && !ci32)) return fail(PM_INVALID_ARGUMENT, "compact indices: null output");
    const int64_t work = std::max(a->rows + 1, a->nnz);
    const int grid = static_cast<int>(std::min<int64_t>(ceil_div(work, kBlock), 8192));
    hipLaunchKernelGGL(k_compact, dim3(grid), dim3(kBlock), 0, stream, a->rows, a->nnz, a->row_offsets,
                       a->col_indices, ro32, ci32);
    PM_HIP(hipGetLastError());
    return PM_OK;
}

int pm_true_relres(const pm_csr* a, const void* b, const void* x, int32_t vec_dtype, double* relres,
                   pm_workspace* ws, hipStream_t stream)
{
    PM_TRY(check_csr(a, "true residual"));
    if (!relres || !ws) return fail(PM_INVALID_ARGUMENT, "true residual: null output or workspace");
    if (a->rows != a->cols) return fail(PM_INVALID_ARGUMENT, "true residual: matrix must be square");
    PM_TRY(ws->bind(stream));
    if (a->rows == 0)
    {
        *relres = 0.0;
        return PM_OK;
    }
    PM_TRY(ws->scratch.reserve(4096));
    PM_TRY(ws->partials.reserve(sizeof(double) * kMaxPartials * 8));
    unsigned* counter = reinterpret_cast<unsigned*>(ws->scratch.as<char>() + 1024);
    double* out = reinterpret_cast<double*>(ws->scratch.as<char>() + 2048);
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, stream, counter);
    PM_TRY(relres_launch(a, b, x, vec_dtype, ws->partials.as<double>(), counter, out, stream));
    double host[2] = {0, 0};
    PM_HIP(hipMemcpyAsync(host, out, sizeof(host), hipMemcpyDeviceToHost, stream));
    PM_HIP(hipStreamSynchronize(stream));
    *relres = host[1] > 0.0 ? std::sqrt(host[0] / host[1]) : std::sqrt(host[0]);
    return PM_OK;
}

} // extern "C"
