// Shared helpers of the gfx950 library: status plumbing for the C ABI,
// deterministic wavefront/workgroup reductions, XCD-aware block numbering and
// the grow-only device buffers the workspaces own.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "plamatrix_amd.h"

namespace pm
{

constexpr int kWave = 64;       // CDNA wavefront width
constexpr int kBlock = 256;     // threads per workgroup for streaming kernels
constexpr int kXcds = 8;        // MI355X: 8 XCDs, blocks dispatched round-robin
constexpr int kMaxPartials = 2048;  // upper bound on workgroups that publish partial sums

// ---- host-side status ----------------------------------------------------
void set_error(int code, const std::string& message);
int fail(int code, const char* fmt, ...);
int check_hip(hipError_t err, const char* what, const char* file, int line);

#define PM_HIP(call)                                                                     \
    do                                                                                   \
    {                                                                                    \
        hipError_t pm_err_ = (call);                                                     \
        if (pm_err_ != hipSuccess) return ::pm::check_hip(pm_err_, #call, __FILE__, __LINE__); \
    } while (0)

#define PM_TRY(expr)                         \
    do                                       \
    {                                        \
        int pm_status_ = (expr);             \
        if (pm_status_ != PM_OK) return pm_status_; \
    } while (0)

// ---- XCD-aware block numbering -------------------------------------------
// The dispatcher hands workgroup b to XCD (b % 8).  Renumbering so that each
// XCD walks one contiguous range of logical blocks keeps neighbouring rows (and
// the x entries they gather) in the same L2.
__device__ __forceinline__ int64_t xcd_logical_block(int64_t b, int64_t nb)
{
    const int64_t q = nb / kXcds;
    const int64_t r = nb % kXcds;
    const int64_t xcd = b % kXcds;
    const int64_t idx = b / kXcds;
    return xcd * q + (xcd < r ? xcd : r) + idx;
}

// LDS hand-off between lanes of one wavefront.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- deterministic reductions ---------------------------------------------
template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Sum within aligned groups of L lanes (L a power of two <= 64); every lane of
// the group ends with the group sum.  Fixed butterfly order.
template <int L, typename T>
__device__ __forceinline__ T group_sum(T v)
{
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Workgroup sum of NV doubles; result valid in thread 0.  `lds` needs
// NV * (blockDim/64) doubles.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds)
{
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    const int waves = blockDim.x / kWave;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0)
    {
#pragma unroll
        for (int k = 0; k < NV; ++k) lds[k * 16 + wave] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
#pragma unroll
        for (int k = 0; k < NV; ++k)
        {
            double s = 0.0;
            for (int w = 0; w < waves; ++w) s += lds[k * 16 + w];
            v[k] = s;
        }
    }
}

// Publish this workgroup's NV partial sums; the last workgroup to arrive sums
// all partials in index order (so the result does not depend on arrival order)
// and returns true in every one of its threads, with the totals in `total`
// (valid in all threads of that workgroup).  `counter` returns to zero.
template <int NV>
__device__ __forceinline__ bool last_block_total(double (&v)[NV], double* partials, unsigned* counter,
                                                 double (&total)[NV])
{
    __shared__ double lds[NV * 16];
    __shared__ int is_last;
    __shared__ double result[NV];
    const int nblocks = gridDim.x;
    block_sum<NV>(v, lds);
    if (threadIdx.x == 0)
    {
#pragma unroll
        for (int k = 0; k < NV; ++k) partials[k * kMaxPartials + blockIdx.x] = v[k];
        __threadfence();
        const unsigned arrived = atomicAdd(counter, 1u);
        is_last = (arrived == static_cast<unsigned>(nblocks - 1));
    }
    __syncthreads();
    if (!is_last) return false;
    __threadfence();
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k)
    {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
            s += partials[k * kMaxPartials + b];
        acc[k] = s;
    }
    __syncthreads();
    block_sum<NV>(acc, lds);
    if (threadIdx.x == 0)
    {
#pragma unroll
        for (int k = 0; k < NV; ++k) result[k] = acc[k];
        *counter = 0u;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) total[k] = result[k];
    return true;
}

// ---- grow-only device buffer ---------------------------------------------
struct DeviceBuffer
{
    void* ptr = nullptr;
    size_t bytes = 0;
    int reserve(size_t need);
    void release();
    template <typename T>
    T* as() const
    {
        return static_cast<T*>(ptr);
    }
};

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

} // namespace pm
