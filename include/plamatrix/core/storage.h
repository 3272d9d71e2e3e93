// Owning host and device arrays with the allocation kinds of the reference's
// containers (core/device_matrix.h:16-282, core/allocator.h): pageable or
// page-locked host memory, ordinary or stream-ordered device memory.  The
// containers in dense/ and sparse/ are built on these two classes.
#pragma once

#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <new>
#include <stdexcept>
#include <utility>

#include <hip/hip_runtime_api.h>

#include "plamatrix/core/status.h"

namespace plamatrix::detail
{

template <typename T>
std::size_t checkedBytes(std::size_t count)
{
    if (count > std::numeric_limits<std::size_t>::max() / sizeof(T))
        throw std::overflow_error("allocation byte count overflows size_t");
    return count * sizeof(T);
}

enum class HostKind
{
    Pageable,
    Pinned,  // hipHostMalloc: the source or target of asynchronous copies
};

template <typename T>
class HostArray
{
public:
    HostArray() = default;
    HostArray(std::size_t count, HostKind kind, bool zero) : _count(count), _kind(kind)
    {
        const std::size_t bytes = checkedBytes<T>(count);
        if (bytes == 0) return;
        if (kind == HostKind::Pinned)
            checkHip(hipHostMalloc(reinterpret_cast<void**>(&_data), bytes, hipHostMallocDefault), "hipHostMalloc");
        else
        {
            _data = static_cast<T*>(std::malloc(bytes));
            if (!_data) throw std::bad_alloc();
        }
        if (zero) std::memset(static_cast<void*>(_data), 0, bytes);
    }
    ~HostArray() { release(); }
    HostArray(const HostArray&) = delete;
    HostArray& operator=(const HostArray&) = delete;
    HostArray(HostArray&& o) noexcept
        : _data(std::exchange(o._data, nullptr)), _count(std::exchange(o._count, 0)),
          _kind(std::exchange(o._kind, HostKind::Pageable))
    {
    }
    HostArray& operator=(HostArray&& o) noexcept
    {
        if (this != &o)
        {
            release();
            _data = std::exchange(o._data, nullptr);
            _count = std::exchange(o._count, 0);
            _kind = std::exchange(o._kind, HostKind::Pageable);
        }
        return *this;
    }

    T* data() const noexcept { return _data; }
    std::size_t size() const noexcept { return _count; }
    HostKind kind() const noexcept { return _kind; }

private:
    void release() noexcept
    {
        if (!_data) return;
        if (_kind == HostKind::Pinned)
            (void)hipHostFree(_data);
        else
            std::free(_data);
        _data = nullptr;
    }

    T* _data = nullptr;
    std::size_t _count = 0;
    HostKind _kind = HostKind::Pageable;
};

enum class GpuKind
{
    Ordinary,       // hipMalloc
    StreamOrdered,  // hipMallocAsync on the owning stream
};

template <typename T>
class GpuArray
{
public:
    GpuArray() = default;
    GpuArray(std::size_t count, GpuKind kind, hipStream_t stream) : _count(count), _kind(kind), _stream(stream)
    {
        const std::size_t bytes = checkedBytes<T>(count);
        if (bytes == 0) return;
        if (kind == GpuKind::StreamOrdered)
            checkHip(hipMallocAsync(reinterpret_cast<void**>(&_data), bytes, stream), "hipMallocAsync");
        else
            checkHip(hipMalloc(reinterpret_cast<void**>(&_data), bytes), "hipMalloc");
    }
    ~GpuArray() { releaseNoThrow(); }
    GpuArray(const GpuArray&) = delete;
    GpuArray& operator=(const GpuArray&) = delete;
    GpuArray(GpuArray&& o) noexcept
        : _data(std::exchange(o._data, nullptr)), _count(std::exchange(o._count, 0)),
          _kind(std::exchange(o._kind, GpuKind::Ordinary)), _stream(std::exchange(o._stream, nullptr))
    {
    }
    GpuArray& operator=(GpuArray&& o) noexcept
    {
        if (this != &o)
        {
            releaseNoThrow();
            _data = std::exchange(o._data, nullptr);
            _count = std::exchange(o._count, 0);
            _kind = std::exchange(o._kind, GpuKind::Ordinary);
            _stream = std::exchange(o._stream, nullptr);
        }
        return *this;
    }

    T* data() const noexcept { return _data; }
    std::size_t size() const noexcept { return _count; }
    GpuKind kind() const noexcept { return _kind; }
    hipStream_t stream() const noexcept { return _stream; }

    // Enqueue the free of stream-ordered storage on its owning stream.  On
    // failure the array keeps its storage, so the caller can retry.
    void releaseAsync()
    {
        if (_data) checkHip(hipFreeAsync(_data, _stream), "hipFreeAsync");
        _data = nullptr;
        _count = 0;
        _kind = GpuKind::Ordinary;
        _stream = nullptr;
    }

private:
    // Destructor path: the owning stream may already be gone, so stream-ordered
    // storage is freed after a device-wide synchronisation on the null stream.
    void releaseNoThrow() noexcept
    {
        if (!_data) return;
        if (_kind == GpuKind::StreamOrdered)
        {
            if (hipDeviceSynchronize() == hipSuccess && hipFreeAsync(_data, nullptr) == hipSuccess)
            {
                (void)hipStreamSynchronize(nullptr);
                _data = nullptr;
                return;
            }
            (void)hipGetLastError();
        }
        (void)hipFree(_data);
        _data = nullptr;
    }

    T* _data = nullptr;
    std::size_t _count = 0;
    GpuKind _kind = GpuKind::Ordinary;
    hipStream_t _stream = nullptr;
};

} // namespace plamatrix::detail
