// CSR matrices on the host and in HBM (sparse/csr_matrix.h:22-825 of the
// reference): int64 row offsets and column indices, Scalar values.  The GPU
// matrix also keeps 32-bit index mirrors whenever nnz and cols fit, which is
// what the gfx950 kernels stream.
#pragma once

#include <cmath>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "plamatrix/core/status.h"
#include "plamatrix/core/types.h"
#include "plamatrix_amd.h"

namespace plamatrix
{

template <typename Scalar, Device Dev>
class CSRMatrix;

namespace detail
{
template <typename Scalar>
constexpr int32_t dtypeOf()
{
    static_assert(std::is_same_v<Scalar, float> || std::is_same_v<Scalar, double>,
                  "the HIP path supports float and double values");
    return std::is_same_v<Scalar, double> ? PM_F64 : PM_F32;
}

// RAII device allocation.
template <typename T>
struct DeviceArray
{
    T* ptr = nullptr;
    DeviceArray() = default;
    explicit DeviceArray(std::size_t n)
    {
        if (n) checkHip(hipMalloc(reinterpret_cast<void**>(&ptr), sizeof(T) * n), "hipMalloc");
    }
    ~DeviceArray()
    {
        if (ptr) (void)hipFree(ptr);
    }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    DeviceArray(DeviceArray&& o) noexcept : ptr(std::exchange(o.ptr, nullptr)) {}
    DeviceArray& operator=(DeviceArray&& o) noexcept
    {
        if (this != &o)
        {
            if (ptr) (void)hipFree(ptr);
            ptr = std::exchange(o.ptr, nullptr);
        }
        return *this;
    }
};
} // namespace detail

template <typename Scalar>
class CSRMatrix<Scalar, Device::CPU>
{
public:
    CSRMatrix() = default;
    CSRMatrix(Index rows, Index cols, Index nnz) : _rows(rows), _cols(cols), _nnz(nnz)
    {
        if (rows < 0 || cols < 0 || nnz < 0) throw std::invalid_argument("CSRMatrix dimensions must be non-negative");
        _rowOffsets.assign(static_cast<std::size_t>(rows + 1), 0);
        _colIndices.assign(static_cast<std::size_t>(nnz), 0);
        _values.assign(static_cast<std::size_t>(nnz), Scalar(0));
    }

    Index rows() const noexcept { return _rows; }
    Index cols() const noexcept { return _cols; }
    Index nnz() const noexcept { return _nnz; }
    Scalar* values() noexcept { return _values.data(); }
    const Scalar* values() const noexcept { return _values.data(); }
    Index* colIndices() noexcept { return _colIndices.data(); }
    const Index* colIndices() const noexcept { return _colIndices.data(); }
    Index* rowOffsets() noexcept { return _rowOffsets.data(); }
    const Index* rowOffsets() const noexcept { return _rowOffsets.data(); }

    // validateHostStructure (csr_matrix.h:633-667): offsets start at 0, end at
    // nnz and never decrease; columns in range; values finite.  Column order
    // within a row and duplicates are allowed.
    void validateStructure() const
    {
        if (_rowOffsets.empty() || _rowOffsets.front() != 0 || _rowOffsets.back() != _nnz)
            throw std::invalid_argument("CSRMatrix row offsets must start at 0 and end at nnz");
        for (Index r = 0; r < _rows; ++r)
            if (_rowOffsets[static_cast<std::size_t>(r + 1)] < _rowOffsets[static_cast<std::size_t>(r)])
                throw std::invalid_argument("CSRMatrix row offsets must be non-decreasing");
        for (Index k = 0; k < _nnz; ++k)
        {
            const Index c = _colIndices[static_cast<std::size_t>(k)];
            if (c < 0 || c >= _cols) throw std::invalid_argument("CSRMatrix column index is out of range");
            if (!std::isfinite(_values[static_cast<std::size_t>(k)]))
                throw std::invalid_argument("CSRMatrix values must be finite");
        }
    }

    CSRMatrix<Scalar, Device::GPU> toGpu() const;

private:
    Index _rows = 0, _cols = 0, _nnz = 0;
    std::vector<Index> _rowOffsets;
    std::vector<Index> _colIndices;
    std::vector<Scalar> _values;
};

template <typename Scalar>
class CSRMatrix<Scalar, Device::GPU>
{
public:
    CSRMatrix() = default;
    CSRMatrix(Index rows, Index cols, Index nnz)
        : _rows(rows), _cols(cols), _nnz(nnz), _rowOffsets(static_cast<std::size_t>(rows + 1)),
          _colIndices(static_cast<std::size_t>(nnz)), _values(static_cast<std::size_t>(nnz))
    {
        if (rows < 0 || cols < 0 || nnz < 0) throw std::invalid_argument("CSRMatrix dimensions must be non-negative");
        detail::checkHip(hipMemset(_rowOffsets.ptr, 0, sizeof(Index) * static_cast<std::size_t>(rows + 1)),
                         "hipMemset");
        if (nnz <= INT32_MAX && cols <= INT32_MAX)
        {
            _rowOffsets32 = detail::DeviceArray<int32_t>(static_cast<std::size_t>(rows + 1));
            _colIndices32 = detail::DeviceArray<int32_t>(static_cast<std::size_t>(nnz));
        }
    }

    Index rows() const noexcept { return _rows; }
    Index cols() const noexcept { return _cols; }
    Index nnz() const noexcept { return _nnz; }
    Scalar* values() noexcept
    {
        _mirrorsCurrent = false;
        return _values.ptr;
    }
    const Scalar* values() const noexcept { return _values.ptr; }
    Index* colIndices() noexcept
    {
        _mirrorsCurrent = false;
        return _colIndices.ptr;
    }
    const Index* colIndices() const noexcept { return _colIndices.ptr; }
    Index* rowOffsets() noexcept
    {
        _mirrorsCurrent = false;
        return _rowOffsets.ptr;
    }
    const Index* rowOffsets() const noexcept { return _rowOffsets.ptr; }

    // Device structural validation (validateStructure, csr_matrix.h:197-237);
    // synchronises the stream.
    void validateStructure(hipStream_t stream = nullptr) const
    {
        pm_workspace* ws = nullptr;
        detail::check(pm_workspace_create(&ws));
        const pm_csr d = descriptor(false);
        const int status = pm_csr_validate(&d, ws, stream);
        pm_workspace_destroy(ws);
        detail::check(status);
    }

    CSRMatrix<Scalar, Device::CPU> toCpu() const
    {
        CSRMatrix<Scalar, Device::CPU> out(_rows, _cols, _nnz);
        detail::checkHip(hipMemcpy(out.rowOffsets(), _rowOffsets.ptr, sizeof(Index) * static_cast<std::size_t>(_rows + 1),
                                   hipMemcpyDeviceToHost),
                         "hipMemcpy");
        if (_nnz > 0)
        {
            detail::checkHip(hipMemcpy(out.colIndices(), _colIndices.ptr, sizeof(Index) * static_cast<std::size_t>(_nnz),
                                       hipMemcpyDeviceToHost),
                             "hipMemcpy");
            detail::checkHip(hipMemcpy(out.values(), _values.ptr, sizeof(Scalar) * static_cast<std::size_t>(_nnz),
                                       hipMemcpyDeviceToHost),
                             "hipMemcpy");
        }
        return out;
    }

    // C ABI descriptor.  With `compact` the 32-bit mirrors are refreshed (on
    // `stream`) when the int64 arrays may have changed since the last refresh.
    pm_csr descriptor(bool compact = true, hipStream_t stream = nullptr) const
    {
        pm_csr d{};
        d.rows = _rows;
        d.cols = _cols;
        d.nnz = _nnz;
        d.row_offsets = _rowOffsets.ptr;
        d.col_indices = _colIndices.ptr;
        d.values = _values.ptr;
        d.dtype = detail::dtypeOf<Scalar>();
        if (compact && _rowOffsets32.ptr)
        {
            if (!_mirrorsCurrent)
            {
                detail::check(pm_csr_compact_indices(&d, _rowOffsets32.ptr, _colIndices32.ptr, stream));
                _mirrorsCurrent = true;
            }
            d.row_offsets32 = _rowOffsets32.ptr;
            d.col_indices32 = _colIndices32.ptr;
        }
        return d;
    }

    // The arrays were written by device work that also refreshed the mirrors.
    void markMirrorsCurrent() noexcept { _mirrorsCurrent = true; }

private:
    Index _rows = 0, _cols = 0, _nnz = 0;
    detail::DeviceArray<Index> _rowOffsets;
    detail::DeviceArray<Index> _colIndices;
    detail::DeviceArray<Scalar> _values;
    detail::DeviceArray<int32_t> _rowOffsets32;
    detail::DeviceArray<int32_t> _colIndices32;
    mutable bool _mirrorsCurrent = false;
};

template <typename Scalar>
CSRMatrix<Scalar, Device::GPU> CSRMatrix<Scalar, Device::CPU>::toGpu() const
{
    validateStructure();
    CSRMatrix<Scalar, Device::GPU> out(_rows, _cols, _nnz);
    detail::checkHip(hipMemcpy(out.rowOffsets(), _rowOffsets.data(), sizeof(Index) * _rowOffsets.size(),
                               hipMemcpyHostToDevice),
                     "hipMemcpy");
    if (_nnz > 0)
    {
        detail::checkHip(hipMemcpy(out.colIndices(), _colIndices.data(), sizeof(Index) * _colIndices.size(),
                                   hipMemcpyHostToDevice),
                         "hipMemcpy");
        detail::checkHip(hipMemcpy(out.values(), _values.data(), sizeof(Scalar) * _values.size(),
                                   hipMemcpyHostToDevice),
                         "hipMemcpy");
    }
    return out;
}

} // namespace plamatrix
