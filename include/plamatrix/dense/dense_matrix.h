// Column-major dense containers (dense/dense_matrix.h:14-287 of the
// reference), reduced to what the sparse solve path uses: vectors and small
// column-major matrices on the host or in HBM, with the reference's
// allocation kinds (page-locked host storage, stream-ordered device storage)
// and its synchronous and stream-ordered transfers.
#pragma once

#include <algorithm>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "plamatrix/core/status.h"
#include "plamatrix/core/storage.h"
#include "plamatrix/core/types.h"

namespace plamatrix
{

template <typename Scalar, Device Dev>
class DenseMatrix;

namespace detail
{
inline std::size_t denseCount(Index rows, Index cols)
{
    if (rows < 0 || cols < 0) throw std::invalid_argument("DenseMatrix dimensions must be non-negative");
    if (cols > 0 && rows > std::numeric_limits<Index>::max() / cols)
        throw std::overflow_error("DenseMatrix element count overflows Index");
    return static_cast<std::size_t>(rows * cols);
}

inline void checkDenseTransfer(const char* op, Index rows, Index cols, Index want_rows, Index want_cols)
{
    if (rows != want_rows || cols != want_cols)
        throw std::runtime_error(std::string("DenseMatrix::") + op + " output dimension mismatch: output is " +
                                 std::to_string(rows) + "x" + std::to_string(cols) + ", expected " +
                                 std::to_string(want_rows) + "x" + std::to_string(want_cols));
}
} // namespace detail

template <typename Scalar>
class DenseMatrix<Scalar, Device::CPU>
{
public:
    DenseMatrix() = default;
    DenseMatrix(Index rows, Index cols)
        : _rows(rows), _cols(cols), _data(detail::denseCount(rows, cols), detail::HostKind::Pageable, true)
    {
    }

    /// Page-locked host storage, zero-initialised: the source or target of
    /// asynchronous copies that overlap with host work.
    static DenseMatrix pinned(Index rows, Index cols)
    {
        DenseMatrix m;
        m._rows = rows;
        m._cols = cols;
        m._data = detail::HostArray<Scalar>(detail::denseCount(rows, cols), detail::HostKind::Pinned, true);
        return m;
    }

    Index rows() const noexcept { return _rows; }
    Index cols() const noexcept { return _cols; }
    Index size() const noexcept { return _rows * _cols; }
    Scalar* data() noexcept { return _data.data(); }
    const Scalar* data() const noexcept { return _data.data(); }
    bool isPinnedHost() const noexcept { return _data.data() && _data.kind() == detail::HostKind::Pinned; }

    Scalar operator()(Index row, Index col) const { return _data.data()[offset(row, col)]; }
    Scalar& operator()(Index row, Index col) { return _data.data()[offset(row, col)]; }
    void setValue(Index row, Index col, Scalar value) { (*this)(row, col) = value; }
    Scalar getValue(Index row, Index col) const { return (*this)(row, col); }
    void fill(Scalar value) { std::fill_n(_data.data(), static_cast<std::size_t>(size()), value); }

    DenseMatrix<Scalar, Device::GPU> toGpu() const;
    /// Copy into new GPU storage on `stream`; synchronise before the host
    /// storage changes or the result is used on another stream.
    DenseMatrix<Scalar, Device::GPU> toGpuAsync(hipStream_t stream = nullptr) const;
    /// Copy into existing GPU storage of the same shape on `stream`.
    void copyToGpuAsync(DenseMatrix<Scalar, Device::GPU>& output, hipStream_t stream = nullptr) const;

private:
    std::size_t offset(Index row, Index col) const
    {
        if (row < 0 || row >= _rows || col < 0 || col >= _cols)
            throw std::out_of_range("DenseMatrix index is out of bounds");
        return static_cast<std::size_t>(row + col * _rows);
    }

    Index _rows = 0;
    Index _cols = 0;
    detail::HostArray<Scalar> _data;
};

// Device storage is move-only and freed on destruction.  Ordinary storage is
// zero-initialised; uninitialized() and uninitializedAsync() skip the fill.
template <typename Scalar>
class DenseMatrix<Scalar, Device::GPU>
{
public:
    DenseMatrix() = default;
    DenseMatrix(Index rows, Index cols)
        : _rows(rows), _cols(cols),
          _data(detail::denseCount(rows, cols), detail::GpuKind::Ordinary, nullptr)
    {
        if (_data.data())
        {
            detail::checkHip(hipMemset(_data.data(), 0, sizeof(Scalar) * _data.size()), "hipMemset");
            detail::checkHip(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
        }
    }

    static DenseMatrix uninitialized(Index rows, Index cols)
    {
        DenseMatrix m;
        m._rows = rows;
        m._cols = cols;
        m._data = detail::GpuArray<Scalar>(detail::denseCount(rows, cols), detail::GpuKind::Ordinary, nullptr);
        return m;
    }

    /// Stream-ordered storage owned by `stream`: every use must be ordered on
    /// that stream, and closeAsyncAllocation() enqueues its release there
    /// (dense_matrix.h:55-87 of the reference).
    static DenseMatrix uninitializedAsync(Index rows, Index cols, hipStream_t stream = nullptr)
    {
        DenseMatrix m;
        m._rows = rows;
        m._cols = cols;
        m._data = detail::GpuArray<Scalar>(detail::denseCount(rows, cols), detail::GpuKind::StreamOrdered, stream);
        return m;
    }

    Index rows() const noexcept { return _rows; }
    Index cols() const noexcept { return _cols; }
    Index size() const noexcept { return _rows * _cols; }
    Scalar* data() noexcept { return _data.data(); }
    const Scalar* data() const noexcept { return _data.data(); }

    bool isAsyncAllocation() const noexcept
    {
        return _data.data() && _data.kind() == detail::GpuKind::StreamOrdered;
    }
    hipStream_t asyncAllocationStream() const noexcept { return isAsyncAllocation() ? _data.stream() : nullptr; }

    /// Enqueue the release of stream-ordered storage on its stream; the matrix
    /// becomes 0x0.  Empty matrices are a no-op; ordinary storage is a
    /// std::logic_error (device_matrix.h:134-175 of the reference).
    void closeAsyncAllocation()
    {
        if (!_data.data())
        {
            _rows = _cols = 0;
            _data = detail::GpuArray<Scalar>();
            return;
        }
        if (_data.kind() != detail::GpuKind::StreamOrdered)
            throw std::logic_error("DenseMatrix::closeAsyncAllocation requires a stream-ordered GPU allocation");
        _data.releaseAsync();
        _rows = _cols = 0;
    }

    void fill(Scalar value)
    {
        std::vector<Scalar> host(static_cast<std::size_t>(size()), value);
        if (!host.empty())
            detail::checkHip(hipMemcpy(_data.data(), host.data(), sizeof(Scalar) * host.size(), hipMemcpyHostToDevice),
                             "hipMemcpy");
    }

    DenseMatrix<Scalar, Device::CPU> toCpu() const
    {
        DenseMatrix<Scalar, Device::CPU> out(_rows, _cols);
        if (size() > 0)
            detail::checkHip(hipMemcpy(out.data(), _data.data(), sizeof(Scalar) * static_cast<std::size_t>(size()),
                                       hipMemcpyDeviceToHost),
                             "hipMemcpy");
        return out;
    }

    /// Copy to new host storage on `stream`; synchronise before reading it.
    DenseMatrix<Scalar, Device::CPU> toCpuAsync(hipStream_t stream = nullptr) const
    {
        DenseMatrix<Scalar, Device::CPU> out(_rows, _cols);
        copyToCpuAsync(out, stream);
        return out;
    }

    void copyToCpuAsync(DenseMatrix<Scalar, Device::CPU>& output, hipStream_t stream = nullptr) const
    {
        detail::checkDenseTransfer("copyToCpuAsync", output.rows(), output.cols(), _rows, _cols);
        if (size() == 0) return;
        detail::checkHip(hipMemcpyAsync(output.data(), _data.data(), sizeof(Scalar) * static_cast<std::size_t>(size()),
                                        hipMemcpyDeviceToHost, stream),
                         "hipMemcpyAsync");
    }

private:
    Index _rows = 0;
    Index _cols = 0;
    detail::GpuArray<Scalar> _data;
};

template <typename Scalar>
DenseMatrix<Scalar, Device::GPU> DenseMatrix<Scalar, Device::CPU>::toGpu() const
{
    auto out = DenseMatrix<Scalar, Device::GPU>::uninitialized(_rows, _cols);
    if (size() > 0)
        detail::checkHip(hipMemcpy(out.data(), _data.data(), sizeof(Scalar) * static_cast<std::size_t>(size()),
                                   hipMemcpyHostToDevice),
                         "hipMemcpy");
    return out;
}

template <typename Scalar>
DenseMatrix<Scalar, Device::GPU> DenseMatrix<Scalar, Device::CPU>::toGpuAsync(hipStream_t stream) const
{
    auto out = DenseMatrix<Scalar, Device::GPU>::uninitialized(_rows, _cols);
    copyToGpuAsync(out, stream);
    return out;
}

template <typename Scalar>
void DenseMatrix<Scalar, Device::CPU>::copyToGpuAsync(DenseMatrix<Scalar, Device::GPU>& output,
                                                      hipStream_t stream) const
{
    detail::checkDenseTransfer("copyToGpuAsync", output.rows(), output.cols(), _rows, _cols);
    if (size() == 0) return;
    detail::checkHip(hipMemcpyAsync(output.data(), _data.data(), sizeof(Scalar) * static_cast<std::size_t>(size()),
                                    hipMemcpyHostToDevice, stream),
                     "hipMemcpyAsync");
}

} // namespace plamatrix
