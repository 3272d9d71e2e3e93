// Column-major dense containers (dense/dense_matrix.h:14-287 of the
// reference), reduced to what the sparse solve path uses: vectors and small
// column-major matrices on the host or in HBM.
#pragma once

#include <algorithm>
#include <cstddef>
#include <stdexcept>
#include <utility>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "plamatrix/core/status.h"
#include "plamatrix/core/types.h"

namespace plamatrix
{

template <typename Scalar, Device Dev>
class DenseMatrix;

template <typename Scalar>
class DenseMatrix<Scalar, Device::CPU>
{
public:
    DenseMatrix() = default;
    DenseMatrix(Index rows, Index cols) : _rows(rows), _cols(cols)
    {
        if (rows < 0 || cols < 0) throw std::invalid_argument("DenseMatrix dimensions must be non-negative");
        _data.assign(static_cast<std::size_t>(rows * cols), Scalar(0));
    }

    Index rows() const noexcept { return _rows; }
    Index cols() const noexcept { return _cols; }
    Index size() const noexcept { return _rows * _cols; }
    Scalar* data() noexcept { return _data.data(); }
    const Scalar* data() const noexcept { return _data.data(); }

    Scalar operator()(Index row, Index col) const { return _data[static_cast<std::size_t>(offset(row, col))]; }
    Scalar& operator()(Index row, Index col) { return _data[static_cast<std::size_t>(offset(row, col))]; }
    void setValue(Index row, Index col, Scalar value) { (*this)(row, col) = value; }
    Scalar getValue(Index row, Index col) const { return (*this)(row, col); }
    void fill(Scalar value) { std::fill(_data.begin(), _data.end(), value); }

    DenseMatrix<Scalar, Device::GPU> toGpu() const;

private:
    Index offset(Index row, Index col) const
    {
        if (row < 0 || row >= _rows || col < 0 || col >= _cols)
            throw std::out_of_range("DenseMatrix index is out of bounds");
        return row + col * _rows;
    }

    Index _rows = 0;
    Index _cols = 0;
    std::vector<Scalar> _data;
};

// Device storage is zero-initialised, move-only and freed on destruction.
template <typename Scalar>
class DenseMatrix<Scalar, Device::GPU>
{
public:
    DenseMatrix() = default;
    DenseMatrix(Index rows, Index cols) : _rows(rows), _cols(cols)
    {
        if (rows < 0 || cols < 0) throw std::invalid_argument("DenseMatrix dimensions must be non-negative");
        const std::size_t bytes = sizeof(Scalar) * static_cast<std::size_t>(rows * cols);
        if (bytes > 0)
        {
            detail::checkHip(hipMalloc(reinterpret_cast<void**>(&_data), bytes), "hipMalloc");
            detail::checkHip(hipMemset(_data, 0, bytes), "hipMemset");
        }
    }
    ~DenseMatrix() { release(); }
    DenseMatrix(const DenseMatrix&) = delete;
    DenseMatrix& operator=(const DenseMatrix&) = delete;
    DenseMatrix(DenseMatrix&& o) noexcept : _rows(o._rows), _cols(o._cols), _data(std::exchange(o._data, nullptr)) {}
    DenseMatrix& operator=(DenseMatrix&& o) noexcept
    {
        if (this != &o)
        {
            release();
            _rows = o._rows;
            _cols = o._cols;
            _data = std::exchange(o._data, nullptr);
        }
        return *this;
    }

    Index rows() const noexcept { return _rows; }
    Index cols() const noexcept { return _cols; }
    Index size() const noexcept { return _rows * _cols; }
    Scalar* data() noexcept { return _data; }
    const Scalar* data() const noexcept { return _data; }

    void fill(Scalar value)
    {
        std::vector<Scalar> host(static_cast<std::size_t>(size()), value);
        if (!host.empty())
            detail::checkHip(hipMemcpy(_data, host.data(), sizeof(Scalar) * host.size(), hipMemcpyHostToDevice),
                             "hipMemcpy");
    }

    DenseMatrix<Scalar, Device::CPU> toCpu() const
    {
        DenseMatrix<Scalar, Device::CPU> out(_rows, _cols);
        if (size() > 0)
            detail::checkHip(hipMemcpy(out.data(), _data, sizeof(Scalar) * static_cast<std::size_t>(size()),
                                       hipMemcpyDeviceToHost),
                             "hipMemcpy");
        return out;
    }

private:
    void release() noexcept
    {
        if (_data) (void)hipFree(_data);
        _data = nullptr;
    }

    Index _rows = 0;
    Index _cols = 0;
    Scalar* _data = nullptr;
};

template <typename Scalar>
DenseMatrix<Scalar, Device::GPU> DenseMatrix<Scalar, Device::CPU>::toGpu() const
{
    DenseMatrix<Scalar, Device::GPU> out(_rows, _cols);
    if (size() > 0)
        detail::checkHip(hipMemcpy(out.data(), _data.data(), sizeof(Scalar) * _data.size(), hipMemcpyHostToDevice),
                         "hipMemcpy");
    return out;
}

} // namespace plamatrix
