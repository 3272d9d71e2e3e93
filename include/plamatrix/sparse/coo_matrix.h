// COO triplet list (sparse/coo_matrix.h:13-74 of the reference).  Triplets
// are kept on the host in insertion order.  toCsr() of a GPU matrix converts
// them on the device (stable (row, col) sort and left fold of duplicates,
// bit-identical to the host conversion); a CPU matrix converts on the host.
#pragma once

#include <stdexcept>
#include <vector>

#include "plamatrix/dense/dense_matrix.h"
#include "plamatrix/sparse/csr_matrix.h"
#include "plamatrix/sparse/sparse_ops.h"

namespace plamatrix
{

template <typename Scalar, Device Dev>
class COOMatrix
{
public:
    COOMatrix(Index rows, Index cols) : _rows(rows), _cols(cols)
    {
        if (rows < 0 || cols < 0) throw std::invalid_argument("COOMatrix dimensions must be non-negative");
    }

    void add(Index row, Index col, Scalar value)
    {
        if (row < 0 || row >= _rows || col < 0 || col >= _cols)
            throw std::out_of_range("COOMatrix triplet index is out of bounds");
        _rowIndices.push_back(row);
        _colIndices.push_back(col);
        _values.push_back(value);
    }

    Index rows() const noexcept { return _rows; }
    Index cols() const noexcept { return _cols; }
    Index nnz() const noexcept { return static_cast<Index>(_values.size()); }

    CSRMatrix<Scalar, Dev> toCsr() const
    {
        if constexpr (Dev == Device::CPU)
        {
            return cooToCsr<Scalar>(_rows, _cols, _rowIndices, _colIndices, _values);
        }
        else
        {
            const Index n = nnz();
            DenseMatrix<Index, Device::CPU> r(n, 1), c(n, 1);
            DenseMatrix<Scalar, Device::CPU> v(n, 1);
            for (Index i = 0; i < n; ++i)
            {
                r.data()[i] = _rowIndices[static_cast<std::size_t>(i)];
                c.data()[i] = _colIndices[static_cast<std::size_t>(i)];
                v.data()[i] = _values[static_cast<std::size_t>(i)];
            }
            SparseOpsWorkspace ws;
            return cooToCsr<Scalar>(_rows, _cols, r.toGpu(), c.toGpu(), v.toGpu(), ws);
        }
    }

private:
    Index _rows = 0, _cols = 0;
    std::vector<Index> _rowIndices;
    std::vector<Index> _colIndices;
    std::vector<Scalar> _values;
};

} // namespace plamatrix
