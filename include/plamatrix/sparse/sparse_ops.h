// COO -> CSR and CSR SpMV on the GPU (sparse/sparse_ops.h:13-188 of the
// reference) over the C ABI in plamatrix_amd.h.
#pragma once

#include <stdexcept>
#include <utility>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "plamatrix/core/status.h"
#include "plamatrix/dense/dense_matrix.h"
#include "plamatrix/sparse/csr_matrix.h"
#include "plamatrix_amd.h"

namespace plamatrix
{

/// Reusable device scratch for sparse operations: move-only, grow-only, not
/// thread-safe, bound to the first stream it is used on (sparse_ops.h:13-77).
class SparseOpsWorkspace
{
public:
    SparseOpsWorkspace() { detail::check(pm_workspace_create(&_ws)); }
    ~SparseOpsWorkspace()
    {
        if (_ws) pm_workspace_destroy(_ws);
    }
    SparseOpsWorkspace(const SparseOpsWorkspace&) = delete;
    SparseOpsWorkspace& operator=(const SparseOpsWorkspace&) = delete;
    SparseOpsWorkspace(SparseOpsWorkspace&& o) noexcept : _ws(std::exchange(o._ws, nullptr)) {}
    SparseOpsWorkspace& operator=(SparseOpsWorkspace&& o) noexcept
    {
        if (this != &o)
        {
            if (_ws) pm_workspace_destroy(_ws);
            _ws = std::exchange(o._ws, nullptr);
        }
        return *this;
    }

    /// Release storage; the owning stream must have completed its work.
    void closeAsyncAllocation() { detail::check(pm_workspace_close(_ws)); }
    /// Deferred status of the last asynchronous COO conversion.
    void checkStatus(const char* = nullptr) { detail::check(pm_workspace_check_status(_ws)); }
    std::size_t capacityBytes() const noexcept { return pm_workspace_capacity_bytes(_ws); }
    pm_workspace* handle() const noexcept { return _ws; }

private:
    pm_workspace* _ws = nullptr;
};

/// Synchronously convert device COO triplets to sorted, duplicate-combined
/// GPU CSR (bit-identical to the CPU conversion).
template <typename Scalar>
CSRMatrix<Scalar, Device::GPU> cooToCsr(Index rows, Index cols, const DenseMatrix<Index, Device::GPU>& row_indices,
                                        const DenseMatrix<Index, Device::GPU>& col_indices,
                                        const DenseMatrix<Scalar, Device::GPU>& values, SparseOpsWorkspace& workspace)
{
    if (row_indices.size() != values.size() || col_indices.size() != values.size())
        throw std::invalid_argument("cooToCsr row, column, and value arrays must have equal lengths");
    hipStream_t stream = nullptr;
    detail::check(pm_coo_to_csr_prepare(rows, cols, values.size(), row_indices.data(), col_indices.data(),
                                        values.data(), detail::dtypeOf<Scalar>(), PM_COO_CHECK_FINITE,
                                        workspace.handle(), stream));
    int64_t nnz = 0;
    detail::check(pm_coo_to_csr_nnz(workspace.handle(), &nnz));
    CSRMatrix<Scalar, Device::GPU> out(rows, cols, nnz);
    detail::check(pm_coo_to_csr_finish(out.rowOffsets(), out.colIndices(), out.values(), nnz, workspace.handle(),
                                       stream));
    detail::checkHip(hipStreamSynchronize(stream), "hipStreamSynchronize");
    workspace.checkStatus("cooToCsr");
    return out;
}

/// Enqueue the conversion into preallocated output of exactly the combined
/// nnz; synchronise, then call workspace.checkStatus() before use.
template <typename Scalar>
void cooToCsrAsync(Index rows, Index cols, const DenseMatrix<Index, Device::GPU>& row_indices,
                   const DenseMatrix<Index, Device::GPU>& col_indices, const DenseMatrix<Scalar, Device::GPU>& values,
                   CSRMatrix<Scalar, Device::GPU>& output, SparseOpsWorkspace& workspace, hipStream_t stream = nullptr)
{
    if (row_indices.size() != values.size() || col_indices.size() != values.size())
        throw std::invalid_argument("cooToCsr row, column, and value arrays must have equal lengths");
    if (output.rows() != rows || output.cols() != cols)
        throw std::invalid_argument("cooToCsrAsync output dimensions do not match");
    detail::check(pm_coo_to_csr_async(rows, cols, values.size(), row_indices.data(), col_indices.data(),
                                      values.data(), detail::dtypeOf<Scalar>(), PM_COO_CHECK_FINITE, output.rowOffsets(),
                                      output.colIndices(), output.values(), output.nnz(), workspace.handle(), stream));
}

/// Enqueue y = A x on `stream` without host synchronisation.
template <typename Scalar>
void spmvAsync(const CSRMatrix<Scalar, Device::GPU>& csr, const DenseMatrix<Scalar, Device::GPU>& x,
               DenseMatrix<Scalar, Device::GPU>& output, SparseOpsWorkspace& workspace, hipStream_t stream = nullptr)
{
    // dimension mismatches are runtime_error, aliasing invalid_argument, as in
    // the reference (sparse_ops_cpu.cpp:138-156, sparse_ops_test.cpp:791-800)
    if (x.rows() != csr.cols() || x.cols() != 1)
        throw std::runtime_error("spmv GPU dimension mismatch");
    if (output.rows() != csr.rows() || output.cols() != 1)
        throw std::runtime_error("spmv GPU output dimension mismatch");
    const pm_csr d = csr.descriptor(true, stream);
    detail::check(pm_spmv(&d, x.data(), output.data(), workspace.handle(), stream));
}

/// Synchronous y = A x into existing storage.
template <typename Scalar>
void spmv(const CSRMatrix<Scalar, Device::GPU>& csr, const DenseMatrix<Scalar, Device::GPU>& x,
          DenseMatrix<Scalar, Device::GPU>& output)
{
    SparseOpsWorkspace ws;
    spmvAsync(csr, x, output, ws, nullptr);
    detail::checkHip(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
}

/// Synchronous y = A x.
template <typename Scalar>
DenseMatrix<Scalar, Device::GPU> spmv(const CSRMatrix<Scalar, Device::GPU>& csr,
                                      const DenseMatrix<Scalar, Device::GPU>& x)
{
    DenseMatrix<Scalar, Device::GPU> y(csr.rows(), 1);
    spmv(csr, x, y);
    return y;
}

/// Enqueue Y = A B (column-major B and Y) on `stream` without host
/// synchronisation; the matrix streams once for all columns.
template <typename Scalar>
void spmmAsync(const CSRMatrix<Scalar, Device::GPU>& csr, const DenseMatrix<Scalar, Device::GPU>& B,
               DenseMatrix<Scalar, Device::GPU>& output, SparseOpsWorkspace& workspace, hipStream_t stream = nullptr)
{
    if (B.rows() != csr.cols()) throw std::runtime_error("spmm dimension mismatch");
    if (output.rows() != csr.rows() || output.cols() != B.cols())
        throw std::runtime_error("spmm output dimension mismatch");
    const pm_csr d = csr.descriptor(true, stream);
    detail::check(pm_spmm(&d, B.data(), B.cols(), output.data(), workspace.handle(), stream));
}

/// Synchronous Y = A B into existing storage.
template <typename Scalar>
void spmm(const CSRMatrix<Scalar, Device::GPU>& csr, const DenseMatrix<Scalar, Device::GPU>& B,
          DenseMatrix<Scalar, Device::GPU>& output)
{
    SparseOpsWorkspace ws;
    spmmAsync(csr, B, output, ws, nullptr);
    detail::checkHip(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
}

/// Synchronous Y = A B.
template <typename Scalar>
DenseMatrix<Scalar, Device::GPU> spmm(const CSRMatrix<Scalar, Device::GPU>& csr,
                                      const DenseMatrix<Scalar, Device::GPU>& B)
{
    DenseMatrix<Scalar, Device::GPU> y(csr.rows(), B.cols());
    spmm(csr, B, y);
    return y;
}

} // namespace plamatrix
