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
#include "plamatrix/sparse/host_sparse.h"
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
    SparseOpsWorkspace(SparseOpsWorkspace&& o) noexcept
        : _ws(std::exchange(o._ws, nullptr)), _statusOutput(std::exchange(o._statusOutput, nullptr)),
          _statusFinalize(std::exchange(o._statusFinalize, nullptr))
    {
    }
    SparseOpsWorkspace& operator=(SparseOpsWorkspace&& o) noexcept
    {
        if (this != &o)
        {
            if (_ws) pm_workspace_destroy(_ws);
            _ws = std::exchange(o._ws, nullptr);
            _statusOutput = std::exchange(o._statusOutput, nullptr);
            _statusFinalize = std::exchange(o._statusFinalize, nullptr);
        }
        return *this;
    }

    /// Release storage (a no-op after a move); the owning stream must have
    /// completed its work, else std::logic_error.
    void closeAsyncAllocation()
    {
        if (_ws) detail::check(pm_workspace_close(_ws));
    }

    /// Deferred status of the last asynchronous COO conversion: throws its
    /// error, if any, and marks the conversion's output trusted on success.
    /// Before the stream has finished it is a std::logic_error and the status
    /// stays pending.
    void checkStatus(const char* = nullptr)
    {
        if (!_ws) return;
        const int status = pm_workspace_check_status(_ws);
        if (status == PM_LOGIC_ERROR) detail::throwStatus(status);
        if (_statusOutput && _statusFinalize) _statusFinalize(_statusOutput, status == PM_OK);
        _statusOutput = nullptr;
        _statusFinalize = nullptr;
        detail::check(status);
    }
    std::size_t capacityBytes() const noexcept { return pm_workspace_capacity_bytes(_ws); }
    pm_workspace* handle() const noexcept { return _ws; }

    // The next checkStatus() settles the trust of `output`.
    template <typename Scalar>
    void beginStatus(CSRMatrix<Scalar, Device::GPU>* output) noexcept
    {
        _statusOutput = output;
        _statusFinalize = [](void* m, bool trusted) {
            detail::CSRMatrixAccess::completeAsyncWrite(*static_cast<CSRMatrix<Scalar, Device::GPU>*>(m), trusted);
        };
    }

private:
    pm_workspace* _ws = nullptr;
    void* _statusOutput = nullptr;
    void (*_statusFinalize)(void*, bool) = nullptr;
};

/// Convert host COO triplets to sorted, duplicate-combined CSR storage
/// (sparse_ops.h:79-85 of the reference; host matrices only).
template <typename Scalar>
CSRMatrix<Scalar, Device::CPU> cooToCsr(Index rows, Index cols, const std::vector<Index>& row_indices,
                                        const std::vector<Index>& col_indices, const std::vector<Scalar>& values)
{
    std::vector<std::size_t> order;
    std::vector<Index> row_start;
    detail::host::orderTriplets(rows, cols, row_indices, col_indices, values.size(), order, row_start);
    CSRMatrix<Scalar, Device::CPU> out;
    detail::host::foldInto(rows, col_indices, values, order, row_start,
                           [&](Index nnz, Index*& ro, Index*& ci, Scalar*& v) {
                               out = CSRMatrix<Scalar, Device::CPU>(rows, cols, nnz);
                               ro = out.rowOffsets();
                               ci = out.colIndices();
                               v = out.values();
                           });
    detail::CSRMatrixAccess::markTrusted(out);
    return out;
}

/// Convert host COO triplets into an existing host CSR matrix whose shape and
/// nnz must match the combined triplets (sparse_ops.h:87-96; runtime_error
/// otherwise, after the triplets themselves were checked).
template <typename Scalar>
void cooToCsr(Index rows, Index cols, const std::vector<Index>& row_indices, const std::vector<Index>& col_indices,
              const std::vector<Scalar>& values, CSRMatrix<Scalar, Device::CPU>& output)
{
    std::vector<std::size_t> order;
    std::vector<Index> row_start;
    detail::host::orderTriplets(rows, cols, row_indices, col_indices, values.size(), order, row_start);
    detail::host::foldInto(rows, col_indices, values, order, row_start,
                           [&](Index nnz, Index*& ro, Index*& ci, Scalar*& v) {
                               if (output.rows() != rows || output.cols() != cols)
                                   throw std::runtime_error("cooToCsr output dimension mismatch: output is " +
                                                            std::to_string(output.rows()) + "x" +
                                                            std::to_string(output.cols()) + ", expected " +
                                                            std::to_string(rows) + "x" + std::to_string(cols));
                               if (output.nnz() != nnz)
                                   throw std::runtime_error("cooToCsr output nnz mismatch: output has " +
                                                            std::to_string(output.nnz()) + ", expected " +
                                                            std::to_string(nnz));
                               ro = output.rowOffsets();
                               ci = output.colIndices();
                               v = output.values();
                           });
    detail::CSRMatrixAccess::markTrusted(output);
}

/// y = A x for host matrices (sparse_ops.h:122-131): each row left to right
/// in Scalar.
template <typename Scalar>
void spmv(const CSRMatrix<Scalar, Device::CPU>& csr, const DenseMatrix<Scalar, Device::CPU>& x,
          DenseMatrix<Scalar, Device::CPU>& output)
{
    if (x.rows() != csr.cols() || x.cols() != 1)
        throw std::runtime_error("spmv dimension mismatch: CSR is " + std::to_string(csr.rows()) + "x" +
                                 std::to_string(csr.cols()) + ", x is " + std::to_string(x.rows()) + "x" +
                                 std::to_string(x.cols()) + ", expected " + std::to_string(csr.cols()) + "x1");
    if (output.rows() != csr.rows() || output.cols() != 1)
        throw std::runtime_error("spmv output dimension mismatch: output is " + std::to_string(output.rows()) + "x" +
                                 std::to_string(output.cols()) + ", expected " + std::to_string(csr.rows()) + "x1");
    if (x.data() != nullptr && x.data() == output.data())
        throw std::invalid_argument("spmv input and output data must not alias");
    detail::host::spmvRows(csr.rows(), csr.rowOffsets(), csr.colIndices(), csr.values(), x.data(), output.data());
}

template <typename Scalar>
DenseMatrix<Scalar, Device::CPU> spmv(const CSRMatrix<Scalar, Device::CPU>& csr,
                                      const DenseMatrix<Scalar, Device::CPU>& x)
{
    DenseMatrix<Scalar, Device::CPU> y(csr.rows(), 1);
    spmv(csr, x, y);
    return y;
}

/// Y = A B for host matrices, B and Y column-major (sparse_ops.h:133-146).
template <typename Scalar>
void spmm(const CSRMatrix<Scalar, Device::CPU>& csr, const DenseMatrix<Scalar, Device::CPU>& B,
          DenseMatrix<Scalar, Device::CPU>& output)
{
    if (B.rows() != csr.cols())
        throw std::runtime_error("spmm dimension mismatch: CSR is " + std::to_string(csr.rows()) + "x" +
                                 std::to_string(csr.cols()) + ", B is " + std::to_string(B.rows()) + "x" +
                                 std::to_string(B.cols()));
    if (output.rows() != csr.rows() || output.cols() != B.cols())
        throw std::runtime_error("spmm output dimension mismatch: output is " + std::to_string(output.rows()) + "x" +
                                 std::to_string(output.cols()) + ", expected " + std::to_string(csr.rows()) + "x" +
                                 std::to_string(B.cols()));
    if (B.data() != nullptr && B.data() == output.data())
        throw std::invalid_argument("spmm input and output data must not alias");
    for (Index j = 0; j < B.cols(); ++j)
        detail::host::spmvRows(csr.rows(), csr.rowOffsets(), csr.colIndices(), csr.values(), B.data() + j * B.rows(),
                               output.data() + j * output.rows());
}

template <typename Scalar>
DenseMatrix<Scalar, Device::CPU> spmm(const CSRMatrix<Scalar, Device::CPU>& csr,
                                      const DenseMatrix<Scalar, Device::CPU>& B)
{
    DenseMatrix<Scalar, Device::CPU> y(csr.rows(), B.cols());
    spmm(csr, B, y);
    return y;
}

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
    using Access = detail::CSRMatrixAccess;
    Access::beginAsyncWrite(out, stream, "cooToCsr");
    detail::check(pm_coo_to_csr_finish(Access::rowOffsets(out), Access::colIndices(out), Access::values(out), nnz,
                                       workspace.handle(), stream));
    detail::checkHip(hipStreamSynchronize(stream), "hipStreamSynchronize");
    workspace.checkStatus("cooToCsr");
    Access::completeAsyncWrite(out, true);
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
    detail::checkStorageStream("cooToCsrAsync rows", row_indices, stream);
    detail::checkStorageStream("cooToCsrAsync columns", col_indices, stream);
    detail::checkStorageStream("cooToCsrAsync values", values, stream);
    detail::checkStorageStream("cooToCsrAsync output", output, stream);
    if (output.rows() != rows || output.cols() != cols)
        throw std::runtime_error("cooToCsrAsync output dimensions do not match");
    using Access = detail::CSRMatrixAccess;
    if (values.data() != nullptr && values.data() == Access::values(output))
        throw std::invalid_argument("cooToCsrAsync values and output must not alias");
    Access::beginAsyncWrite(output, stream, "cooToCsrAsync");
    detail::check(pm_coo_to_csr_async(rows, cols, values.size(), row_indices.data(), col_indices.data(),
                                      values.data(), detail::dtypeOf<Scalar>(), PM_COO_CHECK_FINITE,
                                      Access::rowOffsets(output), Access::colIndices(output), Access::values(output),
                                      output.nnz(), workspace.handle(), stream));
    workspace.beginStatus(&output);
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
    if (x.data() != nullptr && x.data() == output.data())
        throw std::invalid_argument("spmv GPU input and output data must not alias");
    detail::checkStorageStream("spmvAsync CSR", csr, stream);
    detail::checkCsrStructureStream("spmvAsync CSR", csr, stream);
    detail::checkStorageStream("spmvAsync input", x, stream);
    detail::checkStorageStream("spmvAsync output", output, stream);
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
    if (B.data() != nullptr && B.data() == output.data())
        throw std::invalid_argument("spmm GPU input and output data must not alias");
    detail::checkStorageStream("spmmAsync CSR", csr, stream);
    detail::checkCsrStructureStream("spmmAsync CSR", csr, stream);
    detail::checkStorageStream("spmmAsync input", B, stream);
    detail::checkStorageStream("spmmAsync output", output, stream);
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
