// CSR matrices on the host and in HBM (sparse/csr_matrix.h:22-825 of the
// reference): int64 row offsets and column indices, Scalar values.  The GPU
// matrix also keeps 32-bit index mirrors whenever nnz and cols fit, which is
// what the gfx950 kernels stream.
//
// Structure trust (csr_matrix.h:181-237, 690-741): a matrix is trusted when
// its structure was validated (or produced by a conversion that guarantees
// it), no mutable array pointer has been handed out, and no asynchronous write
// into it is pending on another stream.  SpMV/SpMM and the fixed-iteration
// solvers require trust; the adaptive solvers revalidate untrusted input.
// Every asynchronous copy is ordered on one stream, and stream-ordered
// storage may only be used on the stream that owns it.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

#include <hip/hip_runtime_api.h>

#include "plamatrix/core/status.h"
#include "plamatrix/core/storage.h"
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

// Trust and pending-write bookkeeping shared by the host and device matrices.
class CsrTrust
{
public:
    /// Trusted structure and no outstanding mutable pointer or async write.
    bool hasValidatedStructure() const noexcept { return _validated && !_escaped && !_pending; }
    /// Trusted structure that may be consumed in order on `stream`.
    bool isStructureUsableOnStream(hipStream_t stream) const noexcept
    {
        return _validated && !_escaped && (!_pending || _pendingStream == stream);
    }

protected:
    explicit CsrTrust(bool validated = true) : _validated(validated) {}
    CsrTrust(CsrTrust&& o) noexcept
        : _validated(std::exchange(o._validated, true)), _escaped(std::exchange(o._escaped, false)),
          _pending(std::exchange(o._pending, false)), _pendingStream(std::exchange(o._pendingStream, nullptr))
    {
    }
    CsrTrust& operator=(CsrTrust&& o) noexcept
    {
        if (this != &o)
        {
            _validated = std::exchange(o._validated, true);
            _escaped = std::exchange(o._escaped, false);
            _pending = std::exchange(o._pending, false);
            _pendingStream = std::exchange(o._pendingStream, nullptr);
        }
        return *this;
    }

    void escape() noexcept
    {
        _validated = false;
        _escaped = true;
    }
    void checkPendingWriteStream(const char* op, hipStream_t stream) const
    {
        if (_pending && _pendingStream != stream)
            throw std::logic_error(std::string("CSRMatrix::") + op +
                                   " cannot overlap a pending async write from another stream");
    }
    void beginAsyncWrite(const char* op, hipStream_t stream) const
    {
        checkPendingWriteStream(op, stream);
        _validated = false;
        _pending = true;
        _pendingStream = stream;
    }
    void completeAsyncWrite(bool trusted) const noexcept
    {
        _validated = trusted;
        _pending = false;
        _pendingStream = nullptr;
    }
    void resetTrust() noexcept
    {
        _validated = true;
        _escaped = false;
        _pending = false;
        _pendingStream = nullptr;
    }

    mutable bool _validated = true;
    mutable bool _escaped = false;
    mutable bool _pending = false;
    mutable hipStream_t _pendingStream = nullptr;
};

inline void checkCsrShape(Index rows, Index cols, Index nnz)
{
    if (rows < 0 || cols < 0) throw std::invalid_argument("CSRMatrix dimensions must be non-negative");
    if (nnz < 0) throw std::invalid_argument("CSRMatrix nnz must be non-negative");
    if (nnz > 0 && (rows == 0 || cols == 0))
        throw std::invalid_argument("CSRMatrix cannot store non-zero entries with zero rows or columns");
}

inline void checkCsrTransfer(const char* op, Index rows, Index cols, Index nnz, Index want_rows, Index want_cols,
                             Index want_nnz)
{
    if (rows != want_rows || cols != want_cols || nnz != want_nnz)
        throw std::runtime_error(std::string("CSRMatrix::") + op + " output mismatch: output is " +
                                 std::to_string(rows) + "x" + std::to_string(cols) + " with nnz=" +
                                 std::to_string(nnz) + ", expected " + std::to_string(want_rows) + "x" +
                                 std::to_string(want_cols) + " with nnz=" + std::to_string(want_nnz));
}

struct CSRMatrixAccess;
} // namespace detail

template <typename Scalar>
class CSRMatrix<Scalar, Device::CPU> : public detail::CsrTrust
{
public:
    CSRMatrix() = default;
    CSRMatrix(Index rows, Index cols, Index nnz) : CSRMatrix(rows, cols, nnz, detail::HostKind::Pageable) {}

    /// Page-locked host storage: required by toGpuAsync/copyToGpuAsync and as
    /// the target of copyToCpuAsync.
    static CSRMatrix pinned(Index rows, Index cols, Index nnz)
    {
        return CSRMatrix(rows, cols, nnz, detail::HostKind::Pinned);
    }

    CSRMatrix(CSRMatrix&& o) noexcept
        : detail::CsrTrust(std::move(o)), _rows(std::exchange(o._rows, 0)), _cols(std::exchange(o._cols, 0)),
          _nnz(std::exchange(o._nnz, 0)), _rowOffsets(std::move(o._rowOffsets)), _colIndices(std::move(o._colIndices)),
          _values(std::move(o._values))
    {
    }
    CSRMatrix& operator=(CSRMatrix&& o) noexcept
    {
        if (this != &o)
        {
            detail::CsrTrust::operator=(std::move(o));
            _rows = std::exchange(o._rows, 0);
            _cols = std::exchange(o._cols, 0);
            _nnz = std::exchange(o._nnz, 0);
            _rowOffsets = std::move(o._rowOffsets);
            _colIndices = std::move(o._colIndices);
            _values = std::move(o._values);
        }
        return *this;
    }

    Index rows() const noexcept { return _rows; }
    Index cols() const noexcept { return _cols; }
    Index nnz() const noexcept { return _nnz; }
    /// Mutable access marks the structure untrusted; validateStructure() then
    /// checks it again on every call.
    Scalar* values() noexcept
    {
        escape();
        return _values.data();
    }
    const Scalar* values() const noexcept { return _values.data(); }
    Index* colIndices() noexcept
    {
        escape();
        return _colIndices.data();
    }
    const Index* colIndices() const noexcept { return _colIndices.data(); }
    Index* rowOffsets() noexcept
    {
        escape();
        return _rowOffsets.data();
    }
    const Index* rowOffsets() const noexcept { return _rowOffsets.data(); }
    bool isPinnedHost() const noexcept { return _rowOffsets.kind() == detail::HostKind::Pinned; }

    /// Validate offsets, columns and finite values (csr_matrix.h:197-237,
    /// 633-667 of the reference).  A pending async download must have been
    /// issued on `stream`; it is synchronised first.
    void validateStructure(hipStream_t stream = nullptr) const
    {
        if (_pending && _pendingStream != stream)
            throw std::logic_error("CSRMatrix::validateStructure must use the stream of the pending async copy");
        if (hasValidatedStructure()) return;
        if (_pending)
        {
            detail::checkHip(hipStreamSynchronize(stream), "hipStreamSynchronize");
            _pending = false;
            _pendingStream = nullptr;
        }
        validateHost();
        _validated = true;
    }

    CSRMatrix<Scalar, Device::GPU> toGpu() const;
    /// Upload into new stream-ordered GPU storage owned by `stream`.
    CSRMatrix<Scalar, Device::GPU> toGpuAsync(hipStream_t stream = nullptr) const;
    /// Upload into existing GPU storage of the same shape on `stream`.
    void copyToGpuAsync(CSRMatrix<Scalar, Device::GPU>& output, hipStream_t stream = nullptr) const;

private:
    friend class CSRMatrix<Scalar, Device::GPU>;
    friend struct detail::CSRMatrixAccess;

    CSRMatrix(Index rows, Index cols, Index nnz, detail::HostKind kind)
        : detail::CsrTrust(nnz == 0), _rows(rows), _cols(cols), _nnz(nnz)
    {
        detail::checkCsrShape(rows, cols, nnz);
        _rowOffsets = detail::HostArray<Index>(static_cast<std::size_t>(rows) + 1, kind, true);
        _colIndices = detail::HostArray<Index>(static_cast<std::size_t>(nnz), kind, true);
        _values = detail::HostArray<Scalar>(static_cast<std::size_t>(nnz), kind, false);
    }

    void validateHost() const
    {
        const Index* ro = _rowOffsets.data();
        const Index* ci = _colIndices.data();
        const Scalar* v = _values.data();
        if (!ro || ro[0] != 0 || ro[_rows] != _nnz)
            throw std::invalid_argument("CSRMatrix row offsets must start at zero and end at nnz");
        for (Index r = 0; r < _rows; ++r)
            if (ro[r] < 0 || ro[r + 1] < ro[r] || ro[r + 1] > _nnz)
                throw std::invalid_argument("CSRMatrix invalid row offsets at row " + std::to_string(r));
        for (Index k = 0; k < _nnz; ++k)
        {
            if (ci[k] < 0 || ci[k] >= _cols)
                throw std::invalid_argument("CSRMatrix column index out of range at position " + std::to_string(k));
            if (!std::isfinite(static_cast<double>(v[k])))
                throw std::invalid_argument("CSRMatrix value is not finite at position " + std::to_string(k));
        }
    }

    Index _rows = 0, _cols = 0, _nnz = 0;
    detail::HostArray<Index> _rowOffsets;
    detail::HostArray<Index> _colIndices;
    detail::HostArray<Scalar> _values;
};

template <typename Scalar>
class CSRMatrix<Scalar, Device::GPU> : public detail::CsrTrust
{
public:
    CSRMatrix() = default;
    /// Ordinary device storage with zeroed indices.
    CSRMatrix(Index rows, Index cols, Index nnz)
        : CSRMatrix(rows, cols, nnz, detail::GpuKind::Ordinary, nullptr)
    {
        if (_colIndices.data())
            detail::checkHip(hipMemset(_colIndices.data(), 0, sizeof(Index) * _colIndices.size()), "hipMemset");
        detail::checkHip(hipMemset(_rowOffsets.data(), 0, sizeof(Index) * _rowOffsets.size()), "hipMemset");
        detail::checkHip(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
    }

    /// Uninitialised stream-ordered storage owned by `stream`
    /// (csr_matrix.h:55-75 of the reference).  Call closeAsyncAllocation()
    /// before destroying a non-default stream.
    static CSRMatrix uninitializedAsync(Index rows, Index cols, Index nnz, hipStream_t stream = nullptr)
    {
        return CSRMatrix(rows, cols, nnz, detail::GpuKind::StreamOrdered, stream);
    }

    CSRMatrix(CSRMatrix&& o) noexcept
        : detail::CsrTrust(std::move(o)), _rows(std::exchange(o._rows, 0)), _cols(std::exchange(o._cols, 0)),
          _nnz(std::exchange(o._nnz, 0)), _rowOffsets(std::move(o._rowOffsets)), _colIndices(std::move(o._colIndices)),
          _values(std::move(o._values)), _rowOffsets32(std::move(o._rowOffsets32)),
          _colIndices32(std::move(o._colIndices32)), _mirrorsCurrent(std::exchange(o._mirrorsCurrent, false))
    {
    }
    /// Releases the current arrays (stream-ordered ones after a device-wide
    /// synchronisation) and takes over the source's.
    CSRMatrix& operator=(CSRMatrix&& o) noexcept
    {
        if (this != &o)
        {
            detail::CsrTrust::operator=(std::move(o));
            _rows = std::exchange(o._rows, 0);
            _cols = std::exchange(o._cols, 0);
            _nnz = std::exchange(o._nnz, 0);
            _rowOffsets = std::move(o._rowOffsets);
            _colIndices = std::move(o._colIndices);
            _values = std::move(o._values);
            _rowOffsets32 = std::move(o._rowOffsets32);
            _colIndices32 = std::move(o._colIndices32);
            _mirrorsCurrent = std::exchange(o._mirrorsCurrent, false);
        }
        return *this;
    }

    Index rows() const noexcept { return _rows; }
    Index cols() const noexcept { return _cols; }
    Index nnz() const noexcept { return _nnz; }
    /// Mutable access marks the structure untrusted for good: SpMV and the
    /// fixed-iteration solvers then reject the matrix, the adaptive solvers
    /// revalidate it on every call.
    Scalar* values() noexcept
    {
        touch();
        return _values.data();
    }
    const Scalar* values() const noexcept { return _values.data(); }
    Index* colIndices() noexcept
    {
        touch();
        return _colIndices.data();
    }
    const Index* colIndices() const noexcept { return _colIndices.data(); }
    Index* rowOffsets() noexcept
    {
        touch();
        return _rowOffsets.data();
    }
    const Index* rowOffsets() const noexcept { return _rowOffsets.data(); }

    bool isAsyncAllocation() const noexcept
    {
        return _rowOffsets.data() && _rowOffsets.kind() == detail::GpuKind::StreamOrdered;
    }
    hipStream_t asyncAllocationStream() const noexcept { return isAsyncAllocation() ? _rowOffsets.stream() : nullptr; }

    /// Enqueue the release of stream-ordered storage on its owning stream; the
    /// matrix becomes 0x0.  Repeated calls are harmless; ordinary storage is a
    /// std::logic_error (csr_matrix.h:252-296 of the reference).
    void closeAsyncAllocation()
    {
        if (!_rowOffsets.data() && !_colIndices.data() && !_values.data())
        {
            _rows = _cols = _nnz = 0;
            resetTrust();
            return;
        }
        if (!isAsyncAllocation())
            throw std::logic_error("CSRMatrix::closeAsyncAllocation requires stream-ordered storage");
        _values.releaseAsync();
        _colIndices.releaseAsync();
        _rowOffsets.releaseAsync();
        _colIndices32.releaseAsync();
        _rowOffsets32.releaseAsync();
        _rows = _cols = _nnz = 0;
        _mirrorsCurrent = false;
        resetTrust();
    }

    /// Validate the structure on the device (synchronises `stream`).  A
    /// pending async upload must have been issued on `stream`.
    void validateStructure(hipStream_t stream = nullptr) const
    {
        if (_pending && _pendingStream != stream)
            throw std::logic_error("CSRMatrix::validateStructure must use the stream of the pending async copy");
        if (hasValidatedStructure()) return;
        checkAsyncStream("validateStructure", stream);
        pm_workspace* ws = nullptr;
        detail::check(pm_workspace_create(&ws));
        const pm_csr d = descriptor(false);
        const int status = pm_csr_validate(&d, ws, stream);
        pm_workspace_destroy(ws);
        detail::check(status);
        _validated = true;
        _pending = false;
        _pendingStream = nullptr;
    }

    /// Synchronous download; rejected for stream-ordered storage.
    CSRMatrix<Scalar, Device::CPU> toCpu() const
    {
        if (isAsyncAllocation())
            throw std::logic_error("CSRMatrix::toCpu cannot synchronously read stream-ordered storage; "
                                   "use toCpuAsync on the owning stream");
        if (_pending)
        {
            detail::checkHip(hipStreamSynchronize(_pendingStream), "hipStreamSynchronize");
            _pending = false;
            _pendingStream = nullptr;
        }
        CSRMatrix<Scalar, Device::CPU> out(_rows, _cols, _nnz);
        copyArrays(out, hipMemcpyDeviceToHost, nullptr, false);
        out._validated = hasValidatedStructure();
        return out;
    }

    /// Download into new page-locked host storage on `stream`; synchronise
    /// before reading it.
    CSRMatrix<Scalar, Device::CPU> toCpuAsync(hipStream_t stream = nullptr) const
    {
        checkAsyncStream("toCpuAsync", stream);
        auto out = CSRMatrix<Scalar, Device::CPU>::pinned(_rows, _cols, _nnz);
        copyToCpuAsync(out, stream);
        return out;
    }

    /// Download into existing page-locked host storage on `stream`.
    void copyToCpuAsync(CSRMatrix<Scalar, Device::CPU>& output, hipStream_t stream = nullptr) const
    {
        detail::checkCsrTransfer("copyToCpuAsync", output.rows(), output.cols(), output.nnz(), _rows, _cols, _nnz);
        checkAsyncStream("copyToCpuAsync", stream);
        checkPendingWriteStream("copyToCpuAsync", stream);
        if (!output.isPinnedHost())
            throw std::invalid_argument("CSRMatrix::copyToCpuAsync requires pinned CPU output storage");
        output.beginAsyncWrite("copyToCpuAsync", stream);
        copyArrays(output, hipMemcpyDeviceToHost, stream, true);
    }

    /// C ABI descriptor.  With `compact` the 32-bit mirrors are refreshed (on
    /// `stream`) when the int64 arrays may have changed since the last refresh.
    pm_csr descriptor(bool compact = true, hipStream_t stream = nullptr) const
    {
        pm_csr d{};
        d.rows = _rows;
        d.cols = _cols;
        d.nnz = _nnz;
        d.row_offsets = _rowOffsets.data();
        d.col_indices = _colIndices.data();
        d.values = _values.data();
        d.dtype = detail::dtypeOf<Scalar>();
        if (compact && _rowOffsets32.data())
        {
            if (!_mirrorsCurrent)
            {
                detail::check(pm_csr_compact_indices(&d, _rowOffsets32.data(), _colIndices32.data(), stream));
                _mirrorsCurrent = true;
            }
            d.row_offsets32 = _rowOffsets32.data();
            d.col_indices32 = _colIndices32.data();
        }
        return d;
    }

    /// Stream-ordered storage must be used on the stream that owns it.
    void checkAsyncStream(const char* op, hipStream_t stream) const
    {
        if (isAsyncAllocation() && stream != _rowOffsets.stream())
            throw std::logic_error(std::string("CSRMatrix::") + op +
                                   " must use the stream that owns the async allocation");
    }

private:
    friend class CSRMatrix<Scalar, Device::CPU>;
    friend struct detail::CSRMatrixAccess;

    CSRMatrix(Index rows, Index cols, Index nnz, detail::GpuKind kind, hipStream_t stream)
        : detail::CsrTrust(nnz == 0), _rows(rows), _cols(cols), _nnz(nnz)
    {
        detail::checkCsrShape(rows, cols, nnz);
        _rowOffsets = detail::GpuArray<Index>(static_cast<std::size_t>(rows) + 1, kind, stream);
        _colIndices = detail::GpuArray<Index>(static_cast<std::size_t>(nnz), kind, stream);
        _values = detail::GpuArray<Scalar>(static_cast<std::size_t>(nnz), kind, stream);
        if (nnz <= INT32_MAX && cols <= INT32_MAX)
        {
            _rowOffsets32 = detail::GpuArray<int32_t>(static_cast<std::size_t>(rows) + 1, kind, stream);
            _colIndices32 = detail::GpuArray<int32_t>(static_cast<std::size_t>(nnz), kind, stream);
        }
    }

    void touch() noexcept
    {
        escape();
        _mirrorsCurrent = false;
    }

    // Copies all three arrays between this matrix and a host matrix.
    void copyArrays(const CSRMatrix<Scalar, Device::CPU>& host, hipMemcpyKind kind, hipStream_t stream,
                    bool async) const
    {
        const bool down = kind == hipMemcpyDeviceToHost;
        auto copy = [&](void* dst_dev, void* dst_host, std::size_t bytes) {
            if (bytes == 0) return;
            void* dst = down ? dst_host : dst_dev;
            const void* src = down ? dst_dev : dst_host;
            if (async)
                detail::checkHip(hipMemcpyAsync(dst, src, bytes, kind, stream), "hipMemcpyAsync");
            else
                detail::checkHip(hipMemcpy(dst, src, bytes, kind), "hipMemcpy");
        };
        const std::size_t nz = static_cast<std::size_t>(_nnz);
        copy(_values.data(), host._values.data(), detail::checkedBytes<Scalar>(nz));
        copy(_colIndices.data(), host._colIndices.data(), detail::checkedBytes<Index>(nz));
        copy(_rowOffsets.data(), host._rowOffsets.data(),
             detail::checkedBytes<Index>(static_cast<std::size_t>(_rows) + 1));
    }

    Index _rows = 0, _cols = 0, _nnz = 0;
    detail::GpuArray<Index> _rowOffsets;
    detail::GpuArray<Index> _colIndices;
    detail::GpuArray<Scalar> _values;
    detail::GpuArray<int32_t> _rowOffsets32;
    detail::GpuArray<int32_t> _colIndices32;
    mutable bool _mirrorsCurrent = false;
};

template <typename Scalar>
CSRMatrix<Scalar, Device::GPU> CSRMatrix<Scalar, Device::CPU>::toGpu() const
{
    validateStructure();
    CSRMatrix<Scalar, Device::GPU> out(_rows, _cols, _nnz, detail::GpuKind::Ordinary, nullptr);
    out.copyArrays(*this, hipMemcpyHostToDevice, nullptr, false);
    out.resetTrust();
    return out;
}

template <typename Scalar>
CSRMatrix<Scalar, Device::GPU> CSRMatrix<Scalar, Device::CPU>::toGpuAsync(hipStream_t stream) const
{
    if (!isPinnedHost()) throw std::invalid_argument("CSRMatrix::toGpuAsync requires pinned CPU input storage");
    validateStructure();
    auto out = CSRMatrix<Scalar, Device::GPU>::uninitializedAsync(_rows, _cols, _nnz, stream);
    copyToGpuAsync(out, stream);
    return out;
}

template <typename Scalar>
void CSRMatrix<Scalar, Device::CPU>::copyToGpuAsync(CSRMatrix<Scalar, Device::GPU>& output, hipStream_t stream) const
{
    detail::checkCsrTransfer("copyToGpuAsync", output.rows(), output.cols(), output.nnz(), _rows, _cols, _nnz);
    if (!isPinnedHost()) throw std::invalid_argument("CSRMatrix::copyToGpuAsync requires pinned CPU input storage");
    validateStructure();
    output.checkAsyncStream("copyToGpuAsync", stream);
    output.beginAsyncWrite("copyToGpuAsync", stream);
    output.copyArrays(*this, hipMemcpyHostToDevice, stream, true);
    output._mirrorsCurrent = false;
    output._validated = true;
}

namespace detail
{
// Library-internal writers: device conversions that fill a matrix without
// handing out mutable pointers, so the result can stay trusted.
struct CSRMatrixAccess
{
    template <typename Scalar>
    static Scalar* values(CSRMatrix<Scalar, Device::GPU>& m) noexcept
    {
        return m._values.data();
    }
    template <typename Scalar>
    static Index* colIndices(CSRMatrix<Scalar, Device::GPU>& m) noexcept
    {
        return m._colIndices.data();
    }
    template <typename Scalar>
    static Index* rowOffsets(CSRMatrix<Scalar, Device::GPU>& m) noexcept
    {
        return m._rowOffsets.data();
    }
    template <typename Scalar>
    static void beginAsyncWrite(CSRMatrix<Scalar, Device::GPU>& m, hipStream_t stream, const char* op)
    {
        m.beginAsyncWrite(op, stream);
        m._mirrorsCurrent = false;
    }
    template <typename Scalar>
    static void completeAsyncWrite(CSRMatrix<Scalar, Device::GPU>& m, bool trusted) noexcept
    {
        m.completeAsyncWrite(trusted);
    }
    /// A host matrix just written by a conversion that guarantees a valid
    /// structure (the host cooToCsr).
    template <typename Scalar>
    static void markTrusted(CSRMatrix<Scalar, Device::CPU>& m) noexcept
    {
        m.resetTrust();
    }
};

// Consumers' stream rules (sparse_ops.cu:74-97 of the reference).
template <typename M>
void checkStorageStream(const char* name, const M& m, hipStream_t stream)
{
    if (m.isAsyncAllocation() && m.asyncAllocationStream() != stream)
        throw std::logic_error(std::string(name) + " must be used on the stream that owns its async allocation");
}

template <typename Scalar>
void checkCsrStructureStream(const char* name, const CSRMatrix<Scalar, Device::GPU>& m, hipStream_t stream)
{
    if (!m.isStructureUsableOnStream(stream))
        throw std::logic_error(std::string(name) + " requires trusted CSR structure with no cross-stream pending write");
}
} // namespace detail

} // namespace plamatrix
