"""ctypes binding of the C ABI in ``include/plamatrix_amd.h``.

Device memory comes from torch tensors (``tensor.data_ptr()``) and work is
enqueued on torch's current HIP stream.  Every call goes through
``libplamatrix_amd.so``; if the library is missing, ``lib()`` raises instead of
falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "libplamatrix_amd.so")

PM_F32, PM_F64 = 0, 1
PM_COO_CHECK_FINITE = 1

_i32, _i64, _f64, _p = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p

KINDS = {1: "invalid_argument", 2: "out_of_range", 3: "overflow_error", 4: "logic_error",
         5: "runtime_error", 6: "device_error"}


class PlamatrixError(Exception):
    def __init__(self, code: int, message: str):
        self.code = code
        self.kind = KINDS.get(code, "error")
        super().__init__(f"{self.kind}: {message}")


class pm_csr(ctypes.Structure):
    _fields_ = [("rows", _i64), ("cols", _i64), ("nnz", _i64), ("row_offsets", _p), ("col_indices", _p),
                ("values", _p), ("dtype", _i32), ("row_offsets32", _p), ("col_indices32", _p)]


class pm_solver_options(ctypes.Structure):
    _fields_ = [("max_iterations", _i32), ("relative_tolerance", _f64), ("absolute_tolerance", _f64),
                ("use_jacobi", _i32), ("require_convergence", _i32)]


class pm_solver_report(ctypes.Structure):
    _fields_ = [("converged", _i32), ("iterations", _i32), ("initial_residual", _f64), ("final_residual", _f64)]


class pm_solve_policy(ctypes.Structure):
    _fields_ = [("chunk_iterations", _i32), ("use_graph", _i32), ("refresh_interval", _i32),
                ("lanes_per_row", _i32)]


class pm_refine_report(ctypes.Structure):
    _fields_ = [("converged", _i32), ("outer_iterations", _i32), ("inner_iterations", _i32),
                ("initial_relres", _f64), ("final_relres", _f64)]


class pm_batch_report(ctypes.Structure):
    _fields_ = [("converged", _i32), ("iterations", _i32), ("initial_residual", _f64), ("final_residual", _f64),
                ("status", _i32), ("pad", _i32), ("bad_row", _i64)]


class pm_batch_refine_report(ctypes.Structure):
    _fields_ = [("converged", _i32), ("outer_iterations", _i32), ("inner_iterations", _i64),
                ("initial_relres", _f64), ("final_relres", _f64), ("last_inner_status", _i32), ("pad", _i32)]


class pm_normal_equations(ctypes.Structure):
    _fields_ = [("primary_count", _i64), ("eliminated_count", _i64), ("primary_size", _i64),
                ("eliminated_size", _i64), ("primary_diagonal", _p), ("eliminated_diagonal", _p),
                ("primary_gradient", _p), ("eliminated_gradient", _p), ("cross_count", _i64),
                ("cross_primary", _p), ("cross_eliminated", _p), ("cross_values", _p),
                ("primary_cross_count", _i64), ("primary_cross_rows", _p), ("primary_cross_cols", _p),
                ("primary_cross_values", _p)]


class pm_residual_batch(ctypes.Structure):
    _fields_ = [("count", _i64), ("residual_size", _i64), ("primary", _p), ("eliminated", _p),
                ("primary_jacobians", _p), ("eliminated_jacobians", _p), ("residuals", _p), ("weights", _p),
                ("huber_delta", _f64)]


PM_SCHUR_BLOCK_PCG = 0
PM_SCHUR_DENSE_CHOLESKY = 1


class pm_schur_options(ctypes.Structure):
    _fields_ = [("max_iterations", _i32), ("use_initial_guess", _i32), ("relative_tolerance", _f64),
                ("absolute_tolerance", _f64), ("minimum_diagonal", _f64), ("linear_backend", _i32), ("flags", _i32)]


class pm_dense_report(ctypes.Structure):
    _fields_ = [("not_positive_definite", _i32), ("pad", _i32), ("rhs_norm", _f64), ("residual_norm", _f64)]


class pm_schur_report(ctypes.Structure):
    _fields_ = [("converged", _i32), ("iterations", _i32), ("initial_residual_norm", _f64),
                ("final_residual_norm", _f64), ("status", _i32), ("pattern_reused", _i32),
                ("assembly_on_device", _i32), ("mixed_precision_used", _i32), ("pattern_build_count", _i64),
                ("message", ctypes.c_char * 96)]


# Every symbol declared in include/plamatrix_amd.h (checked by the CPU tests).
EXPORTS = [
    "pm_last_error", "pm_version", "pm_workspace_create", "pm_workspace_destroy", "pm_workspace_close",
    "pm_workspace_capacity_bytes", "pm_workspace_capacity_rows", "pm_workspace_solver_diagonals", "pm_workspace_solver_value_streams", "pm_workspace_solver_form", "pm_workspace_solver_preconditioner", "pm_csr_validate", "pm_spmm", "pm_csr_compact_indices",
    "pm_coo_to_csr_prepare", "pm_coo_to_csr_nnz", "pm_coo_to_csr_finish", "pm_coo_to_csr_async",
    "pm_workspace_check_status", "pm_spmv", "pm_spmv_lanes", "pm_autotune_spmv", "pm_tuned_lanes", "pm_tune_record", "pm_true_relres", "pm_cg", "pm_pcg",
    "pm_block_pcg", "pm_solve_fixed_async", "pm_fixed_finalize", "pm_fixed_state_destroy", "pm_pcg_refine",
    "pm_schur_workspace_create", "pm_schur_workspace_destroy", "pm_schur_pattern_build_count",
    "pm_schur_pattern_size", "pm_normal_equations_assemble", "pm_schur_solve", "pm_schur_assemble", "pm_schur_system",
    "pm_batch_layout", "pm_batch_pcg", "pm_batch_pcg_refine", "pm_pcg_factored", "pm_pcg_refine_factored",
    "pm_fsai_parameter_count", "pm_fsai_initial_parameters", "pm_fsai_net_create", "pm_fsai_net_destroy",
    "pm_fsai_net_set_parameters", "pm_fsai_net_get_parameters", "pm_fsai_graph_create", "pm_fsai_graph_destroy",
    "pm_fsai_graph_pool_release", "pm_fsai_graph_pool_bytes",
    "pm_fsai_forward", "pm_fsai_factors", "pm_fsai_loss_grad", "pm_fsai_train_step", "pm_dense_cholesky_solve",
    "pm_batch_pcg_factored", "pm_batch_pcg_refine_factored", "pm_fsai_batch_select",
    "pm_batch_block_pcg", "pm_batch_block_pcg_refine", "pm_normal_equations_add_primary",
    "pm_spmm_noalloc",
]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY):
            raise RuntimeError(f"{LIBRARY} is not built; run __graft_entry__.build() (HIP library required)")
        # torch ships its own HIP runtime under the same soname: load it first so
        # the library binds to that one runtime (loaded the other way round,
        # torch later finds no devices)
        import torch  # noqa: F401

        _lib = ctypes.CDLL(LIBRARY)
        _lib.pm_last_error.restype = ctypes.c_char_p
        _lib.pm_workspace_capacity_bytes.restype = ctypes.c_size_t
        _lib.pm_workspace_capacity_bytes.argtypes = [_p]
        _lib.pm_workspace_capacity_rows.restype = _i64
        _lib.pm_workspace_capacity_rows.argtypes = [_p]
        _lib.pm_workspace_solver_diagonals.restype = _i32
        _lib.pm_workspace_solver_diagonals.argtypes = [_p]
        _lib.pm_workspace_solver_value_streams.restype = _i32
        _lib.pm_workspace_solver_value_streams.argtypes = [_p]
        _lib.pm_workspace_solver_form.restype = _i32
        _lib.pm_workspace_solver_form.argtypes = [_p]
        _lib.pm_workspace_solver_preconditioner.restype = _i32
        _lib.pm_workspace_solver_preconditioner.argtypes = [_p]
        _lib.pm_workspace_destroy.argtypes = [_p]
        _lib.pm_workspace_close.argtypes = [_p]
        _lib.pm_workspace_check_status.argtypes = [_p]
        _lib.pm_fixed_state_destroy.argtypes = [_p]
        _lib.pm_schur_workspace_destroy.argtypes = [_p]
        _lib.pm_schur_pattern_build_count.argtypes = [_p]
        _lib.pm_schur_pattern_build_count.restype = _i64
        _lib.pm_fsai_parameter_count.restype = _i32
        _lib.pm_fsai_net_destroy.argtypes = [_p]
        _lib.pm_fsai_graph_destroy.argtypes = [_p]
    return _lib


def check(code: int):
    if code != 0:
        raise PlamatrixError(code, lib().pm_last_error().decode())


def current_stream():
    import torch

    return _p(torch.cuda.current_stream().cuda_stream)


class Workspace:
    """Owns a pm_workspace (device scratch, solver state, cached launch plans)."""

    def __init__(self):
        h = _p()
        check(lib().pm_workspace_create(ctypes.byref(h)))
        self.handle = h

    def close(self):
        check(lib().pm_workspace_close(self.handle))

    def check_status(self):
        check(lib().pm_workspace_check_status(self.handle))

    @property
    def capacity_bytes(self) -> int:
        return int(lib().pm_workspace_capacity_bytes(self.handle))

    @property
    def capacity_rows(self) -> int:
        return int(lib().pm_workspace_capacity_rows(self.handle))

    @property
    def solver_diagonals(self) -> int:
        """Diagonals of the layout the last solve streamed A in (0: CSR)."""
        return int(lib().pm_workspace_solver_diagonals(self.handle))

    @property
    def solver_value_streams(self) -> int:
        """Value arrays the last solve's diagonal-layout row sums read per row
        (fewer than the diagonals when A is mirrored; 0: CSR)."""
        return int(lib().pm_workspace_solver_value_streams(self.handle))

    @property
    def solver_form(self) -> int:
        """Iteration form the last solve started in: 3 classic, 2 / 1 single-reduction
        CG or Jacobi-PCG in two kernels / one, 4 / 5 single-reduction factored PCG in
        three kernels / two."""
        return int(lib().pm_workspace_solver_form(self.handle))

    @property
    def solver_preconditioner(self) -> str:
        """Preconditioner the last adaptive solve ran: "none", "jacobi",
        "block" or "factored" (a unit-diagonal Jacobi solve runs as "none")."""
        return ("none", "jacobi", "block", "factored")[int(lib().pm_workspace_solver_preconditioner(self.handle))]

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and _lib is not None:
            _lib.pm_workspace_destroy(h)
            self.handle = None


def dtype_code(t) -> int:
    import torch

    if t.dtype == torch.float32:
        return PM_F32
    if t.dtype == torch.float64:
        return PM_F64
    raise TypeError(f"unsupported dtype {t.dtype}")


@dataclass
class DeviceCSR:
    """CSR arrays on the device (torch tensors) plus optional int32 mirrors."""

    rows: int
    cols: int
    row_offsets: object
    col_indices: object
    values: object
    row_offsets32: object = None
    col_indices32: object = None

    @property
    def nnz(self) -> int:
        return int(self.values.numel())

    def descriptor(self) -> pm_csr:
        def ptr(t):
            return None if t is None else t.data_ptr()

        return pm_csr(self.rows, self.cols, self.nnz, ptr(self.row_offsets), ptr(self.col_indices),
                      ptr(self.values), dtype_code(self.values), ptr(self.row_offsets32), ptr(self.col_indices32))

    def compact(self):
        """Build the int32 index mirrors the kernels stream when present."""
        import torch

        ro = torch.empty(self.rows + 1, dtype=torch.int32, device=self.values.device)
        ci = torch.empty(max(self.nnz, 1), dtype=torch.int32, device=self.values.device)
        d = self.descriptor()
        check(lib().pm_csr_compact_indices(ctypes.byref(d), _p(ro.data_ptr()), _p(ci.data_ptr()), current_stream()))
        self.row_offsets32, self.col_indices32 = ro, ci
        return self

    def without_compact(self) -> "DeviceCSR":
        return DeviceCSR(self.rows, self.cols, self.row_offsets, self.col_indices, self.values)


def to_device_csr(host, device="cuda", dtype=None, compact=False) -> DeviceCSR:
    import numpy as np
    import torch

    vals = host.values if dtype is None else host.values.astype(dtype)
    csr = DeviceCSR(host.rows, host.cols,
                    torch.from_numpy(np.ascontiguousarray(host.row_offsets, np.int64)).to(device),
                    torch.from_numpy(np.ascontiguousarray(host.col_indices, np.int64)).to(device),
                    torch.from_numpy(np.ascontiguousarray(vals)).to(device))
    return csr.compact() if compact else csr


def options(max_iterations=1000, relative_tolerance=1e-6, absolute_tolerance=0.0, use_jacobi=True,
            require_convergence=False) -> pm_solver_options:
    return pm_solver_options(int(max_iterations), float(relative_tolerance), float(absolute_tolerance),
                             int(bool(use_jacobi)), int(bool(require_convergence)))


def policy(chunk_iterations=0, use_graph=-1, refresh_interval=0, lanes_per_row=0) -> pm_solve_policy:
    return pm_solve_policy(int(chunk_iterations), int(use_graph), int(refresh_interval), int(lanes_per_row))


def _ptr(t):
    return _p(t.data_ptr()) if t is not None else None


def spmv(a: DeviceCSR, x, y, ws: Workspace | None = None, lanes: int | None = None):
    """y = A x.  lanes=None uses the tuned variant; 0 forces the wave-stream
    kernel, 1..64 the lanes-per-row kernel."""
    d = a.descriptor()
    h = ws.handle if ws else None
    if lanes is not None:
        check(lib().pm_spmv_lanes(ctypes.byref(d), _ptr(x), _ptr(y), _i32(lanes), h, current_stream()))
    else:
        check(lib().pm_spmv(ctypes.byref(d), _ptr(x), _ptr(y), h, current_stream()))
    return y


def spmm(a: DeviceCSR, b, y, ws: Workspace | None = None):
    """Y = A B (pm_spmm) with B (cols x k) and Y (rows x k) column-major:
    torch tensors of shape (k, cols) and (k, rows), one column per row."""
    k = int(b.shape[0]) if b.dim() == 2 else 1
    d = a.descriptor()
    ws = ws if ws is not None else Workspace()
    check(lib().pm_spmm(ctypes.byref(d), _ptr(b), _i64(k), _ptr(y), ws.handle, current_stream()))


def validate(a: DeviceCSR, ws: Workspace):
    d = a.descriptor()
    check(lib().pm_csr_validate(ctypes.byref(d), ws.handle, current_stream()))


def true_relres(a: DeviceCSR, b, x, ws: Workspace) -> float:
    d = a.descriptor()
    out = _f64()
    check(lib().pm_true_relres(ctypes.byref(d), _ptr(b), _ptr(x), _i32(dtype_code(x)), ctypes.byref(out), ws.handle,
                               current_stream()))
    return out.value


@dataclass
class DenseReport:
    not_positive_definite: bool
    rhs_norm: float
    residual_norm: float


def dense_cholesky_solve(a, b, x, ws: Workspace) -> DenseReport:
    """Dense FP64 Cholesky solve of an SPD CSR system (lower triangle used).
    a: DeviceCSR or pm_csr descriptor; b, x: tensors or device addresses."""
    d = a if isinstance(a, pm_csr) else a.descriptor()

    def addr(v):
        return _p(v) if isinstance(v, int) else _ptr(v)

    r = pm_dense_report()
    check(lib().pm_dense_cholesky_solve(ctypes.byref(d), addr(b), addr(x), ctypes.byref(r), ws.handle,
                                        current_stream()))
    return DenseReport(bool(r.not_positive_definite), r.rhs_norm, r.residual_norm)


@dataclass
class Report:
    converged: bool
    iterations: int
    initial_residual: float
    final_residual: float


def _report(r: pm_solver_report) -> Report:
    return Report(bool(r.converged), r.iterations, r.initial_residual, r.final_residual)


def _system_vectors(a, *vs):
    """The solve entry points read and write vectors of A's dtype and row
    count through raw pointers: refuse any other before the call."""
    for v in vs:
        if v is None:
            continue
        if v.dtype != a.values.dtype or v.numel() < a.rows or not v.is_contiguous():
            raise ValueError(f"solve vector must be a contiguous {a.values.dtype} vector of {a.rows} entries, "
                             f"got {v.dtype} with shape {tuple(v.shape)}")


def _solve(fn, a, b, x, ws, pol, extra=(), **opts):
    _system_vectors(a, b, x)
    d = a.descriptor()
    o = options(**opts)
    rep = pm_solver_report()
    p = pol if pol is not None else policy()
    check(fn(ctypes.byref(d), _ptr(b), _ptr(x), *extra, ctypes.byref(o), ctypes.byref(p), ctypes.byref(rep),
             ws.handle, current_stream()))
    return _report(rep)


def cg(a, b, x, ws, pol=None, **opts) -> Report:
    return _solve(lib().pm_cg, a, b, x, ws, pol, **opts)


def pcg(a, b, x, ws, pol=None, **opts) -> Report:
    return _solve(lib().pm_pcg, a, b, x, ws, pol, **opts)


def block_pcg(a, b, x, inverse_blocks, block_size, ws, pol=None, **opts) -> Report:
    return _solve(lib().pm_block_pcg, a, b, x, ws, pol, (_ptr(inverse_blocks), _i64(block_size)), **opts)


class FixedState:
    def __init__(self, handle):
        self.handle = handle

    def finalize(self, **opts) -> Report:
        o = options(**opts)
        rep = pm_solver_report()
        check(lib().pm_fixed_finalize(self.handle, ctypes.byref(o), ctypes.byref(rep)))
        return _report(rep)

    def __del__(self):
        if getattr(self, "handle", None) is not None and _lib is not None:
            _lib.pm_fixed_state_destroy(self.handle)
            self.handle = None


def solve_fixed_async(a, b, x, iterations, preconditioned, ws) -> FixedState:
    d = a.descriptor()
    h = _p()
    check(lib().pm_solve_fixed_async(ctypes.byref(d), _ptr(b), _ptr(x), _i32(iterations), _i32(int(preconditioned)),
                                     ctypes.byref(h), ws.handle, current_stream()))
    return FixedState(h)


@dataclass
class RefineReport:
    converged: bool
    outer_iterations: int
    inner_iterations: int
    initial_relres: float
    final_relres: float


def pcg_refine(a, b64, x64, ws, target=1e-8, max_outer=10, pol=None, **inner) -> RefineReport:
    d = a.descriptor()
    o = options(**inner)
    p = pol if pol is not None else policy()
    rep = pm_refine_report()
    check(lib().pm_pcg_refine(ctypes.byref(d), _ptr(b64), _ptr(x64), ctypes.byref(o), _f64(target), _i32(max_outer),
                              ctypes.byref(p), ctypes.byref(rep), ws.handle, current_stream()))
    return RefineReport(bool(rep.converged), rep.outer_iterations, rep.inner_iterations, rep.initial_relres,
                        rep.final_relres)


def coo_to_csr(rows, cols, r, c, v, ws: Workspace, check_finite=True) -> DeviceCSR:
    """Synchronous device conversion (prepare, read nnz, allocate, finish)."""
    import torch

    flags = PM_COO_CHECK_FINITE if check_finite else 0
    stream = current_stream()
    count = int(v.numel())
    check(lib().pm_coo_to_csr_prepare(_i64(rows), _i64(cols), _i64(count), _ptr(r), _ptr(c), _ptr(v),
                                      _i32(dtype_code(v)), _i32(flags), ws.handle, stream))
    nnz = _i64()
    check(lib().pm_coo_to_csr_nnz(ws.handle, ctypes.byref(nnz)))
    dev = v.device
    ro = torch.empty(rows + 1, dtype=torch.int64, device=dev)
    ci = torch.empty(max(nnz.value, 1), dtype=torch.int64, device=dev)
    vals = torch.empty(max(nnz.value, 1), dtype=v.dtype, device=dev)
    check(lib().pm_coo_to_csr_finish(_ptr(ro), _ptr(ci), _ptr(vals), _i64(nnz.value), ws.handle, stream))
    torch.cuda.current_stream().synchronize()
    ws.check_status()
    return DeviceCSR(rows, cols, ro, ci[:nnz.value], vals[:nnz.value])


def coo_to_csr_async(rows, cols, r, c, v, out: DeviceCSR, ws: Workspace, check_finite=True):
    flags = PM_COO_CHECK_FINITE if check_finite else 0
    check(lib().pm_coo_to_csr_async(_i64(rows), _i64(cols), _i64(int(v.numel())), _ptr(r), _ptr(c), _ptr(v),
                                    _i32(dtype_code(v)), _i32(flags), _ptr(out.row_offsets), _ptr(out.col_indices),
                                    _ptr(out.values), _i64(out.nnz), ws.handle, current_stream()))


def autotune_spmv(a: DeviceCSR, x, y, repeats=10):
    d = a.descriptor()
    best = _i32()
    ms = ctypes.c_float()
    check(lib().pm_autotune_spmv(ctypes.byref(d), _ptr(x), _ptr(y), _i32(repeats), ctypes.byref(best),
                                 ctypes.byref(ms), current_stream()))
    return best.value, ms.value


# ---- damped Schur complement ----------------------------------------------------


class SchurWorkspace:
    """Owns a pm_schur_workspace (cached device pattern and numerics)."""

    def __init__(self):
        h = _p()
        check(lib().pm_schur_workspace_create(ctypes.byref(h)))
        self.handle = h

    @property
    def pattern_build_count(self) -> int:
        return int(lib().pm_schur_pattern_build_count(self.handle))

    def pattern_size(self):
        s, v, t = _i64(), _i64(), _i64()
        check(lib().pm_schur_pattern_size(self.handle, ctypes.byref(s), ctypes.byref(v), ctypes.byref(t)))
        return s.value, v.value, t.value

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and _lib is not None:
            _lib.pm_schur_workspace_destroy(h)
            self.handle = None


@dataclass
class NormalEquations:
    """Device normal equations: a pm_normal_equations descriptor plus the
    tensors that own (or alias) its arrays."""
    desc: pm_normal_equations
    keep: tuple = ()

    def cross_count(self) -> int:
        return int(self.desc.cross_count)


def assemble_normal_equations(sw: SchurWorkspace, primary_count, eliminated_count, primary_size, eliminated_size,
                              primary, eliminated, primary_jacobians, eliminated_jacobians, residuals,
                              weights=None, huber_delta=1.0) -> NormalEquations:
    """Device BlockNormalEquations::addResidualBlocks over a batch; arrays are
    device tensors (int64 ids, float64 values)."""
    rs = residuals.shape[1] if residuals.dim() == 2 else 1
    b = pm_residual_batch(int(primary.numel()), rs, _ptr(primary), _ptr(eliminated), _ptr(primary_jacobians),
                          _ptr(eliminated_jacobians), _ptr(residuals),
                          _ptr(weights) if weights is not None else None, float(huber_delta))
    out = pm_normal_equations()
    check(lib().pm_normal_equations_assemble(sw.handle, _i64(primary_count), _i64(eliminated_count),
                                             _i64(primary_size), _i64(eliminated_size), ctypes.byref(b),
                                             ctypes.byref(out), current_stream()))
    return NormalEquations(out, (primary, eliminated, primary_jacobians, eliminated_jacobians, residuals, weights))


class pm_primary_residual_batch(ctypes.Structure):
    _fields_ = [("count", _i64), ("residual_size", _i64), ("blocks_per_residual", _i64), ("primary", _p),
                ("jacobians", _p), ("residuals", _p), ("weights", _p), ("huber_delta", _f64)]


def add_primary_terms(sw: SchurWorkspace, ne: NormalEquations, primary, jacobians, residuals, weights=None,
                      huber_delta=1.0) -> NormalEquations:
    """Device addPrimaryTerms over a batch: primary (count x k int64 ids),
    jacobians (count x k x rs x ps float64), residuals (count x rs float64)
    device tensors; returns new equations (the input is unchanged)."""
    count, k = int(primary.shape[0]), int(primary.shape[1]) if primary.dim() == 2 else 1
    rs = int(residuals.shape[1]) if residuals.dim() == 2 else 1
    b = pm_primary_residual_batch(count, rs, k, _ptr(primary), _ptr(jacobians), _ptr(residuals),
                                  _ptr(weights) if weights is not None else None, float(huber_delta))
    out = pm_normal_equations()
    check(lib().pm_normal_equations_add_primary(sw.handle, ctypes.byref(ne.desc), ctypes.byref(b),
                                                ctypes.byref(out), current_stream()))
    return NormalEquations(out, (ne.keep, primary, jacobians, residuals, weights))


def upload_normal_equations(pc, ec, ps, es, exported: dict, device="cuda") -> NormalEquations:
    """Normal equations from host arrays (the layout of BlockNormalEquations'
    members, e.g. oracle.NormalEquations.export())."""
    import torch

    def f64(a):
        return torch.as_tensor(a, dtype=torch.float64).to(device).contiguous()

    def i64(a):
        return torch.as_tensor(a, dtype=torch.int64).to(device).contiguous()

    t = dict(B=f64(exported["primary_diagonal"]), C=f64(exported["eliminated_diagonal"]),
             gp=f64(exported["primary_gradient"]), ge=f64(exported["eliminated_gradient"]),
             cp=i64(exported["cross_primary"]), ce=i64(exported["cross_eliminated"]),
             W=f64(exported["cross_values"]), pr=i64(exported["primary_cross_rows"]),
             pcc=i64(exported["primary_cross_cols"]), pv=f64(exported["primary_cross_values"]))
    d = pm_normal_equations(pc, ec, ps, es, _ptr(t["B"]), _ptr(t["C"]), _ptr(t["gp"]), _ptr(t["ge"]),
                            len(exported["cross_primary"]), _ptr(t["cp"]), _ptr(t["ce"]), _ptr(t["W"]),
                            len(exported["primary_cross_rows"]), _ptr(t["pr"]), _ptr(t["pcc"]), _ptr(t["pv"]))
    return NormalEquations(d, tuple(t.values()))


def normal_equations_arrays(ne: NormalEquations) -> dict:
    """Host copies of the device normal equations (for checks)."""
    d = ne.desc
    pc, ec, ps, es = d.primary_count, d.eliminated_count, d.primary_size, d.eliminated_size
    nc = d.cross_count
    return dict(primary_diagonal=device_array(d.primary_diagonal, pc * ps * ps),
                eliminated_diagonal=device_array(d.eliminated_diagonal, ec * es * es),
                primary_gradient=device_array(d.primary_gradient, pc * ps),
                eliminated_gradient=device_array(d.eliminated_gradient, ec * es),
                cross_primary=device_array(d.cross_primary, nc, "int64"),
                cross_eliminated=device_array(d.cross_eliminated, nc, "int64"),
                cross_values=device_array(d.cross_values, nc * ps * es),
                primary_cross_rows=device_array(d.primary_cross_rows, d.primary_cross_count, "int64"),
                primary_cross_cols=device_array(d.primary_cross_cols, d.primary_cross_count, "int64"),
                primary_cross_values=device_array(d.primary_cross_values, d.primary_cross_count * ps * ps))


def lib_memcpy(dst: int, src: int, nbytes: int) -> int:
    """Device-to-device copy on the current stream, synchronised (the HIP
    runtime symbols resolve through the library's own dependencies)."""
    import torch

    stream = _p(torch.cuda.current_stream().cuda_stream)
    err = lib().hipMemcpyAsync(_p(dst), _p(src), ctypes.c_size_t(nbytes), ctypes.c_int(3), stream)
    if err == 0:
        err = lib().hipStreamSynchronize(stream)
    if err != 0:
        raise PlamatrixError(6, f"hipMemcpyAsync failed with HIP error {err}")
    return 0


@dataclass
class SchurReport:
    converged: bool
    iterations: int
    initial_residual_norm: float
    final_residual_norm: float
    status: int
    pattern_reused: bool
    assembly_on_device: bool
    pattern_build_count: int
    message: str
    mixed_precision_used: bool = False


PM_SCHUR_TOLERANCE_SUM = 1
PM_SCHUR_MIXED_PRECISION = 2


def schur_solve(sw: SchurWorkspace, ne: NormalEquations, damping, primary_step, eliminated_step, ws: Workspace,
                max_iterations=100, relative_tolerance=1e-8, absolute_tolerance=1e-12, minimum_diagonal=1e-12,
                use_initial_guess=False, dense=False, tolerance_sum=False, mixed_precision=False) -> SchurReport:
    """dense=True selects the dense Cholesky backend (the reference's DenseCpu);
    tolerance_sum=True the Cpu backend's stop rule abs + rel ||r0||;
    mixed_precision=True the FP32-seeded solve (useMixedPrecision)."""
    flags = (PM_SCHUR_TOLERANCE_SUM if tolerance_sum else 0) | (PM_SCHUR_MIXED_PRECISION if mixed_precision else 0)
    o = pm_schur_options(int(max_iterations), int(bool(use_initial_guess)), float(relative_tolerance),
                         float(absolute_tolerance), float(minimum_diagonal),
                         PM_SCHUR_DENSE_CHOLESKY if dense else PM_SCHUR_BLOCK_PCG, flags)
    r = pm_schur_report()
    check(lib().pm_schur_solve(sw.handle, ctypes.byref(ne.desc), _f64(damping), ctypes.byref(o),
                               _ptr(primary_step), _ptr(eliminated_step), ctypes.byref(r), ws.handle,
                               current_stream()))
    return SchurReport(bool(r.converged), r.iterations, r.initial_residual_norm, r.final_residual_norm, r.status,
                       bool(r.pattern_reused), bool(r.assembly_on_device), r.pattern_build_count,
                       r.message.decode(), bool(r.mixed_precision_used))


def schur_assemble(sw: SchurWorkspace, ne: NormalEquations, damping, minimum_diagonal=1e-12) -> SchurReport:
    r = pm_schur_report()
    check(lib().pm_schur_assemble(sw.handle, ctypes.byref(ne.desc), _f64(damping), _f64(minimum_diagonal),
                                  ctypes.byref(r), current_stream()))
    return SchurReport(bool(r.converged), r.iterations, r.initial_residual_norm, r.final_residual_norm, r.status,
                       bool(r.pattern_reused), bool(r.assembly_on_device), r.pattern_build_count,
                       r.message.decode())


def schur_system(sw: SchurWorkspace):
    """(pm_csr descriptor, rhs pointer, preconditioner pointer) of the last solve."""
    a = pm_csr()
    rhs = _p()
    pre = _p()
    check(lib().pm_schur_system(sw.handle, ctypes.byref(a), ctypes.byref(rhs), ctypes.byref(pre)))
    return a, rhs.value, pre.value


def device_array(ptr: int, n: int, dtype="float64"):
    """Host numpy copy of n elements at a device pointer."""
    import numpy as np
    import torch

    tdt = {"float64": torch.float64, "float32": torch.float32, "int64": torch.int64, "int32": torch.int32}[dtype]
    t = torch.empty(max(n, 1), dtype=tdt, device="cuda")
    if n > 0:
        check(lib_memcpy(t.data_ptr(), ptr, n * t.element_size()))
    return t[:n].cpu().numpy().astype(np.dtype(dtype))


# ---- batched PCG -----------------------------------------------------------------


def batch_layout(sizes):
    """Row starts (64-aligned) and total rows for systems of the given sizes."""
    import numpy as np

    sizes = np.ascontiguousarray(sizes, dtype=np.int64)
    begin = np.zeros(len(sizes), dtype=np.int64)
    total = _i64()
    check(lib().pm_batch_layout(_i64(len(sizes)), sizes.ctypes.data_as(_p), begin.ctypes.data_as(_p),
                                ctypes.byref(total)))
    return begin, total.value


@dataclass
class BatchReport:
    converged: bool
    iterations: int
    initial_residual: float
    final_residual: float
    status: int
    bad_row: int


def batch_pcg(a: "DeviceCSR", row_begin, row_end, b, x, ws: Workspace, pol=None, **opts):
    import numpy as np

    d = a.descriptor()
    o = options(**opts)
    p = pol if pol is not None else policy()
    row_begin = np.ascontiguousarray(row_begin, dtype=np.int64)
    row_end = np.ascontiguousarray(row_end, dtype=np.int64)
    systems = len(row_begin)
    reps = (pm_batch_report * max(systems, 1))()
    check(lib().pm_batch_pcg(ctypes.byref(d), _i64(systems), row_begin.ctypes.data_as(_p),
                             row_end.ctypes.data_as(_p), _ptr(b), _ptr(x), ctypes.byref(o), ctypes.byref(p), reps,
                             ws.handle, current_stream()))
    return [BatchReport(bool(r.converged), r.iterations, r.initial_residual, r.final_residual, r.status, r.bad_row)
            for r in reps[:systems]]


@dataclass
class BatchRefineReport:
    converged: bool
    outer_iterations: int
    inner_iterations: int
    initial_relres: float
    final_relres: float
    last_inner_status: int


def batch_pcg_refine(a: "DeviceCSR", row_begin, row_end, b64, x64, ws: Workspace, target=1e-6, max_outer=10,
                     pol=None, **inner):
    """Batched FP64 refinement: FP32 batched PCG inner solves, FP64 x and
    residuals; one report per system."""
    import numpy as np

    d = a.descriptor()
    o = options(**inner)
    p = pol if pol is not None else policy()
    row_begin = np.ascontiguousarray(row_begin, dtype=np.int64)
    row_end = np.ascontiguousarray(row_end, dtype=np.int64)
    systems = len(row_begin)
    reps = (pm_batch_refine_report * max(systems, 1))()
    check(lib().pm_batch_pcg_refine(ctypes.byref(d), _i64(systems), row_begin.ctypes.data_as(_p),
                                    row_end.ctypes.data_as(_p), _ptr(b64), _ptr(x64), ctypes.byref(o), _f64(target),
                                    _i32(max_outer), ctypes.byref(p), reps, ws.handle, current_stream()))
    return [BatchRefineReport(bool(r.converged), r.outer_iterations, r.inner_iterations, r.initial_relres,
                              r.final_relres, r.last_inner_status) for r in reps[:systems]]


def batch_block_pcg(a: "DeviceCSR", row_begin, row_end, block_sizes, b, x, ws: Workspace, pol=None, **opts):
    """Batched block-Jacobi PCG: per-system block sizes (1 = Jacobi)."""
    import numpy as np

    d = a.descriptor()
    o = options(**opts)
    p = pol if pol is not None else policy()
    row_begin = np.ascontiguousarray(row_begin, dtype=np.int64)
    row_end = np.ascontiguousarray(row_end, dtype=np.int64)
    bs = np.ascontiguousarray(block_sizes, dtype=np.int32)
    systems = len(row_begin)
    reps = (pm_batch_report * max(systems, 1))()
    check(lib().pm_batch_block_pcg(ctypes.byref(d), _i64(systems), row_begin.ctypes.data_as(_p),
                                   row_end.ctypes.data_as(_p), bs.ctypes.data_as(_p), _ptr(b), _ptr(x),
                                   ctypes.byref(o), ctypes.byref(p), reps, ws.handle, current_stream()))
    return [BatchReport(bool(r.converged), r.iterations, r.initial_residual, r.final_residual, r.status, r.bad_row)
            for r in reps[:systems]]


def batch_block_pcg_refine(a: "DeviceCSR", row_begin, row_end, block_sizes, b64, x64, ws: Workspace, target=1e-6,
                           max_outer=10, pol=None, **inner):
    """Batched FP64 refinement with per-system block-Jacobi inner solves."""
    import numpy as np

    d = a.descriptor()
    o = options(**inner)
    p = pol if pol is not None else policy()
    row_begin = np.ascontiguousarray(row_begin, dtype=np.int64)
    row_end = np.ascontiguousarray(row_end, dtype=np.int64)
    bs = np.ascontiguousarray(block_sizes, dtype=np.int32)
    systems = len(row_begin)
    reps = (pm_batch_refine_report * max(systems, 1))()
    check(lib().pm_batch_block_pcg_refine(ctypes.byref(d), _i64(systems), row_begin.ctypes.data_as(_p),
                                          row_end.ctypes.data_as(_p), bs.ctypes.data_as(_p), _ptr(b64), _ptr(x64),
                                          ctypes.byref(o), _f64(target), _i32(max_outer), ctypes.byref(p), reps,
                                          ws.handle, current_stream()))
    return [BatchRefineReport(bool(r.converged), r.outer_iterations, r.inner_iterations, r.initial_relres,
                              r.final_relres, r.last_inner_status) for r in reps[:systems]]


# ---- learned factorized preconditioner ---------------------------------------------


def fsai_parameter_count() -> int:
    return int(lib().pm_fsai_parameter_count())


def fsai_initial_parameters(seed: int = 0):
    import numpy as np

    out = np.zeros(fsai_parameter_count(), dtype=np.float32)
    check(lib().pm_fsai_initial_parameters(ctypes.c_uint64(seed), out.ctypes.data_as(_p), _i64(len(out))))
    return out


class FsaiNet:
    """Owns a pm_fsai_net (FP32 master parameters, Adam state, BF16 weights)."""

    def __init__(self, seed: int = 0):
        h = _p()
        check(lib().pm_fsai_net_create(ctypes.c_uint64(seed), ctypes.byref(h)))
        self.handle = h

    @property
    def parameters(self):
        import numpy as np

        out = np.zeros(fsai_parameter_count(), dtype=np.float32)
        check(lib().pm_fsai_net_get_parameters(self.handle, out.ctypes.data_as(_p), _i64(len(out))))
        return out

    @parameters.setter
    def parameters(self, values):
        import numpy as np

        values = np.ascontiguousarray(values, dtype=np.float32)
        check(lib().pm_fsai_net_set_parameters(self.handle, values.ctypes.data_as(_p), _i64(len(values))))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and _lib is not None:
            _lib.pm_fsai_net_destroy(h)
            self.handle = None


class FsaiGraph:
    """Owns a pm_fsai_graph for one SPD matrix (a DeviceCSR with 32-bit indices)."""

    def __init__(self, a: "DeviceCSR"):
        self.a = a
        self.desc = a.descriptor()
        h = _p()
        check(lib().pm_fsai_graph_create(ctypes.byref(self.desc), ctypes.byref(h), current_stream()))
        self.handle = h

    def factors(self):
        g, gt = pm_csr(), pm_csr()
        check(lib().pm_fsai_factors(self.handle, ctypes.byref(g), ctypes.byref(gt)))
        return g, gt

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and _lib is not None:
            _lib.pm_fsai_graph_destroy(h)
            self.handle = None


def fsai_forward(net: FsaiNet, graph: FsaiGraph):
    check(lib().pm_fsai_forward(net.handle, graph.handle, current_stream()))


def fsai_loss_grad(net: FsaiNet, graph: FsaiGraph, probes: int = 8, seed: int = 0):
    import numpy as np

    loss = ctypes.c_double()
    grad = np.zeros(fsai_parameter_count(), dtype=np.float32)
    check(lib().pm_fsai_loss_grad(net.handle, graph.handle, _i32(probes), ctypes.c_uint64(seed), ctypes.byref(loss),
                                  grad.ctypes.data_as(_p), current_stream()))
    return loss.value, grad


def fsai_train_step(net: FsaiNet, graph: FsaiGraph, probes: int = 8, seed: int = 0, lr: float = 1e-3) -> float:
    loss = ctypes.c_double()
    check(lib().pm_fsai_train_step(net.handle, graph.handle, _i32(probes), ctypes.c_uint64(seed),
                                   ctypes.c_float(lr), ctypes.byref(loss), current_stream()))
    return loss.value


def pcg_factored(a: "DeviceCSR", b, x, graph: FsaiGraph, ws: Workspace, pol=None, **opts) -> "Report":
    _system_vectors(a, b, x)
    d = a.descriptor()
    g, gt = graph.factors()
    o = options(**opts)
    p = pol if pol is not None else policy()
    r = pm_solver_report()
    check(lib().pm_pcg_factored(ctypes.byref(d), _ptr(b), _ptr(x), ctypes.byref(g), ctypes.byref(gt),
                                ctypes.byref(o), ctypes.byref(p), ctypes.byref(r), ws.handle, current_stream()))
    return _report(r)


def pcg_refine_factored(a: "DeviceCSR", b64, x64, graph: FsaiGraph, ws: Workspace, target=1e-8, max_outer=10,
                        pol=None, **inner) -> "RefineReport":
    d = a.descriptor()
    g, gt = graph.factors()
    o = options(**inner)
    p = pol if pol is not None else policy()
    rep = pm_refine_report()
    check(lib().pm_pcg_refine_factored(ctypes.byref(d), _ptr(b64), _ptr(x64), ctypes.byref(g), ctypes.byref(gt),
                                       ctypes.byref(o), _f64(target), _i32(max_outer), ctypes.byref(p),
                                       ctypes.byref(rep), ws.handle, current_stream()))
    return RefineReport(bool(rep.converged), rep.outer_iterations, rep.inner_iterations, rep.initial_relres,
                        rep.final_relres)


def fsai_batch_select(graph: FsaiGraph, row_begin, row_end, probes: int = 8, seed: int = 0):
    """Per-system choice between the learned factor and Jacobi on a batched
    graph: a uint8 device tensor (1 = Jacobi) and the host losses (systems x 2:
    learned, Jacobi)."""
    import numpy as np
    import torch

    row_begin = np.ascontiguousarray(row_begin, dtype=np.int64)
    row_end = np.ascontiguousarray(row_end, dtype=np.int64)
    systems = len(row_begin)
    flags = torch.zeros(max(systems, 1), dtype=torch.uint8, device="cuda")
    losses = np.zeros((max(systems, 1), 2), dtype=np.float64)
    check(lib().pm_fsai_batch_select(graph.handle, _i64(systems), row_begin.ctypes.data_as(_p),
                                     row_end.ctypes.data_as(_p), _i32(probes), ctypes.c_uint64(seed),
                                     _ptr(flags), losses.ctypes.data_as(_p), current_stream()))
    return flags[:systems], losses[:systems]


def batch_pcg_factored(a: "DeviceCSR", row_begin, row_end, b, x, graph: FsaiGraph, ws: Workspace, jacobi=None,
                       pol=None, factors=None, **opts):
    """Batched PCG with the graph's factors (or explicit (G, G^T) descriptors);
    `jacobi` is an optional uint8 device tensor of per-system Jacobi flags."""
    import numpy as np

    d = a.descriptor()
    g, gt = factors if factors is not None else graph.factors()
    o = options(**opts)
    p = pol if pol is not None else policy()
    row_begin = np.ascontiguousarray(row_begin, dtype=np.int64)
    row_end = np.ascontiguousarray(row_end, dtype=np.int64)
    systems = len(row_begin)
    reps = (pm_batch_report * max(systems, 1))()
    check(lib().pm_batch_pcg_factored(ctypes.byref(d), _i64(systems), row_begin.ctypes.data_as(_p),
                                      row_end.ctypes.data_as(_p), _ptr(b), _ptr(x), ctypes.byref(g), ctypes.byref(gt),
                                      _ptr(jacobi) if jacobi is not None else None, ctypes.byref(o), ctypes.byref(p),
                                      reps, ws.handle, current_stream()))
    return [BatchReport(bool(r.converged), r.iterations, r.initial_residual, r.final_residual, r.status, r.bad_row)
            for r in reps[:systems]]


def batch_pcg_refine_factored(a: "DeviceCSR", row_begin, row_end, b64, x64, graph: FsaiGraph, ws: Workspace,
                              jacobi=None, target=1e-6, max_outer=10, pol=None, factors=None, **inner):
    """Batched FP64 refinement with the factored preconditioner in the FP32
    inner batched solves; one report per system."""
    import numpy as np

    d = a.descriptor()
    g, gt = factors if factors is not None else graph.factors()
    o = options(**inner)
    p = pol if pol is not None else policy()
    row_begin = np.ascontiguousarray(row_begin, dtype=np.int64)
    row_end = np.ascontiguousarray(row_end, dtype=np.int64)
    systems = len(row_begin)
    reps = (pm_batch_refine_report * max(systems, 1))()
    check(lib().pm_batch_pcg_refine_factored(ctypes.byref(d), _i64(systems), row_begin.ctypes.data_as(_p),
                                             row_end.ctypes.data_as(_p), _ptr(b64), _ptr(x64), ctypes.byref(g),
                                             ctypes.byref(gt), _ptr(jacobi) if jacobi is not None else None,
                                             ctypes.byref(o), _f64(target), _i32(max_outer), ctypes.byref(p), reps,
                                             ws.handle, current_stream()))
    return [BatchRefineReport(bool(r.converged), r.outer_iterations, r.inner_iterations, r.initial_relres,
                              r.final_relres, r.last_inner_status) for r in reps[:systems]]


def csr_arrays(d: pm_csr):
    """Host copies (row_offsets, col_indices, values) of a CSR descriptor's arrays."""
    ro = device_array(d.row_offsets32, d.rows + 1, "int32").astype("int64")
    ci = device_array(d.col_indices32, d.nnz, "int32").astype("int64")
    vals = device_array(d.values, d.nnz, "float64" if d.dtype == PM_F64 else "float32")
    return ro, ci, vals
