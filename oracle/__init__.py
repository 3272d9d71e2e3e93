"""ctypes binding of the oracle (oracle/*.c) and of the optional reference build.

Test infrastructure only: imported by tests/, by ``__graft_entry__.smoke()`` as
the checker, and by ``bench.py``'s ``cpu_baseline`` leg.  The product package
(``plamatrix_amd``) never imports this module.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libplamatrix_ref.so")

_i64 = ctypes.c_int64
_p = ctypes.c_void_p


class OracleError(Exception):
    """Raised with the reference's exception class name and message."""

    KINDS = {1: "invalid_argument", 2: "out_of_range", 3: "overflow_error",
             4: "logic_error", 5: "runtime_error"}

    def __init__(self, code: int, message: str):
        self.code = code
        self.kind = self.KINDS.get(code, "error")
        super().__init__(f"{self.kind}: {message}")


class _Options(ctypes.Structure):
    _fields_ = [("max_iterations", ctypes.c_int), ("relative_tolerance", ctypes.c_double),
                ("absolute_tolerance", ctypes.c_double), ("use_jacobi", ctypes.c_int),
                ("require_convergence", ctypes.c_int)]


class _Report(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int), ("iterations", ctypes.c_int),
                ("initial_residual", ctypes.c_double), ("final_residual", ctypes.c_double)]


class _SchurOptions(ctypes.Structure):
    _fields_ = [("max_iterations", ctypes.c_int), ("relative_tolerance", ctypes.c_double),
                ("absolute_tolerance", ctypes.c_double), ("minimum_diagonal", ctypes.c_double),
                ("use_initial_guess", ctypes.c_int)]


class _SchurReport(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int), ("iterations", ctypes.c_int),
                ("initial_residual_norm", ctypes.c_double), ("final_residual_norm", ctypes.c_double),
                ("status", ctypes.c_int)]


@dataclass
class Report:
    converged: bool
    iterations: int
    initial_residual: float
    final_residual: float


@dataclass
class CSR:
    rows: int
    cols: int
    row_offsets: np.ndarray
    col_indices: np.ndarray
    values: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.row_offsets[-1]) if self.rows >= 0 else 0


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(_p)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library not built: {_LIB_PATH} (run `make -C oracle`)")
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.or_last_error.restype = ctypes.c_char_p
        _lib.or_true_relres_f64.restype = ctypes.c_double
        _lib.or_true_relres_f32.restype = ctypes.c_double
        _lib.or_true_relres_f32x64.restype = ctypes.c_double
        _lib.or_ne_create.restype = _p
        _lib.or_ne_create.argtypes = [_i64, _i64, _i64, _i64]
        _lib.or_ne_destroy.argtypes = [_p]
        _lib.or_ne_cross_count.restype = _i64
        _lib.or_ne_cross_count.argtypes = [_p]
        _lib.or_ne_primary_cross_count.restype = _i64
        _lib.or_ne_primary_cross_count.argtypes = [_p]
    return _lib


def ref_available() -> bool:
    return os.path.exists(_REF_PATH)


def ref():
    """The reference's own CPU build (oracle/_ref), when it was built here."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("reference build not available")
        _ref = ctypes.CDLL(_REF_PATH)
        _ref.ref_last_error.restype = ctypes.c_char_p
    return _ref


def _check(code: int, library=None):
    if code != 0:
        library = library or lib()
        getter = library.or_last_error if library is lib() else library.ref_last_error
        raise OracleError(code, getter().decode())


def _suffix(dtype) -> str:
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def coo_to_csr(rows, cols, row_indices, col_indices, values, use_ref=False) -> CSR:
    """cooToCsr (sparse_ops_cpu.cpp:180-191)."""
    r = np.ascontiguousarray(row_indices, dtype=np.int64)
    c = np.ascontiguousarray(col_indices, dtype=np.int64)
    v = np.ascontiguousarray(values)
    if v.dtype not in (np.float32, np.float64):
        v = v.astype(np.float64)
    count = len(v)
    if len(r) != count or len(c) != count:
        raise OracleError(1, "cooToCsr row, column, and value arrays must have equal lengths")
    ro = np.zeros(max(rows, 0) + 1, dtype=np.int64)
    co = np.zeros(max(count, 1), dtype=np.int64)
    vo = np.zeros(max(count, 1), dtype=v.dtype)
    nnz = _i64(0)
    library = ref() if use_ref else lib()
    fn = getattr(library, ("ref_" if use_ref else "or_") + "coo_to_csr_" + _suffix(v.dtype))
    _check(fn(_i64(rows), _i64(cols), _i64(count), _ptr(r), _ptr(c), _ptr(v), _ptr(ro), _ptr(co), _ptr(vo),
              ctypes.byref(nnz)), library)
    n = nnz.value
    return CSR(rows, cols, ro, co[:n].copy(), vo[:n].copy())


def spmv(csr: CSR, x: np.ndarray, use_ref=False) -> np.ndarray:
    """spmv (sparse_ops_cpu.cpp:214-233), Scalar accumulation."""
    x = np.ascontiguousarray(x, dtype=csr.values.dtype)
    y = np.zeros(csr.rows, dtype=csr.values.dtype)
    sfx = _suffix(csr.values.dtype)
    if use_ref:
        _check(getattr(ref(), "ref_spmv_" + sfx)(_i64(csr.rows), _i64(csr.cols), _i64(csr.nnz),
                                                  _ptr(csr.row_offsets), _ptr(csr.col_indices),
                                                  _ptr(csr.values), _ptr(x), _ptr(y)), ref())
    else:
        _check(getattr(lib(), "or_spmv_" + sfx)(_i64(csr.rows), _ptr(csr.row_offsets), _ptr(csr.col_indices),
                                                 _ptr(csr.values), _ptr(x), _ptr(y)))
    return y


def spmm(csr: CSR, b: np.ndarray, use_ref=False) -> np.ndarray:
    """spmm (sparse_ops_cpu.cpp:244-268): B (cols x k) and the result (rows x k)
    column-major, Scalar accumulation per column."""
    b = np.asfortranarray(np.asarray(b, dtype=csr.values.dtype).reshape(csr.cols, -1))
    k = b.shape[1]
    y = np.zeros((csr.rows, k), dtype=csr.values.dtype, order="F")
    sfx = _suffix(csr.values.dtype)
    if use_ref:
        _check(getattr(ref(), "ref_spmm_" + sfx)(_i64(csr.rows), _i64(csr.cols), _i64(csr.nnz),
                                                  _ptr(csr.row_offsets), _ptr(csr.col_indices),
                                                  _ptr(csr.values), _i64(k), _ptr(b), _ptr(y)), ref())
    else:
        _check(getattr(lib(), "or_spmm_" + sfx)(_i64(csr.rows), _i64(csr.cols), _i64(k), _ptr(csr.row_offsets),
                                                 _ptr(csr.col_indices), _ptr(csr.values), _ptr(b), _ptr(y)))
    return y


def _options(max_iterations=1000, relative_tolerance=1e-6, absolute_tolerance=0.0, use_jacobi=True,
             require_convergence=False) -> _Options:
    return _Options(int(max_iterations), float(relative_tolerance), float(absolute_tolerance),
                    int(bool(use_jacobi)), int(bool(require_convergence)))


def _solve(csr: CSR, rhs, x0, preconditioned, use_ref=False, **opts):
    dt = csr.values.dtype
    b = np.ascontiguousarray(rhs, dtype=dt)
    x = np.array(x0 if x0 is not None else np.zeros(csr.cols), dtype=dt)
    o = _options(**opts)
    rep = _Report()
    sfx = _suffix(dt)
    if use_ref:
        library = ref()
        fn = getattr(library, "ref_solve_" + sfx)
        code = fn(_i64(csr.rows), _i64(csr.cols), _i64(csr.nnz), _ptr(csr.row_offsets), _ptr(csr.col_indices),
                  _ptr(csr.values), _ptr(b), _ptr(x), ctypes.byref(o), ctypes.c_int(preconditioned),
                  ctypes.byref(rep))
    else:
        library = lib()
        fn = getattr(library, "or_solve_" + sfx)
        code = fn(_i64(csr.rows), _i64(csr.cols), _i64(csr.nnz), _ptr(csr.row_offsets), _ptr(csr.col_indices),
                  _ptr(csr.values), _ptr(b), _ptr(x), ctypes.byref(o), ctypes.c_int(preconditioned),
                  ctypes.byref(rep))
    _check(code, library)
    return x, Report(bool(rep.converged), rep.iterations, rep.initial_residual, rep.final_residual)


def cg(csr, rhs, x0=None, use_ref=False, **opts):
    """cg (iterative_solver_cpu.cpp:288-295)."""
    return _solve(csr, rhs, x0, 0, use_ref, **opts)


def pcg(csr, rhs, x0=None, use_ref=False, **opts):
    """pcg (iterative_solver_cpu.cpp:297-304)."""
    return _solve(csr, rhs, x0, 1, use_ref, **opts)


def block_pcg(csr, rhs, inverse_blocks, block_size, x0=None, **opts):
    dt = csr.values.dtype
    b = np.ascontiguousarray(rhs, dtype=dt)
    x = np.array(x0 if x0 is not None else np.zeros(csr.cols), dtype=dt)
    inv = np.ascontiguousarray(inverse_blocks, dtype=dt)
    o = _options(**opts)
    rep = _Report()
    fn = getattr(lib(), "or_block_pcg_" + _suffix(dt))
    _check(fn(_i64(csr.rows), _i64(csr.nnz), _ptr(csr.row_offsets), _ptr(csr.col_indices), _ptr(csr.values),
              _ptr(b), _ptr(x), _ptr(inv), _i64(block_size), ctypes.byref(o), ctypes.byref(rep)))
    return x, Report(bool(rep.converged), rep.iterations, rep.initial_residual, rep.final_residual)


def true_relres(csr: CSR, rhs, x) -> float:
    """FP64 ||b - A x|| / ||b||."""
    if csr.values.dtype == np.float64:
        return lib().or_true_relres_f64(_i64(csr.rows), _ptr(csr.row_offsets), _ptr(csr.col_indices),
                                        _ptr(csr.values), _ptr(np.ascontiguousarray(rhs, np.float64)),
                                        _ptr(np.ascontiguousarray(x, np.float64)))
    b = np.ascontiguousarray(rhs, np.float64)
    xx = np.ascontiguousarray(x, np.float64)
    return lib().or_true_relres_f32x64(_i64(csr.rows), _ptr(csr.row_offsets), _ptr(csr.col_indices),
                                       _ptr(csr.values), _ptr(b), _ptr(xx))


def huber(squared_norm: float, delta: float):
    cost = ctypes.c_double()
    weight = ctypes.c_double()
    _check(lib().or_huber_f64(ctypes.c_double(squared_norm), ctypes.c_double(delta), ctypes.byref(cost),
                              ctypes.byref(weight)))
    return cost.value, weight.value


class NormalEquations:
    """BlockNormalEquations<double> (block_schur.h:197-319) restated in C."""

    def __init__(self, primary_count, eliminated_count, primary_size, eliminated_size):
        self._h = lib().or_ne_create(_i64(primary_count), _i64(eliminated_count), _i64(primary_size),
                                     _i64(eliminated_size))
        if not self._h:
            raise OracleError(1, lib().or_last_error().decode())
        self.pc, self.ec, self.ps, self.es = primary_count, eliminated_count, primary_size, eliminated_size

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_ne_destroy(self._h)
            self._h = None

    def add_residual_blocks(self, primary_blocks, primary_jacobians, eliminated_block, eliminated_jacobian,
                            residual, weight=1.0):
        res = np.ascontiguousarray(residual, np.float64).ravel()
        jacs = [np.ascontiguousarray(j, np.float64).ravel() for j in primary_jacobians]
        blocks = np.ascontiguousarray(primary_blocks, np.int64)
        jac_ptrs = (ctypes.c_void_p * max(len(jacs), 1))(*[j.ctypes.data for j in jacs])
        ej = np.ascontiguousarray(eliminated_jacobian, np.float64).ravel()
        _check(lib().or_ne_add_residual_blocks(_p(self._h), _i64(len(blocks)), _ptr(blocks), jac_ptrs,
                                               _i64(eliminated_block), _ptr(ej), _ptr(res), _i64(len(res)),
                                               ctypes.c_double(weight)))

    def add_residual_block(self, primary, eliminated, pj, ej, residual, weight=1.0):
        self.add_residual_blocks([primary], [pj], eliminated, ej, residual, weight)

    def add_residual_batch(self, primary, eliminated, jp, je, r, weights=None):
        primary = np.ascontiguousarray(primary, np.int64)
        eliminated = np.ascontiguousarray(eliminated, np.int64)
        jp = np.ascontiguousarray(jp, np.float64)
        je = np.ascontiguousarray(je, np.float64)
        r = np.ascontiguousarray(r, np.float64)
        rs = r.shape[1] if r.ndim == 2 else 1
        w = None if weights is None else np.ascontiguousarray(weights, np.float64)
        _check(lib().or_ne_add_residual_batch(_p(self._h), _i64(len(primary)), _ptr(primary), _ptr(eliminated),
                                              _ptr(jp), _ptr(je), _ptr(r), _i64(rs), _ptr(w)))

    def add_primary_residual_blocks(self, primary_blocks, primary_jacobians, residual, weight=1.0):
        res = np.ascontiguousarray(residual, np.float64).ravel()
        jacs = [np.ascontiguousarray(j, np.float64).ravel() for j in primary_jacobians]
        blocks = np.ascontiguousarray(primary_blocks, np.int64)
        jac_ptrs = (ctypes.c_void_p * max(len(jacs), 1))(*[j.ctypes.data for j in jacs])
        _check(lib().or_ne_add_primary_residual_blocks(_p(self._h), _i64(len(blocks)), _ptr(blocks), jac_ptrs,
                                                       _ptr(res), _i64(len(res)), ctypes.c_double(weight)))

    def add_primary_residual_block(self, primary, pj, residual, weight=1.0):
        self.add_primary_residual_blocks([primary], [pj], residual, weight)

    def add_eliminated_residual_block(self, eliminated, ej, residual, weight=1.0):
        res = np.ascontiguousarray(residual, np.float64).ravel()
        ej = np.ascontiguousarray(ej, np.float64).ravel()
        _check(lib().or_ne_add_eliminated_residual_block(_p(self._h), _i64(eliminated), _ptr(ej), _ptr(res),
                                                         _i64(len(res)), ctypes.c_double(weight)))

    def merge_from(self, other: "NormalEquations"):
        _check(lib().or_ne_merge(_p(self._h), _p(other._h)))

    def export(self):
        n_cross = lib().or_ne_cross_count(_p(self._h))
        n_pcross = lib().or_ne_primary_cross_count(_p(self._h))
        out = dict(
            primary_diagonal=np.zeros(self.pc * self.ps * self.ps),
            eliminated_diagonal=np.zeros(self.ec * self.es * self.es),
            primary_gradient=np.zeros(self.pc * self.ps),
            eliminated_gradient=np.zeros(self.ec * self.es),
            cross_primary=np.zeros(n_cross, np.int64),
            cross_eliminated=np.zeros(n_cross, np.int64),
            cross_values=np.zeros(n_cross * self.ps * self.es),
            primary_cross_rows=np.zeros(n_pcross, np.int64),
            primary_cross_cols=np.zeros(n_pcross, np.int64),
            primary_cross_values=np.zeros(n_pcross * self.ps * self.ps),
        )
        lib().or_ne_export(_p(self._h), *[_ptr(out[k]) for k in (
            "primary_diagonal", "eliminated_diagonal", "primary_gradient", "eliminated_gradient",
            "cross_primary", "cross_eliminated", "cross_values", "primary_cross_rows", "primary_cross_cols",
            "primary_cross_values")])
        return out

    def solve(self, damping, max_iterations=100, relative_tolerance=1e-8, absolute_tolerance=1e-12,
              minimum_diagonal=1e-12, initial_guess=None):
        o = _SchurOptions(int(max_iterations), float(relative_tolerance), float(absolute_tolerance),
                          float(minimum_diagonal), int(initial_guess is not None))
        dp = np.zeros(self.pc * self.ps) if initial_guess is None else np.array(initial_guess, np.float64)
        de = np.zeros(self.ec * self.es)
        rep = _SchurReport()
        _check(lib().or_schur_solve(_p(self._h), ctypes.c_double(damping), ctypes.byref(o), _ptr(dp), _ptr(de),
                                    ctypes.byref(rep)))
        return dp, de, rep

    def prologue(self, damping, minimum_diagonal=1e-12):
        B = np.zeros(self.pc * self.ps * self.ps)
        Cinv = np.zeros(self.ec * self.es * self.es)
        rhs = np.zeros(self.pc * self.ps)
        Pinv = np.zeros(self.pc * self.ps * self.ps)
        _check(lib().or_schur_prologue(_p(self._h), ctypes.c_double(damping), ctypes.c_double(minimum_diagonal),
                                       _ptr(B), _ptr(Cinv), _ptr(rhs), _ptr(Pinv)))
        return B, Cinv, rhs, Pinv

    def pattern(self):
        slots, values, terms = _i64(), _i64(), _i64()
        lib().or_schur_pattern_size(_p(self._h), ctypes.byref(slots), ctypes.byref(values), ctypes.byref(terms))
        dim = self.pc * self.ps
        out = dict(slot_rows=np.zeros(slots.value, np.int64), slot_cols=np.zeros(slots.value, np.int64),
                   row_offsets=np.zeros(dim + 1, np.int64), col_indices=np.zeros(values.value, np.int64),
                   term_offsets=np.zeros(slots.value + 1, np.int64), term_eliminated=np.zeros(terms.value, np.int64),
                   term_left=np.zeros(terms.value, np.int64), term_right=np.zeros(terms.value, np.int64))
        lib().or_schur_pattern(_p(self._h), *[_ptr(out[k]) for k in (
            "slot_rows", "slot_cols", "row_offsets", "col_indices", "term_offsets", "term_eliminated",
            "term_left", "term_right")])
        return out

    def csr_values(self, damping, minimum_diagonal=1e-12):
        pat = self.pattern()
        values = np.zeros(len(pat["col_indices"]))
        _check(lib().or_schur_csr_values(_p(self._h), ctypes.c_double(damping), ctypes.c_double(minimum_diagonal),
                                         _ptr(values)))
        return CSR(self.pc * self.ps, self.pc * self.ps, pat["row_offsets"], pat["col_indices"], values)


def ref_schur_solve_batch(pc, ec, ps, es, primary, eliminated, jp, je, r, weights, damping, backend=0,
                          max_iterations=100, relative_tolerance=1e-8, absolute_tolerance=1e-12):
    """solveDampedSchurComplement of the reference's own CPU build."""

    class _R(ctypes.Structure):
        _fields_ = [("converged", ctypes.c_int), ("iterations", ctypes.c_int),
                    ("initial", ctypes.c_double), ("final", ctypes.c_double)]

    primary = np.ascontiguousarray(primary, np.int64)
    eliminated = np.ascontiguousarray(eliminated, np.int64)
    jp = np.ascontiguousarray(jp, np.float64)
    je = np.ascontiguousarray(je, np.float64)
    r = np.ascontiguousarray(r, np.float64)
    rs = r.shape[1] if r.ndim == 2 else 1
    w = None if weights is None else np.ascontiguousarray(weights, np.float64)
    dp = np.zeros(pc * ps)
    de = np.zeros(ec * es)
    rep = _R()
    _check(ref().ref_schur_solve_batch(_i64(pc), _i64(ec), _i64(ps), _i64(es), _i64(len(primary)), _ptr(primary),
                                       _ptr(eliminated), _ptr(jp), _ptr(je), _ptr(r), _i64(rs), _ptr(w),
                                       ctypes.c_double(damping), ctypes.c_int(backend), ctypes.c_int(max_iterations),
                                       ctypes.c_double(relative_tolerance), ctypes.c_double(absolute_tolerance),
                                       _ptr(dp), _ptr(de), ctypes.byref(rep)), ref())
    return dp, de, Report(bool(rep.converged), rep.iterations, rep.initial, rep.final)
