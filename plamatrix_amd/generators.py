"""Deterministic synthetic problem generators (host side, numpy).

Every generator is a pure function of its arguments and seed (numpy PCG64), so
the GPU path, the tests and the CPU baseline see identical inputs.

* ``poisson2d_coo``      2-D 5-point Dirichlet Laplacian as shuffled COO
                         triplets with split duplicates (COO->CSR workload).
* ``diffusion3d_csr``    3-D 7-point variable-coefficient diffusion, symmetric
                         diagonal scaling to unit diagonal (SpMV / PCG workload).
* ``lm_schur_instance``  bundle-adjustment-like block least squares: primaries
                         on a ring, each secondary seen by a window of
                         consecutive primaries, Huber-weighted 2-D residuals.
* ``laplacian2d_csr``    2-D Laplacian directly in CSR (batched workload).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class COO:
    rows: int
    cols: int
    row_indices: np.ndarray
    col_indices: np.ndarray
    values: np.ndarray


@dataclass
class CSRHost:
    rows: int
    cols: int
    row_offsets: np.ndarray  # int64 [rows+1]
    col_indices: np.ndarray  # int64 [nnz]
    values: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.row_offsets[-1])


def _stencil_csr(n: int, offsets: list[int], valid: np.ndarray, vals: np.ndarray, dtype) -> CSRHost:
    """Assemble CSR from a (n, k) table whose columns are sorted by offset."""
    cols = np.arange(n, dtype=np.int64)[:, None] + np.asarray(offsets, dtype=np.int64)[None, :]
    counts = valid.sum(axis=1)
    row_offsets = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=row_offsets[1:])
    return CSRHost(n, n, row_offsets, cols[valid].astype(np.int64), vals[valid].astype(dtype))


def laplacian2d_csr(nx: int, ny: int | None = None, dtype=np.float64) -> CSRHost:
    """5-point Dirichlet Laplacian (diagonal 4, neighbours -1), row-major grid."""
    ny = nx if ny is None else ny
    n = nx * ny
    idx = np.arange(n)
    i, j = idx % nx, idx // nx
    offsets = [-nx, -1, 0, 1, nx]
    valid = np.stack([j > 0, i > 0, np.ones(n, bool), i < nx - 1, j < ny - 1], axis=1)
    vals = np.tile(np.array([-1.0, -1.0, 4.0, -1.0, -1.0]), (n, 1))
    return _stencil_csr(n, offsets, valid, vals, dtype)


def poisson2d_coo(nx: int = 256, seed: int = 0x0C0FFEE0, duplicate_fraction: float = 0.1,
                  dtype=np.float64) -> COO:
    """2-D 5-point Laplacian triplets, shuffled, with a fraction of entries split
    into two triplets ``(a, v - a)`` where ``a = v * 2**-k`` is dyadic, so both
    parts and their sum are exact and the folded CSR equals the unsplit matrix."""
    csr = laplacian2d_csr(nx, nx, np.float64)
    rows = np.repeat(np.arange(csr.rows, dtype=np.int64), np.diff(csr.row_offsets))
    cols = csr.col_indices.copy()
    vals = csr.values.copy()
    rng = np.random.default_rng(seed)
    z = len(vals)
    split = rng.permutation(z)[: int(round(duplicate_fraction * z))]
    k = rng.integers(1, 4, len(split))
    first = vals[split] * np.exp2(-k.astype(np.float64))
    second = vals[split] - first
    vals[split] = first
    rows = np.concatenate([rows, rows[split]])
    cols = np.concatenate([cols, cols[split]])
    vals = np.concatenate([vals, second])
    order = rng.permutation(len(vals))
    return COO(csr.rows, csr.cols, rows[order], cols[order], vals[order].astype(dtype))


def diffusion3d_csr(n: int = 128, seed: int = 2, sigma: float = 1.5, z_scale: float = 0.1,
                    dtype=np.float32) -> CSRHost:
    """7-point finite-volume diffusion on an n^3 grid with Dirichlet boundary.

    Face coefficients are log-normal ``exp(sigma * N(0,1))``; z faces are
    multiplied by ``z_scale`` (10:1 anisotropy for 0.1).  The SPD matrix is
    scaled as ``D^-1/2 A D^-1/2`` so its diagonal is exactly one."""
    return diffusion3d_box_csr(n, n, n, seed, sigma, z_scale, dtype)


def diffusion3d_box_csr(nx: int, ny: int, nz: int, seed: int = 2, sigma: float = 1.5, z_scale: float = 0.1,
                        dtype=np.float32) -> CSRHost:
    """diffusion3d_csr on an nx x ny x nz grid (x fastest)."""
    rng = np.random.default_rng(seed)
    kx = np.exp(sigma * rng.standard_normal((nz, ny, nx + 1)))          # [z, y, x-face]
    ky = np.exp(sigma * rng.standard_normal((nz, ny + 1, nx)))          # [z, y-face, x]
    kz = np.exp(sigma * rng.standard_normal((nz + 1, ny, nx))) * z_scale  # [z-face, y, x]
    diag = (kx[:, :, :-1] + kx[:, :, 1:] + ky[:, :-1, :] + ky[:, 1:, :] + kz[:-1] + kz[1:]).reshape(-1)
    N = nx * ny * nz
    idx = np.arange(N)
    x, y, z = idx % nx, (idx // nx) % ny, idx // (nx * ny)
    offsets = [-nx * ny, -nx, -1, 0, 1, nx, nx * ny]
    valid = np.stack([z > 0, y > 0, x > 0, np.ones(N, bool), x < nx - 1, y < ny - 1, z < nz - 1], axis=1)
    vals = np.empty((N, 7))
    vals[:, 0] = -kz[:-1].reshape(-1)          # face between z-1 and z
    vals[:, 1] = -ky[:, :-1, :].reshape(-1)
    vals[:, 2] = -kx[:, :, :-1].reshape(-1)
    vals[:, 3] = diag
    vals[:, 4] = -kx[:, :, 1:].reshape(-1)
    vals[:, 5] = -ky[:, 1:, :].reshape(-1)
    vals[:, 6] = -kz[1:].reshape(-1)
    inv_sqrt = 1.0 / np.sqrt(diag)
    cols = idx[:, None] + np.asarray(offsets)[None, :]
    cols_clipped = np.clip(cols, 0, N - 1)
    vals = vals * inv_sqrt[:, None] * inv_sqrt[cols_clipped]
    vals[:, 3] = 1.0
    return _stencil_csr(N, offsets, valid, vals, dtype)


def batch_sizes(count: int = 512, seed: int = 4, lo: int = 1000, hi: int = 200000) -> np.ndarray:
    """Log-uniform system sizes in [lo, hi]."""
    rng = np.random.default_rng(seed)
    return np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(np.int64)


def grid2d_dims(n: int) -> tuple:
    nx = max(1, int(round(np.sqrt(n))))
    return nx, max(1, int(round(n / nx)))


def grid3d_dims(n: int) -> tuple:
    nx = max(1, int(round(n ** (1.0 / 3.0))))
    ny = nx
    return nx, ny, max(1, int(round(n / (nx * ny))))


def block_diagonal(systems: list, dtype=np.float32, pad_identity: bool = False):
    """Pack CSR systems as the diagonal blocks of one matrix, each starting at
    a multiple of 64 rows.  Padding rows are empty, or carry a unit diagonal
    with pad_identity (so the whole matrix is SPD and a learned graph can be
    built over it in one go).  Returns the CSR and the row_begin / row_end
    arrays."""
    sizes = np.array([s.rows for s in systems], dtype=np.int64)
    begin = np.zeros(len(systems), dtype=np.int64)
    cursor = 0
    for i, n in enumerate(sizes):
        begin[i] = cursor
        cursor += -(-int(n) // 64) * 64
    total = cursor
    end = begin + sizes
    counts = np.zeros(total, dtype=np.int64)
    pad = np.ones(total, dtype=bool)
    for s, b0 in zip(systems, begin):
        counts[b0:b0 + s.rows] = np.diff(s.row_offsets)
        pad[b0:b0 + s.rows] = False
    if pad_identity:
        counts[pad] = 1
    ro = np.zeros(total + 1, dtype=np.int64)
    np.cumsum(counts, out=ro[1:])
    ci = np.zeros(int(ro[-1]), dtype=np.int64)
    vals = np.zeros(int(ro[-1]), dtype=dtype)
    for s, b0 in zip(systems, begin):
        k0, k1 = ro[b0], ro[b0 + s.rows]
        ci[k0:k1] = s.col_indices + b0
        vals[k0:k1] = s.values.astype(dtype)
    if pad_identity:
        rows = np.nonzero(pad)[0]
        ci[ro[rows]] = rows
        vals[ro[rows]] = 1.0
    return CSRHost(total, total, ro, ci, vals), begin, end


def rhs_uniform(n: int, seed: int = 1, dtype=np.float64) -> np.ndarray:
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n).astype(dtype)


def rhs_normal(n: int, seed: int = 3, dtype=np.float32) -> np.ndarray:
    return np.random.default_rng(seed).standard_normal(n).astype(dtype)


def huber_weights(residuals: np.ndarray, delta: float) -> np.ndarray:
    """Vectorised ``evaluateHuberLoss(||r||^2, delta).weight``."""
    s = np.einsum("ij,ij->i", residuals, residuals)
    if delta == 0.0:
        return np.ones(len(s))
    w = np.ones(len(s))
    big = s > delta * delta
    w[big] = delta / np.sqrt(s[big])
    return w


@dataclass
class LMInstance:
    primary_count: int
    eliminated_count: int
    primary_size: int
    eliminated_size: int
    primary: np.ndarray              # int64 [R]
    eliminated: np.ndarray           # int64 [R], non-decreasing
    primary_jacobians: np.ndarray    # [R, rs*ps] row-major rs x ps
    eliminated_jacobians: np.ndarray  # [R, rs*es]
    residuals: np.ndarray            # [R, rs]
    weights: np.ndarray              # [R] Huber IRLS weights
    huber_delta: float
    damping: float


def lm_schur_instance(primary_count: int = 2000, eliminated_count: int = 200000, window: tuple = (3, 12),
                      seed: int = 20, residual_size: int = 2, primary_size: int = 6, eliminated_size: int = 3,
                      sigma: float = 0.5, outlier_fraction: float = 0.05, outlier_scale: float = 20.0,
                      huber_delta: float = 1.0, damping: float = 1e-3, mean_window: float | None = None
                      ) -> LMInstance:
    """Secondary e is seen by k_e consecutive primaries (mod primary_count)
    starting at a uniform position; ``k_e = lo + Binomial(hi - lo, p)`` with
    ``p = 1/3`` (mean 6 for the (3, 12) window), one residual per
    observation, residuals ordered by secondary."""
    rng = np.random.default_rng(seed)
    lo, hi = window
    p = 1.0 / 3.0 if mean_window is None else (mean_window - lo) / max(hi - lo, 1)
    k = lo + rng.binomial(hi - lo, p, eliminated_count)
    k = np.minimum(k, primary_count)
    start = rng.integers(0, primary_count, eliminated_count)
    eliminated = np.repeat(np.arange(eliminated_count, dtype=np.int64), k)
    offsets = np.arange(len(eliminated)) - np.repeat(np.cumsum(k) - k, k)
    primary = ((np.repeat(start, k) + offsets) % primary_count).astype(np.int64)
    r_count = len(primary)
    jp = rng.standard_normal((r_count, residual_size * primary_size))
    je = rng.standard_normal((r_count, residual_size * eliminated_size))
    res = sigma * rng.standard_normal((r_count, residual_size))
    outliers = rng.random(r_count) < outlier_fraction
    res[outliers] *= outlier_scale
    w = huber_weights(res, huber_delta)
    return LMInstance(primary_count, eliminated_count, primary_size, eliminated_size, primary, eliminated,
                      jp, je, res, w, huber_delta, damping)
