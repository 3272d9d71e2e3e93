"""Pins the learned-preconditioner oracle (oracle/fsai.py) on the CPU.

Checks what fixes the restatement independently of the GPU: the initial
parameters give G = D^-1/2 exactly (the Jacobi preconditioner), so the loss
equals the closed form computed from A directly; autograd gradients agree
with central finite differences of the float64 loss; the probe generator is
standard normal; G^T G is SPD for any parameters.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import fsai  # noqa: E402
from plamatrix_amd import generators  # noqa: E402


def small_system(n=4, seed=3):
    a = generators.diffusion3d_box_csr(n, n, 3, seed=seed, dtype=np.float64)
    return a.row_offsets, a.col_indices, a.values


def initial_parameters(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(fsai.PARAMS, np.float32)
    p[fsai.OFF_WIN:fsai.OFF_BIN] = rng.standard_normal(fsai.FEAT * fsai.HID) * np.sqrt(2.0 / fsai.FEAT)
    p[fsai.OFF_BIN:fsai.OFF_WL] = 0.1
    p[fsai.OFF_WL:fsai.OFF_BL] = rng.standard_normal(fsai.LAYERS * 2 * fsai.HID * fsai.HID) * np.sqrt(1.0 / fsai.HID)
    p[fsai.OFF_BL:fsai.OFF_WOFF] = 0.1
    p[fsai.OFF_BDIAG] = np.log(np.e - 1.0)
    return p


def test_initial_g_is_jacobi_and_loss_closed_form():
    ro, ci, vals = small_system()
    p = initial_parameters()
    rows, cols, g = fsai.factors(ro, ci, vals, p, emulate_bf16=False)
    n = len(ro) - 1
    diag = np.array([vals[ro[i]:ro[i + 1]][ci[ro[i]:ro[i + 1]] == i].sum() for i in range(n)])
    on = rows == cols
    np.testing.assert_allclose(g[on], 1.0 / np.sqrt(diag), rtol=1e-6)
    assert np.all(g[~on] == 0.0)
    # loss with G = D^-1/2: (I - D^-1/2 A D^-1/2) applied to the probes
    import scipy.sparse as sp

    A = sp.csr_matrix((vals, ci, ro), shape=(n, n)).toarray()
    Dm = np.diag(1.0 / np.sqrt(diag))
    Z = fsai.probes(n, 8, 5).astype(np.float64)
    R = Z - Dm @ A @ Dm @ Z
    expected = ((R * R).sum(0) / (8 * (Z * Z).sum(0))).sum()
    value, _ = fsai.loss_and_grad(ro, ci, vals, p, 8, 5, emulate_bf16=False)
    assert abs(value - expected) <= 1e-6 * expected  # dinv is FP32, as on the device


def test_autograd_matches_finite_differences():
    ro, ci, vals = small_system(3, seed=1)
    rng = np.random.default_rng(7)
    p = initial_parameters(1).astype(np.float64)
    p[fsai.OFF_WOFF:fsai.OFF_BOFF] = 0.05 * rng.standard_normal(fsai.HID)
    p[fsai.OFF_WDIAG:fsai.OFF_BDIAG] = 0.05 * rng.standard_normal(fsai.HID)
    p[fsai.OFF_ALPHA] = -0.3
    _, grad = fsai.loss_and_grad(ro, ci, vals, p, 4, 11, emulate_bf16=False)
    gr = fsai.graph(ro, ci, vals)
    idx = list(rng.choice(fsai.PARAMS, 24, replace=False)) + [fsai.OFF_ALPHA, fsai.OFF_BDIAG, fsai.OFF_BOFF,
                                                               fsai.OFF_WOFF + 3, fsai.OFF_WIN + 5]
    for k in idx:
        h = 1e-6
        plus, minus = p.copy(), p.copy()
        plus[k] += h
        minus[k] -= h
        fp = float(fsai.loss(gr, torch.tensor(plus), 4, 11, emulate_bf16=False))
        fm = float(fsai.loss(gr, torch.tensor(minus), 4, 11, emulate_bf16=False))
        fd = (fp - fm) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-5 * max(1.0, abs(fd)), (k, fd, grad[k])


def test_probes_are_standard_normal():
    z = fsai.probes(20000, 8, 123).astype(np.float64)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01
    assert not np.array_equal(z, fsai.probes(20000, 8, 124))


def test_factor_gives_spd_preconditioner():
    ro, ci, vals = small_system(3, seed=2)
    rng = np.random.default_rng(3)
    p = initial_parameters(2)
    p[fsai.OFF_WOFF:fsai.OFF_BOFF] = rng.standard_normal(fsai.HID)
    p[fsai.OFF_WDIAG:fsai.OFF_BDIAG] = rng.standard_normal(fsai.HID)
    rows, cols, g = fsai.factors(ro, ci, vals, p)
    n = len(ro) - 1
    G = np.zeros((n, n))
    G[rows, cols] = g
    assert np.all(np.triu(G, 1) == 0) and np.all(np.diag(G) > 0)
    assert np.linalg.eigvalsh(G.T @ G).min() > 0


def _diag_of(ro, ci, vals):
    n = len(ro) - 1
    return np.array([vals[ro[i]:ro[i + 1]][ci[ro[i]:ro[i + 1]] == i].sum() for i in range(n)])


def test_factored_pcg_with_jacobi_factor_reproduces_oracle_pcg():
    """G = D^-1/2 makes G^T G = D^-1, so factored_pcg must follow the oracle's
    Jacobi-PCG (itself bitwise equal to the reference build) iterate by
    iterate: the same iteration count, and every truncated run within 1e-12."""
    import oracle

    a = generators.diffusion3d_box_csr(9, 8, 7, seed=5, dtype=np.float64)
    ro, ci, vals = a.row_offsets, a.col_indices, a.values
    n = len(ro) - 1
    d = _diag_of(ro, ci, vals)
    idx = np.arange(n)
    g = 1.0 / np.sqrt(d)
    b = generators.rhs_uniform(n, seed=9)
    csr = oracle.CSR(n, n, ro, ci, vals)
    x_ref, rep = oracle.pcg(csr, b, relative_tolerance=1e-10, max_iterations=500)
    x, its = fsai.factored_pcg(ro, ci, vals, idx, idx, g, b, relative_tolerance=1e-10, max_iterations=500)
    assert rep.converged and its == rep.iterations
    assert np.linalg.norm(x - x_ref) <= 1e-12 * np.linalg.norm(x_ref)
    for k in (1, 2, 3, 5, 8, 13):
        xk_ref, _ = oracle.pcg(csr, b, relative_tolerance=1e-30, max_iterations=k)
        xk, itk = fsai.factored_pcg(ro, ci, vals, idx, idx, g, b, relative_tolerance=1e-30, max_iterations=k)
        assert itk == k
        assert np.linalg.norm(xk - xk_ref) <= 1e-12 * np.linalg.norm(xk_ref), k


def test_factored_pcg_with_inverse_cholesky_factor_converges_in_one_step():
    """G = L^-1 with A = L L^T gives G^T G = A^-1 exactly: PCG stops after
    one iteration with x = A^-1 b.  The transposed product G G^T is not
    A^-1, and the same run then needs many iterations."""
    import scipy.linalg as sla

    a = generators.diffusion3d_box_csr(5, 4, 3, seed=6, dtype=np.float64)
    ro, ci, vals = a.row_offsets, a.col_indices, a.values
    n = len(ro) - 1
    A = np.zeros((n, n))
    A[np.repeat(np.arange(n), np.diff(ro)), ci] = vals
    L = np.linalg.cholesky(A)
    G = sla.solve_triangular(L, np.eye(n), lower=True)
    rows, cols = np.nonzero(np.tril(np.ones((n, n))))
    b = generators.rhs_uniform(n, seed=2)
    x, its = fsai.factored_pcg(ro, ci, vals, rows, cols, G[rows, cols], b, relative_tolerance=1e-9,
                               max_iterations=200)
    assert its == 1
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)
    # sensitivity: the transposed factor (upper-triangular, applied as G G^T) is not the inverse
    _, its_t = fsai.factored_pcg(ro, ci, vals, cols, rows, G[rows, cols], b, relative_tolerance=1e-9,
                                 max_iterations=200)
    assert its_t > 3


def test_factored_pcg_solution_matches_direct_solve():
    """Any lower G with a positive diagonal gives an SPD M^-1, so PCG
    converges to A^-1 b: compared with scipy's sparse direct solve."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    a = generators.diffusion3d_box_csr(8, 7, 6, seed=8, dtype=np.float64)
    ro, ci, vals = a.row_offsets, a.col_indices, a.values
    n = len(ro) - 1
    rows = np.repeat(np.arange(n), np.diff(ro))
    lower = ci <= rows
    rng = np.random.default_rng(4)
    g = np.where(ci[lower] == rows[lower], 0.5 + rng.random(lower.sum()), 0.2 * rng.standard_normal(lower.sum()))
    b = generators.rhs_uniform(n, seed=3)
    x, its = fsai.factored_pcg(ro, ci, vals, rows[lower], ci[lower], g, b, relative_tolerance=1e-12,
                               max_iterations=2000)
    assert its < 2000
    A = sp.csr_matrix((vals, ci, ro), shape=(n, n)).tocsc()
    x_direct = spla.spsolve(A, b)
    assert np.linalg.norm(x - x_direct) <= 1e-9 * np.linalg.norm(x_direct)
