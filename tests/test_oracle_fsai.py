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
