"""GPU parity of the dense FP64 Cholesky solve: the reference's DenseCpu Schur
backend (block_schur_dense_solver.h:126-398), served on the device.

The factorization has no bitwise CPU counterpart (its trailing updates run on
FP64 MFMA in tile order), so each case is checked against what the mathematics
fixes: the solution of the same symmetric positive definite system by LAPACK
(numpy.linalg.solve) within a forward-error bound stated per case, the residual
norm the kernel reports against the one recomputed from its x, exact
detection of indefinite matrices, and bitwise repeatability.  The Schur
backend is also compared with the reference's own DenseCpu solve
(oracle/_ref, built from the reference's sources) on the same residuals.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from plamatrix_amd import capi, generators  # noqa: E402


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    capi.lib()
    return torch.device("cuda:0")


@pytest.fixture
def ws(dev):
    w = capi.Workspace()
    yield w
    torch.cuda.synchronize()


def dense_to_csr(a, dev, lower_only=False):
    """CSR of every stored entry of a (lower triangle only if asked)."""
    n = a.shape[0]
    mask = np.tril(np.ones_like(a, dtype=bool)) if lower_only else np.ones_like(a, dtype=bool)
    mask &= a != 0.0
    rows, cols = np.nonzero(mask)
    ro = np.zeros(n + 1, dtype=np.int64)
    np.add.at(ro, rows + 1, 1)
    ro = np.cumsum(ro)
    return capi.DeviceCSR(n, n, torch.from_numpy(ro).to(dev), torch.from_numpy(cols.astype(np.int64)).to(dev),
                          torch.from_numpy(a[rows, cols].astype(np.float64)).to(dev))


def spd(n, seed, cond=None):
    rng = np.random.default_rng(seed)
    if cond is None:
        m = rng.standard_normal((n, n))
        return m @ m.T / n + np.eye(n)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.geomspace(1.0, 1.0 / cond, n)
    a = (q * lam) @ q.T
    return (a + a.T) / 2


def solve(a_csr, b, ws, dev):
    bt = torch.from_numpy(b).to(dev)
    xt = torch.empty_like(bt)
    rep = capi.dense_cholesky_solve(a_csr, bt, xt, ws)
    return xt.cpu().numpy(), rep


@pytest.mark.parametrize("n", [1, 5, 127, 128, 129, 256, 257, 300, 511, 700])
def test_random_spd_matches_lapack(dev, ws, n):
    a = spd(n, seed=n)
    b = np.random.default_rng(100 + n).standard_normal(n)
    for compact in (False, True):
        csr = dense_to_csr(a, dev)
        if compact:
            csr = csr.compact()
        x, rep = solve(csr, b, ws, dev)
        assert not rep.not_positive_definite
        ref = np.linalg.solve(a, b)
        # cond(a) <= ~5 here: forward error a few ulps times n
        assert np.linalg.norm(x - ref) <= 1e-13 * np.sqrt(n) * np.linalg.norm(ref) + 1e-300
        assert rep.rhs_norm == pytest.approx(np.linalg.norm(b), rel=1e-14)
        r = np.linalg.norm(b - a @ x)
        assert rep.residual_norm == pytest.approx(r, rel=1e-6, abs=1e-15 * np.linalg.norm(b))
        assert rep.residual_norm <= 1e-13 * np.sqrt(n) * np.linalg.norm(b)


def test_ill_conditioned(dev, ws):
    n = 400
    a = spd(n, seed=3, cond=1e8)
    b = np.random.default_rng(4).standard_normal(n)
    x, rep = solve(dense_to_csr(a, dev), b, ws, dev)
    assert not rep.not_positive_definite
    ref = np.linalg.solve(a, b)
    # backward stable: forward error <= c * eps * cond
    assert np.linalg.norm(x - ref) <= 1e-6 * np.linalg.norm(ref)
    assert rep.residual_norm <= 1e-11 * np.linalg.norm(a, 2) * np.linalg.norm(x)


def test_lower_triangle_and_duplicates(dev, ws):
    # Only the lower triangle is read (the reference's dense-lower storage), and
    # duplicate CSR entries sum, as they do in SpMV.
    n = 150
    a = spd(n, seed=9)
    low = np.tril(a)
    rows, cols = np.nonzero(low)
    vals = low[rows, cols]
    # split every diagonal entry into two duplicates; fill the upper triangle with junk
    rng = np.random.default_rng(10)
    entries = []
    for r, c, v in zip(rows, cols, vals):
        if r == c:
            h = 0.25 * v
            entries += [(r, c, h), (r, c, v - h)]
        else:
            entries.append((r, c, v))
            entries.append((c, r, rng.standard_normal()))
    entries.sort(key=lambda e: (e[0], e[1]))
    er = np.array([e[0] for e in entries], dtype=np.int64)
    ec = np.array([e[1] for e in entries], dtype=np.int64)
    ev = np.array([e[2] for e in entries], dtype=np.float64)
    ro = np.zeros(n + 1, dtype=np.int64)
    np.add.at(ro, er + 1, 1)
    ro = np.cumsum(ro)
    csr = capi.DeviceCSR(n, n, torch.from_numpy(ro).to(dev), torch.from_numpy(ec).to(dev),
                         torch.from_numpy(ev).to(dev))
    b = rng.standard_normal(n)
    bt = torch.from_numpy(b).to(dev)
    xt = torch.empty_like(bt)
    rep = capi.dense_cholesky_solve(csr, bt, xt, ws)
    assert not rep.not_positive_definite
    ref = np.linalg.solve(a, b)
    assert np.linalg.norm(xt.cpu().numpy() - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("n,where", [(3, 1), (300, 200), (700, 650)])
def test_indefinite_is_reported(dev, ws, n, where):
    rng = np.random.default_rng(n)
    if n == 3:
        a = np.diag([1.0, -1.0, 2.0])
    else:
        # the leading minors stay positive definite up to `where`
        a = spd(n, seed=n)
        a[where, where] = -50.0
    x, rep = solve(dense_to_csr(a, dev), rng.standard_normal(n), ws, dev)
    assert rep.not_positive_definite


def test_repeated_solves_are_bitwise_equal(dev, ws):
    n = 700
    a = spd(n, seed=21)
    b = np.random.default_rng(22).standard_normal(n)
    csr = dense_to_csr(a, dev)
    x1, _ = solve(csr, b, ws, dev)
    x2, _ = solve(csr, b, ws, dev)
    assert np.array_equal(x1, x2)


def test_errors(dev, ws):
    a = spd(4, seed=1)
    csr = dense_to_csr(a, dev)
    b = torch.ones(4, dtype=torch.float64, device=dev)
    with pytest.raises(capi.PlamatrixError):
        capi.dense_cholesky_solve(csr, b, b, ws)  # aliasing
    f32 = capi.DeviceCSR(4, 4, csr.row_offsets, csr.col_indices, csr.values.float())
    with pytest.raises(capi.PlamatrixError):
        capi.dense_cholesky_solve(f32, b, torch.empty_like(b), ws)
    rect = capi.DeviceCSR(4, 5, csr.row_offsets, csr.col_indices, csr.values)
    with pytest.raises(capi.PlamatrixError):
        capi.dense_cholesky_solve(rect, b, torch.empty_like(b), ws)
    big = 70000
    empty = capi.DeviceCSR(big, big, torch.zeros(big + 1, dtype=torch.int64, device=dev),
                           torch.zeros(1, dtype=torch.int64, device=dev), torch.zeros(1, dtype=torch.float64, device=dev))
    zb = torch.zeros(big, dtype=torch.float64, device=dev)
    with pytest.raises(capi.PlamatrixError, match="dense limit"):
        capi.dense_cholesky_solve(empty, zb, torch.empty_like(zb), ws)


def schur_instance(dev, sw, pc, ec, seed):
    inst = generators.lm_schur_instance(primary_count=pc, eliminated_count=ec, window=(3, 12), seed=seed)

    def t(v, dt):
        return torch.from_numpy(np.ascontiguousarray(np.asarray(v, dtype=dt))).to(dev)

    ne = capi.assemble_normal_equations(sw, pc, ec, 6, 3, primary=t(inst.primary, np.int64),
                                        eliminated=t(inst.eliminated, np.int64),
                                        primary_jacobians=t(inst.primary_jacobians, np.float64),
                                        eliminated_jacobians=t(inst.eliminated_jacobians, np.float64),
                                        residuals=t(inst.residuals, np.float64), weights=t(inst.weights, np.float64))
    return inst, ne


def reduced_system(sw):
    import scipy.sparse as sp

    a, rhs_ptr, _ = capi.schur_system(sw)
    n = a.rows
    ro = capi.device_array(a.row_offsets, n + 1, "int64")
    ci = capi.device_array(a.col_indices, a.nnz, "int64")
    vals = capi.device_array(a.values, a.nnz)
    return sp.csr_matrix((vals, ci, ro), shape=(n, n)), capi.device_array(rhs_ptr, n)


def test_schur_dense_backend_matches_lapack_and_pcg(dev, ws):
    pc, ec = 40, 2000
    sw = capi.SchurWorkspace()
    inst, ne = schur_instance(dev, sw, pc, ec, seed=3)
    dp = torch.zeros(pc * 6, dtype=torch.float64, device=dev)
    de = torch.zeros(ec * 3, dtype=torch.float64, device=dev)
    rep = capi.schur_solve(sw, ne, inst.damping, dp, de, ws, relative_tolerance=1e-12, absolute_tolerance=0.0,
                           dense=True)
    assert rep.converged and rep.iterations == 1 and rep.message == "converged", rep.message
    S, rhs = reduced_system(sw)
    ref = np.linalg.solve(S.toarray(), rhs)
    assert np.linalg.norm(dp.cpu().numpy() - ref) <= 1e-10 * np.linalg.norm(ref)
    assert rep.final_residual_norm <= 1e-12 * rep.initial_residual_norm
    # the PCG backend at a tight tolerance agrees, back-substitution included
    dp2 = torch.zeros_like(dp)
    de2 = torch.zeros_like(de)
    rep2 = capi.schur_solve(sw, ne, inst.damping, dp2, de2, ws, max_iterations=2000, relative_tolerance=1e-12,
                            absolute_tolerance=0.0)
    assert rep2.converged
    got = np.concatenate([dp.cpu().numpy(), de.cpu().numpy()])
    other = np.concatenate([dp2.cpu().numpy(), de2.cpu().numpy()])
    assert np.linalg.norm(got - other) <= 1e-8 * np.linalg.norm(other)


def test_schur_dense_matches_reference_dense_cpu(dev, ws, ref_lib):
    # the reference's own DenseCpu backend (oracle/_ref, built from its sources)
    # on the same residual batch and weights
    pc, ec = 40, 2000
    sw = capi.SchurWorkspace()
    inst, ne = schur_instance(dev, sw, pc, ec, seed=5)
    dp = torch.zeros(pc * 6, dtype=torch.float64, device=dev)
    de = torch.zeros(ec * 3, dtype=torch.float64, device=dev)
    rep = capi.schur_solve(sw, ne, inst.damping, dp, de, ws, relative_tolerance=1e-12, absolute_tolerance=0.0,
                           dense=True)
    rdp, rde, rrep = ref_lib.ref_schur_solve_batch(pc, ec, 6, 3, inst.primary, inst.eliminated,
                                                   inst.primary_jacobians, inst.eliminated_jacobians, inst.residuals,
                                                   inst.weights, inst.damping, backend=1, relative_tolerance=1e-12,
                                                   absolute_tolerance=0.0)
    assert rep.converged and rrep.converged and rrep.iterations == rep.iterations == 1
    got = np.concatenate([dp.cpu().numpy(), de.cpu().numpy()])
    ref = np.concatenate([rdp, rde])
    assert np.linalg.norm(got - ref) <= 1e-10 * np.linalg.norm(ref)
    assert rep.initial_residual_norm == pytest.approx(rrep.initial_residual, rel=1e-12)


def test_schur_dense_full_size(dev, ws):
    # 2,000 primaries x 6: a 12,000 x 12,000 dense factorization
    inst = generators.lm_schur_instance()
    sw = capi.SchurWorkspace()

    def t(v, dt):
        return torch.from_numpy(np.ascontiguousarray(np.asarray(v, dtype=dt))).to(dev)

    ne = capi.assemble_normal_equations(sw, 2000, 200000, 6, 3, primary=t(inst.primary, np.int64),
                                        eliminated=t(inst.eliminated, np.int64),
                                        primary_jacobians=t(inst.primary_jacobians, np.float64),
                                        eliminated_jacobians=t(inst.eliminated_jacobians, np.float64),
                                        residuals=t(inst.residuals, np.float64), weights=None,
                                        huber_delta=inst.huber_delta)
    dp = torch.zeros(12000, dtype=torch.float64, device=dev)
    de = torch.zeros(600000, dtype=torch.float64, device=dev)
    rep = capi.schur_solve(sw, ne, inst.damping, dp, de, ws, relative_tolerance=1e-10, absolute_tolerance=0.0,
                           dense=True)
    assert rep.converged, rep.message
    S, rhs = reduced_system(sw)
    x = dp.cpu().numpy()
    assert np.linalg.norm(rhs - S @ x) <= 1e-10 * np.linalg.norm(rhs)
    assert rep.final_residual_norm == pytest.approx(np.linalg.norm(rhs - S @ x), rel=1e-3, abs=1e-14 * np.linalg.norm(rhs))
