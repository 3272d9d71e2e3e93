"""GPU parity of the diagonal (DIA) layout the iterative solvers use for large
stencil matrices (plamatrix_amd/csrc/dia.hip).

The layout adds each row's present entries in increasing column order with the
same fused multiply-adds as the CSR row walk, so a solve with it must be
bitwise the solve without it (PLAMATRIX_AMD_DIA=0): same iteration count, same
residual, same solution bits.  That the layout was really used is read from
the workspace's storage, which grows by the diagonal values.  Matrices the
layout does not fit (unsorted rows, too many diagonals) fall back to CSR.
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


def host(x):
    torch.cuda.synchronize()
    return x.detach().cpu().numpy()


def from_scipy(m, dtype):
    m = m.tocsr()
    m.sort_indices()
    return generators.CSRHost(m.shape[0], m.shape[1], m.indptr.astype(np.int64), m.indices.astype(np.int64),
                              m.data.astype(dtype))


def nine_point(nx):
    import scipy.sparse as sp

    t = sp.diags([1.0, 1.0, 1.0], [-1, 0, 1], shape=(nx, nx))
    return 9.0 * sp.identity(nx * nx) - sp.kron(t, t)


def solve_both(dev, h, dtype, monkeypatch, rel, refresh=50):
    a = capi.to_device_csr(h, compact=True)
    b = torch.from_numpy(generators.rhs_normal(h.rows, seed=31).astype(dtype)).to(dev)
    out = {}
    grow = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("PLAMATRIX_AMD_DIA", flag)
        ws = capi.Workspace()
        x = torch.zeros(h.rows, dtype=b.dtype, device=dev)
        before = ws.capacity_bytes
        rep = capi.pcg(a, b, x, ws, capi.policy(refresh_interval=refresh), relative_tolerance=rel,
                       max_iterations=20000)
        torch.cuda.synchronize()
        grow[flag] = ws.capacity_bytes - before
        grow["diagonals" + flag] = ws.solver_diagonals
        out[flag] = (rep, host(x).copy())
    return out, grow


@pytest.mark.parametrize("fuse", ["0", "1"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_diffusion3d_dia_is_bitwise_csr(dev, monkeypatch, dtype, fuse):
    # fuse=1: the direction p = z + beta p formed inside the diagonal SpMV
    monkeypatch.setenv("PLAMATRIX_AMD_FUSE_DIRECTION", fuse)
    h = generators.diffusion3d_csr(48, seed=7, dtype=dtype)  # 110,592 rows, 7 diagonals
    out, grow = solve_both(dev, h, dtype, monkeypatch, 1e-6 if dtype == np.float32 else 1e-10)
    es = np.dtype(dtype).itemsize
    assert grow["1"] - grow["0"] >= 7 * h.rows * es  # the diagonal values were built
    (r0, x0), (r1, x1) = out["0"], out["1"]
    assert r0.converged and r1.converged
    assert r0.iterations == r1.iterations
    assert r0.final_residual == r1.final_residual
    assert np.array_equal(x0, x1)


def test_laplacian2d_five_diagonals(dev, monkeypatch):
    h = generators.laplacian2d_csr(400, dtype=np.float32)  # 160,000 rows
    out, grow = solve_both(dev, h, np.float32, monkeypatch, 1e-6)
    assert grow["1"] - grow["0"] >= 5 * h.rows * 4
    assert out["0"][0].iterations == out["1"][0].iterations
    assert np.array_equal(out["0"][1], out["1"][1])


def test_nine_point_generic_path(dev, monkeypatch):
    h = from_scipy(nine_point(320), np.float64)  # 102,400 rows, 9 diagonals (runtime diagonal count)
    out, grow = solve_both(dev, h, np.float64, monkeypatch, 1e-10)
    assert grow["1"] - grow["0"] >= 9 * h.rows * 8
    assert out["0"][0].iterations == out["1"][0].iterations
    assert np.array_equal(out["0"][1], out["1"][1])


def test_unsorted_rows_fall_back(dev, monkeypatch):
    h = generators.diffusion3d_csr(48, seed=9, dtype=np.float64)
    # reverse the entries of every row: same matrix, columns decreasing
    ro, ci, v = h.row_offsets, h.col_indices.copy(), h.values.copy()
    for r in range(h.rows):
        s, e = ro[r], ro[r + 1]
        ci[s:e] = ci[s:e][::-1]
        v[s:e] = v[s:e][::-1]
    rev = generators.CSRHost(h.rows, h.cols, ro, ci, v)
    out, grow = solve_both(dev, rev, np.float64, monkeypatch, 1e-10)
    assert grow["diagonals1"] == 0  # the CSR kernels ran
    assert out["1"][0].converged and np.array_equal(out["0"][1], out["1"][1])


def test_refinement_with_dia_reaches_target(dev, monkeypatch):
    monkeypatch.setenv("PLAMATRIX_AMD_DIA", "1")
    h = generators.diffusion3d_csr(64, seed=11)  # FP32 matrix, 262,144 rows
    a = capi.to_device_csr(h, compact=True)
    b = generators.rhs_normal(h.rows, seed=12).astype(np.float64)
    x = torch.zeros(h.rows, dtype=torch.float64, device=dev)
    ws = capi.Workspace()
    rep = capi.pcg_refine(a, torch.from_numpy(b).to(dev), x, ws, target=1e-8, max_outer=20,
                          pol=capi.policy(refresh_interval=50), relative_tolerance=1e-3, max_iterations=20000)
    assert rep.converged and rep.final_relres <= 1e-8
    import scipy.sparse as sp

    m = sp.csr_matrix((h.values.astype(np.float64), h.col_indices, h.row_offsets), shape=(h.rows, h.cols))
    xs = host(x)
    assert np.linalg.norm(b - m @ xs) <= 1.2e-8 * np.linalg.norm(b)


def test_learned_factors_on_the_diagonal_layout(dev, monkeypatch):
    # the factored preconditioner's G and G^T (4 diagonals each on a 3-D
    # 7-point pattern) take the diagonal layout too; the classic factored
    # iteration stays bitwise (the single-reduction form needs all three
    # matrices on the layout and has its own tests)
    monkeypatch.setenv("PLAMATRIX_AMD_SINGLE_REDUCTION", "0")
    h = generators.diffusion3d_box_csr(48, 48, 48, seed=13, dtype=np.float32)
    a = capi.to_device_csr(h, compact=True)
    net = capi.FsaiNet(seed=0)
    graph = capi.FsaiGraph(a)
    capi.fsai_forward(net, graph)
    b = torch.from_numpy(generators.rhs_normal(h.rows, seed=14)).to(dev)
    out, grow = {}, {}
    for flag in ("0", "1"):
        monkeypatch.setenv("PLAMATRIX_AMD_DIA", flag)
        ws = capi.Workspace()
        x = torch.zeros_like(b)
        rep = capi.pcg_factored(a, b, x, graph, ws, relative_tolerance=1e-6, max_iterations=5000)
        torch.cuda.synchronize()
        grow[flag] = ws.capacity_bytes
        out[flag] = (rep, host(x).copy())
        assert ws.solver_form == 3
    # A (7 diagonals) plus G and G^T (4 each)
    assert grow["1"] - grow["0"] >= (7 + 4 + 4) * h.rows * 4
    assert out["0"][0].converged and out["0"][0].iterations == out["1"][0].iterations
    assert np.array_equal(out["0"][1], out["1"][1])


def jacobi_outcome(dev, a, b, monkeypatch, unit_flag, rel, kernels=None):
    monkeypatch.setenv("PLAMATRIX_AMD_UNIT_DIAGONAL", unit_flag)
    ws = capi.Workspace()
    x = torch.zeros_like(b)
    try:
        rep = capi.pcg(a, b, x, ws, capi.policy(refresh_interval=50), relative_tolerance=rel, max_iterations=20000)
    except Exception as e:  # the Jacobi diagonal rules reject a row
        torch.cuda.synchronize()
        return ("error", type(e).__name__, str(e))
    torch.cuda.synchronize()
    if kernels is not None:
        assert ws.solver_preconditioner == kernels
    return (rep.converged, rep.iterations, rep.final_residual, host(x).tobytes())


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_unit_diagonal_jacobi_is_bitwise(dev, monkeypatch, dtype):
    # the diffusion matrix is scaled to a unit diagonal: Jacobi's z = 1 * r is
    # r, so the solve runs without the inverse diagonal and must keep its bits
    h = generators.diffusion3d_csr(48, seed=23, dtype=dtype)
    a = capi.to_device_csr(h, compact=True)
    b = torch.from_numpy(generators.rhs_normal(h.rows, seed=24).astype(dtype)).to(dev)
    rel = 1e-6 if dtype == np.float32 else 1e-10
    ref = jacobi_outcome(dev, a, b, monkeypatch, "0", rel, kernels="jacobi")
    got = jacobi_outcome(dev, a, b, monkeypatch, "1", rel, kernels="none")
    assert ref[0] is True and ref == got


@pytest.mark.parametrize("change", ["diagonal", "huge_off_diagonal"])
def test_unit_diagonal_check_is_exact(dev, monkeypatch, change):
    # one row breaks the unit-inverse property: a diagonal of 2, or (FP32) an
    # off-diagonal so large that the Jacobi rules reject the row's diagonal;
    # the outcome must be the Jacobi kernels' either way
    h = generators.diffusion3d_csr(48, seed=25, dtype=np.float32)
    v = h.values.copy()
    row = h.rows // 2
    s, e = h.row_offsets[row], h.row_offsets[row + 1]
    cols = h.col_indices[s:e]
    if change == "diagonal":
        v[s + int(np.nonzero(cols == row)[0][0])] = 2.0
    else:
        v[s + int(np.nonzero(cols != row)[0][0])] = -1e8
    m = generators.CSRHost(h.rows, h.cols, h.row_offsets, h.col_indices, v)
    a = capi.to_device_csr(m, compact=True)
    b = torch.from_numpy(generators.rhs_normal(h.rows, seed=26)).to(dev)
    ref = jacobi_outcome(dev, a, b, monkeypatch, "0", 1e-6)
    got = jacobi_outcome(dev, a, b, monkeypatch, "1", 1e-6, kernels="jacobi" if change == "diagonal" else None)
    assert ref == got
    if change == "huge_off_diagonal":
        assert ref[0] == "error" and "diagonal" in ref[2]


def layout_outcome(dev, h, monkeypatch, dia, mirror, rel=1e-12, iterations=40, unit="1"):
    # a fixed number of iterations (the tolerance is out of reach), so
    # matrices that are not symmetric compare too
    monkeypatch.setenv("PLAMATRIX_AMD_DIA", dia)
    monkeypatch.setenv("PLAMATRIX_AMD_DIA_MIRROR", mirror)
    monkeypatch.setenv("PLAMATRIX_AMD_UNIT_DIAGONAL", unit)
    a = capi.to_device_csr(h, compact=True)
    b = torch.from_numpy(generators.rhs_normal(h.rows, seed=41).astype(h.values.dtype)).to(dev)
    ws = capi.Workspace()
    x = torch.zeros_like(b)
    rep = capi.pcg(a, b, x, ws, capi.policy(refresh_interval=50), relative_tolerance=rel, max_iterations=iterations)
    torch.cuda.synchronize()
    return (rep.iterations, rep.final_residual, host(x).tobytes()), ws.solver_diagonals, ws.solver_value_streams


@pytest.mark.parametrize("kind,diagonals,streams", [
    ("diffusion3d", 7, 3),   # unit diagonal: 3 upper diagonals only
    ("laplacian2d", 5, 3),   # 2 upper diagonals and the main diagonal (4)
    ("nine_point", 9, 5),    # 4 upper diagonals and the main diagonal (8)
])
def test_mirrored_layout_is_bitwise(dev, monkeypatch, kind, diagonals, streams):
    # a symmetric matrix reads its lower diagonals from their upper mirrors
    # and an all-ones main diagonal from no array; the products keep their bits
    if kind == "diffusion3d":
        h = generators.diffusion3d_csr(48, seed=43, dtype=np.float32)
    elif kind == "laplacian2d":
        h = generators.laplacian2d_csr(400, dtype=np.float32)
    else:
        h = from_scipy(nine_point(320), np.float64)
    csr, d0, s0 = layout_outcome(dev, h, monkeypatch, "0", "1")
    full, d1, s1 = layout_outcome(dev, h, monkeypatch, "1", "0")
    mirrored, d2, s2 = layout_outcome(dev, h, monkeypatch, "1", "1")
    assert (d0, s0) == (0, 0)
    assert (d1, s1) == (diagonals, diagonals)
    assert (d2, s2) == (diagonals, streams)
    assert csr == full == mirrored


@pytest.mark.parametrize("change", ["one_ulp", "missing_mirror", "diagonal_not_one"])
def test_mirror_check_is_exact(dev, monkeypatch, change):
    # one entry breaks the property: an off-diagonal value one ulp away from
    # its mirror, a lower entry whose mirror is absent, or a main-diagonal
    # entry that is not exactly 1; the layout must then read that array
    h = generators.diffusion3d_csr(48, seed=45, dtype=np.float32)
    ro, ci, v = h.row_offsets.copy(), h.col_indices.copy(), h.values.copy()
    row = h.rows // 2 + 7
    s, e = ro[row], ro[row + 1]
    cols = ci[s:e]
    if change == "one_ulp":
        k = s + int(np.nonzero(cols == row - 1)[0][0])
        v[k] = np.nextafter(v[k], np.float32(0))
        expect = 6  # every diagonal but the main one
    elif change == "missing_mirror":
        k = s + int(np.nonzero(cols == row - 1)[0][0])
        ci = np.delete(ci, k)
        v = np.delete(v, k)
        ro[row + 1:] -= 1
        expect = 6
    else:
        k = s + int(np.nonzero(cols == row)[0][0])
        v[k] = np.float32(1.0) + np.float32(2.0 ** -23)
        expect = 4  # the mirrors hold, the main diagonal is read
    m = generators.CSRHost(h.rows, h.cols, ro, ci, v)
    # unit="0": the Jacobi kernels run throughout, whatever the diagonal holds
    csr, _, _ = layout_outcome(dev, m, monkeypatch, "0", "1", unit="0")
    got, d, streams = layout_outcome(dev, m, monkeypatch, "1", "1", unit="0")
    assert d == 7 and streams == expect
    assert csr == got


@pytest.mark.parametrize("kind", ["diffusion3d_unit", "laplacian2d_jacobi", "nine_point_fp64"])
def test_one_kernel_iteration_is_bitwise(dev, monkeypatch, kind):
    # on the diagonal layout the single-reduction iteration runs as one kernel
    # (neighbours' new u recomputed from the old r, w, s of those rows, the
    # vectors alternating between two buffers); it must keep the two-kernel
    # form's bits, residual refreshes every 50 iterations included
    if kind == "diffusion3d_unit":
        h = generators.diffusion3d_csr(48, seed=47, dtype=np.float32)  # Jacobi runs as z = r
        rel = 1e-6
    elif kind == "laplacian2d_jacobi":
        h = generators.laplacian2d_csr(400, dtype=np.float32)  # diagonal 4: the Jacobi kernels
        rel = 1e-5
    else:
        h = from_scipy(nine_point(320), np.float64)
        rel = 1e-10
    a = capi.to_device_csr(h, compact=True)
    b = torch.from_numpy(generators.rhs_normal(h.rows, seed=48).astype(h.values.dtype)).to(dev)
    out = {}
    for fused in ("0", "1"):
        monkeypatch.setenv("PLAMATRIX_AMD_CG_FUSED", fused)
        ws = capi.Workspace()
        x = torch.zeros_like(b)
        rep = capi.pcg(a, b, x, ws, capi.policy(refresh_interval=50), relative_tolerance=rel, max_iterations=20000)
        torch.cuda.synchronize()
        assert ws.solver_diagonals > 0
        out[fused] = (rep.converged, rep.iterations, rep.final_residual, host(x).tobytes())
    assert out["0"][0] and out["0"][1] > 50
    assert out["0"] == out["1"]


def test_layout_from_64bit_indices(dev, monkeypatch):
    # a CSR with only 64-bit index arrays (the output of cooToCsr) takes the
    # layout and the one-kernel iteration too, with the CSR kernels' bits
    h = generators.laplacian2d_csr(300, dtype=np.float64)  # 90,000 rows, diagonal 4: Jacobi kernels
    a = capi.to_device_csr(h, compact=False)
    assert a.row_offsets32 is None
    b = torch.from_numpy(generators.rhs_normal(h.rows, seed=51).astype(np.float64)).to(dev)
    out = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("PLAMATRIX_AMD_DIA", flag)
        ws = capi.Workspace()
        x = torch.zeros_like(b)
        rep = capi.pcg(a, b, x, ws, relative_tolerance=1e-8, max_iterations=4000)
        torch.cuda.synchronize()
        out[flag] = (rep.converged, rep.iterations, rep.final_residual, host(x).tobytes())
        if flag == "1":
            assert ws.solver_diagonals == 5 and ws.solver_value_streams == 3 and ws.solver_form == 1
    assert out["0"][0] and out["0"] == out["1"]


def outcome(fn):
    try:
        rep, x = fn()
        return ("ok", rep.converged, rep.iterations, rep.final_residual, host(x).tobytes())
    except capi.PlamatrixError as e:
        torch.cuda.synchronize()
        return ("error", e.kind, str(e))


@pytest.mark.parametrize("shift", [3.9, 3.0])
def test_restart_hands_over_identically(dev, monkeypatch, shift):
    # an indefinite matrix (2-D Laplacian minus shift * I) makes p.Ap
    # non-positive: the single-reduction iteration restarts at iteration 1 and
    # after its chunks in flight the solve continues in the three-kernel form
    # from the vectors the one-kernel iteration left in its first buffers,
    # which reports the breakdown (at iteration 21; the three-kernel form alone
    # stops at 2); one and two kernels per iteration must end identically
    import scipy.sparse as sp

    h = generators.laplacian2d_csr(300, dtype=np.float64)
    m = sp.csr_matrix((h.values, h.col_indices, h.row_offsets), shape=(h.rows, h.rows)) - shift * sp.identity(h.rows)
    a = capi.to_device_csr(from_scipy(m, np.float64), compact=True)
    b = torch.from_numpy(generators.rhs_normal(h.rows, seed=53).astype(np.float64)).to(dev)
    got = {}
    for fused in ("0", "1"):
        monkeypatch.setenv("PLAMATRIX_AMD_CG_FUSED", fused)

        def run():
            ws = capi.Workspace()
            x = torch.zeros_like(b)
            rep = capi.pcg(a, b, x, ws, capi.policy(refresh_interval=50), relative_tolerance=1e-8,
                           max_iterations=3000)
            torch.cuda.synchronize()
            return rep, x

        got[fused] = outcome(run)
    assert got["0"] == got["1"]
