"""GPU parity of COO->CSR, SpMV and CG / PCG / block-PCG through the C ABI.

Every check compares the HIP path with the oracle (oracle/) on the same
seeded inputs, or with the reference's own known answers
(sparse_ops_test.cpp, iterative_solver_test.cpp).  Bars: bit-exact for
COO->CSR (integer structure and folded values); SpMV to FP32/FP64 rounding;
PCG solutions within relative 1e-6 (FP64) / 1e-4 (FP32) of the CPU solve, with
the FP64 true residual recomputed for every result.
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
    del w


def t(a, dev, dtype=None):
    arr = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype=dtype))
    return torch.from_numpy(arr).to(dev)


def host(x):
    torch.cuda.synchronize()
    return x.detach().cpu().numpy()


def kinds(e):
    return e.value.kind


# ---- COO -> CSR -------------------------------------------------------------

@pytest.mark.parametrize("dtype,large", [(np.float32, 1e8), (np.float64, 1e16)])
def test_coo_kat_cancellation(dev, ws, oracle, dtype, large):
    # sparse_ops_test.cpp:320-369
    rows = [2, 0, 2, 1, 2, 0, 2]
    cols = [4, 3, 1, 2, 1, 3, 1]
    vals = np.array([5, 2, large, 4, -large, -2, 1], dtype=dtype)
    r, c, v = t(rows, dev, np.int64), t(cols, dev, np.int64), t(vals, dev)
    csr = capi.coo_to_csr(4, 5, r, c, v, ws)
    assert list(host(csr.row_offsets)) == [0, 1, 2, 4, 4]
    assert list(host(csr.col_indices)) == [3, 2, 1, 4]
    assert list(host(csr.values)) == [0, 4, 1, 5]
    # triplets untouched
    assert list(host(r)) == rows and list(host(c)) == cols and np.array_equal(host(v), vals)


def test_coo_async_reuses_output(dev, ws):
    # sparse_ops_test.cpp:371-430
    r = t([1, 0, 1, 2, 1], dev, np.int64)
    c = t([0, 2, 0, 1, 0], dev, np.int64)
    v = t([2, 3, 4, 5, -1], dev, np.float64)
    out = capi.DeviceCSR(3, 3, torch.empty(4, dtype=torch.int64, device=dev),
                         torch.empty(3, dtype=torch.int64, device=dev), torch.empty(3, dtype=torch.float64, device=dev))
    for _ in range(2):
        capi.coo_to_csr_async(3, 3, r, c, v, out, ws)
        torch.cuda.synchronize()
        ws.check_status()
    assert list(host(out.row_offsets)) == [0, 1, 2, 3]
    assert list(host(out.col_indices)) == [2, 0, 1]
    assert list(host(out.values)) == [3, 5, 5]


def test_coo_deferred_errors(dev, ws):
    # sparse_ops_test.cpp:468-497: nnz mismatch -> runtime_error,
    # out-of-range coordinate -> out_of_range, both deferred to checkStatus
    r = t([0, 0], dev, np.int64)
    c = t([1, 1], dev, np.int64)
    v = t([2.0, 3.0], dev, np.float32)
    wrong = capi.DeviceCSR(2, 2, torch.empty(3, dtype=torch.int64, device=dev),
                           torch.empty(2, dtype=torch.int64, device=dev), torch.empty(2, dtype=torch.float32, device=dev))
    capi.coo_to_csr_async(2, 2, r, c, v, wrong, ws)
    torch.cuda.synchronize()
    with pytest.raises(capi.PlamatrixError) as e:
        ws.check_status()
    assert kinds(e) == "runtime_error"
    bad_row = t([0, 2], dev, np.int64)
    one = capi.DeviceCSR(2, 2, torch.empty(3, dtype=torch.int64, device=dev),
                         torch.empty(1, dtype=torch.int64, device=dev), torch.empty(1, dtype=torch.float32, device=dev))
    capi.coo_to_csr_async(2, 2, bad_row, c, v, one, ws)
    torch.cuda.synchronize()
    with pytest.raises(capi.PlamatrixError) as e:
        ws.check_status()
    assert kinds(e) == "out_of_range" and "source offset 1" in str(e.value)


def test_coo_rejects_non_finite(dev, ws):
    # sparse_ops_test.cpp:499-522 (input and combined value)
    r = t([0, 0], dev, np.int64)
    v = t([np.inf, 1.0], dev, np.float64)
    with pytest.raises(capi.PlamatrixError) as e:
        capi.coo_to_csr(1, 1, r, r, v, ws)
    assert kinds(e) == "invalid_argument"
    big = np.finfo(np.float64).max
    with pytest.raises(capi.PlamatrixError) as e:
        capi.coo_to_csr(1, 1, r, r, t([big, big], dev, np.float64), ws)
    assert kinds(e) == "invalid_argument"


def test_coo_empty(dev, ws):
    # sparse_ops_test.cpp:524-545
    e64 = torch.empty(0, dtype=torch.int64, device=dev)
    csr = capi.coo_to_csr(3, 4, e64, e64, torch.empty(0, dtype=torch.float32, device=dev), ws)
    assert list(host(csr.row_offsets)) == [0, 0, 0, 0]
    assert csr.nnz == 0


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_coo_poisson_shuffled_duplicates_bit_exact(dev, ws, oracle, dtype):
    coo = generators.poisson2d_coo(256, dtype=dtype)
    ref = oracle.coo_to_csr(coo.rows, coo.cols, coo.row_indices, coo.col_indices, coo.values)
    csr = capi.coo_to_csr(coo.rows, coo.cols, t(coo.row_indices, dev), t(coo.col_indices, dev), t(coo.values, dev), ws)
    assert csr.nnz == ref.nnz == 326656
    assert np.array_equal(host(csr.row_offsets), ref.row_offsets)
    assert np.array_equal(host(csr.col_indices), ref.col_indices)
    assert np.array_equal(host(csr.values).view(np.uint8), ref.values.view(np.uint8))


def test_coo_random_heavy_duplicates_bit_exact(dev, ws, oracle):
    rng = np.random.default_rng(77)
    for rows, cols, count in ((1000, 3000, 200000), (5, 7, 5000), (70000, 1, 300000), (1, 1 << 20, 100000)):
        r = rng.integers(0, rows, count)
        c = rng.integers(0, cols, count)
        v = rng.standard_normal(count) * 10.0 ** rng.integers(-12, 13, count)
        ref = oracle.coo_to_csr(rows, cols, r, c, v)
        csr = capi.coo_to_csr(rows, cols, t(r, dev), t(c, dev), t(v, dev), ws, check_finite=False)
        assert np.array_equal(host(csr.row_offsets), ref.row_offsets)
        assert np.array_equal(host(csr.col_indices), ref.col_indices)
        assert np.array_equal(host(csr.values).view(np.uint8), ref.values.view(np.uint8)), (rows, cols, count)


def test_coo_negative_zero_matches_cpu_fold(dev, ws, oracle):
    v = np.array([-0.0, 1.0, -1.0, -0.0])
    r = np.array([0, 1, 1, 2])
    c = np.array([0, 0, 0, 0])
    ref = oracle.coo_to_csr(3, 1, r, c, v)
    csr = capi.coo_to_csr(3, 1, t(r, dev), t(c, dev), t(v, dev), ws)
    assert np.array_equal(host(csr.values).view(np.uint64), ref.values.view(np.uint64))


# ---- SpMV ---------------------------------------------------------------------

@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_spmv_kat(dev, ws, oracle, dtype):
    # sparse_ops_test.cpp:636-670
    ref = oracle.coo_to_csr(3, 4, [2, 0, 1, 0], [3, 2, 1, 0], np.array([5, 2, -1, 3], dtype))
    a = capi.to_device_csr(ref)
    x = t(np.array([1, 2, 3, 4], dtype), dev)
    y = torch.full((3,), -100, dtype=x.dtype, device=dev)
    capi.spmv(a, x, y, ws)
    assert list(host(y)) == [9, -2, 20]
    with pytest.raises(capi.PlamatrixError) as e:
        capi.spmv(a, x, x, ws)
    assert kinds(e) == "invalid_argument"


@pytest.mark.parametrize("lanes", [0, 1, 2, 4, 8, 16, 32, 64])
@pytest.mark.parametrize("compact", [False, True])
def test_spmv_random_all_lane_widths(dev, ws, oracle, lanes, compact):
    rng = np.random.default_rng(lanes)
    rows, cols = 3000, 2500
    counts = rng.integers(0, 90, rows)
    counts[::7] = 0
    r = np.repeat(np.arange(rows), counts)
    c = rng.integers(0, cols, len(r))
    v = rng.standard_normal(len(r))
    ref = oracle.coo_to_csr(rows, cols, r, c, v)
    x = rng.standard_normal(cols)
    exact = oracle.spmv(ref, x)
    a = capi.to_device_csr(ref, compact=compact)
    y = torch.empty(rows, dtype=torch.float64, device=dev)
    capi.spmv(a, t(x, dev), y, ws, lanes=lanes)
    np.testing.assert_allclose(host(y), exact, rtol=1e-12, atol=1e-12)
    ref32 = oracle.CSR(rows, cols, ref.row_offsets, ref.col_indices, ref.values.astype(np.float32))
    a32 = capi.to_device_csr(ref32, compact=compact)
    y32 = torch.empty(rows, dtype=torch.float32, device=dev)
    capi.spmv(a32, t(x.astype(np.float32), dev), y32, ws, lanes=lanes)
    exact32 = oracle.spmv(oracle.CSR(rows, cols, ref.row_offsets, ref.col_indices, ref32.values.astype(np.float64)),
                          x.astype(np.float32).astype(np.float64))
    scale = oracle.spmv(oracle.CSR(rows, cols, ref.row_offsets, ref.col_indices, np.abs(ref32.values.astype(np.float64))),
                        np.abs(x.astype(np.float32).astype(np.float64)))
    assert np.all(np.abs(host(y32) - exact32) <= 2e-6 * scale + 1e-30)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_spmv_stream_long_rows_fallback(dev, ws, oracle, dtype):
    """Rows long enough that a 64-row chunk exceeds the LDS staging capacity
    take the 8-lanes-per-row path; mixed with short rows and empty rows."""
    rng = np.random.default_rng(5)
    rows, cols = 700, 4000
    counts = rng.integers(0, 12, rows)
    counts[100:180] = 300
    counts[::11] = 0
    r = np.repeat(np.arange(rows), counts)
    c = rng.integers(0, cols, len(r))
    v = rng.standard_normal(len(r)).astype(dtype)
    ref = oracle.coo_to_csr(rows, cols, r, c, v)
    x = rng.standard_normal(cols).astype(dtype)
    exact = oracle.spmv(oracle.CSR(rows, cols, ref.row_offsets, ref.col_indices, ref.values.astype(np.float64)),
                        x.astype(np.float64))
    scale = oracle.spmv(oracle.CSR(rows, cols, ref.row_offsets, ref.col_indices,
                                   np.abs(ref.values.astype(np.float64))), np.abs(x.astype(np.float64)))
    for compact in (False, True):
        a = capi.to_device_csr(ref, compact=compact)
        y = torch.empty(rows, dtype=torch.float32 if dtype == np.float32 else torch.float64, device=dev)
        capi.spmv(a, t(x, dev), y, ws, lanes=0)
        eps = 4e-6 if dtype == np.float32 else 1e-14
        assert np.all(np.abs(host(y) - exact) <= eps * scale + 1e-30)


def test_spmv_diffusion3d_full_size_row_sums(dev, ws, oracle):
    """128^3 diffusion (14.6M nnz): y = A 1 against the oracle's FP32 SpMV and
    the FP64 row sums."""
    h = generators.diffusion3d_csr(128)
    assert h.nnz == 14581760
    a = capi.to_device_csr(h, compact=True)
    ones = torch.ones(h.rows, dtype=torch.float32, device=dev)
    y = torch.empty(h.rows, dtype=torch.float32, device=dev)
    capi.spmv(a, ones, y, ws)
    row_sums = np.add.reduceat(h.values.astype(np.float64), h.row_offsets[:-1])
    absum = np.add.reduceat(np.abs(h.values.astype(np.float64)), h.row_offsets[:-1])
    assert np.all(np.abs(host(y) - row_sums) <= 1e-6 * absum)
    x = generators.rhs_normal(h.rows, seed=9)
    capi.spmv(a, t(x, dev), y, ws)
    ref = oracle.spmv(oracle.CSR(h.rows, h.cols, h.row_offsets, h.col_indices, h.values), x)
    np.testing.assert_allclose(host(y), ref, rtol=1e-4, atol=2e-5)


def test_validate_structure(dev, ws):
    good = capi.to_device_csr(generators.laplacian2d_csr(8))
    capi.validate(good, ws)
    bad = capi.DeviceCSR(2, 2, t([0, 2, 1], dev, np.int64), t([0, 1], dev, np.int64), t([1.0, 1.0], dev))
    with pytest.raises(capi.PlamatrixError) as e:
        capi.validate(bad, ws)
    assert kinds(e) == "invalid_argument"
    badc = capi.DeviceCSR(2, 2, t([0, 1, 2], dev, np.int64), t([0, 2], dev, np.int64), t([1.0, 1.0], dev))
    with pytest.raises(capi.PlamatrixError) as e:
        capi.validate(badc, ws)
    assert "position 1" in str(e.value)


# ---- CG / PCG -------------------------------------------------------------------

def poisson1d(oracle, n, dtype=np.float64, scale=1.0):
    rows, cols, vals = [], [], []
    for r in range(n):
        if r > 0:
            rows.append(r); cols.append(r - 1); vals.append(-scale)
        rows.append(r); cols.append(r); vals.append(2 * scale)
        if r + 1 < n:
            rows.append(r); cols.append(r + 1); vals.append(-scale)
    return oracle.coo_to_csr(n, n, rows, cols, np.array(vals, dtype))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("graph", [0, 1])
def test_adaptive_matches_cpu(dev, ws, oracle, dtype, graph):
    # iterative_solver_test.cpp:347-392
    m = poisson1d(oracle, 32, dtype)
    expected = np.arange(1, 33, dtype=dtype)
    rhs = oracle.spmv(m, expected)
    rel = 1e-5 if dtype == np.float32 else 1e-12
    cpu_x, cpu = oracle.pcg(m, rhs, relative_tolerance=rel)
    a = capi.to_device_csr(m)
    x = torch.zeros(32, dtype=torch.float32 if dtype == np.float32 else torch.float64, device=dev)
    rep = capi.pcg(a, t(rhs, dev), x, ws, capi.policy(use_graph=graph), relative_tolerance=rel)
    tol = 2e-4 if dtype == np.float32 else 1e-10
    assert rep.converged
    assert abs(rep.final_residual - cpu.final_residual) <= (2e-4 if dtype == np.float32 else 1e-9)
    gx = host(x)
    for i in range(32):
        assert abs(gx[i] - cpu_x[i]) <= tol * (i + 1)
    x.zero_()
    rep2 = capi.cg(a, t(rhs, dev), x, ws, relative_tolerance=rel)
    assert rep2.converged and rep2.iterations > 0


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_block_pcg_one_iteration(dev, ws, oracle, dtype):
    # iterative_solver_test.cpp:394-424
    m = oracle.coo_to_csr(4, 4, [0, 0, 1, 1, 2, 2, 3, 3], [0, 1, 0, 1, 2, 3, 2, 3],
                          np.array([4, 1, 1, 3, 2, 0.5, 0.5, 1.5], dtype))
    inv = np.array([3 / 11, -1 / 11, -1 / 11, 4 / 11, 1.5 / 2.75, -0.5 / 2.75, -0.5 / 2.75, 2 / 2.75], dtype)
    expected = np.array([1, 2, -1, 3], dtype)
    rhs = oracle.spmv(m, expected)
    a = capi.to_device_csr(m)
    x = torch.zeros(4, dtype=torch.float32 if dtype == np.float32 else torch.float64, device=dev)
    rep = capi.block_pcg(a, t(rhs, dev), x, t(inv, dev), 2, ws,
                         relative_tolerance=1e-5 if dtype == np.float32 else 1e-12)
    assert rep.converged and rep.iterations == 1
    np.testing.assert_allclose(host(x), expected, atol=2e-4 if dtype == np.float32 else 1e-10)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_zero_iterations_and_nonconvergence(dev, ws, oracle, dtype):
    # iterative_solver_test.cpp:426-461
    m = poisson1d(oracle, 16, dtype)
    a = capi.to_device_csr(m)
    exact = np.full(16, 2, dtype)
    x = t(exact, dev)
    rep = capi.pcg(a, t(oracle.spmv(m, exact), dev), x, ws)
    assert rep.converged and rep.iterations == 0
    x.zero_()
    b = t(np.ones(16, dtype), dev)
    rep = capi.cg(a, b, x, ws, max_iterations=1, relative_tolerance=0.0)
    assert not rep.converged and rep.iterations == 1
    x.zero_()
    with pytest.raises(capi.PlamatrixError) as e:
        capi.cg(a, b, x, ws, max_iterations=1, relative_tolerance=0.0, require_convergence=True)
    assert kinds(e) == "runtime_error"


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_workspace_resize_planned_sizes(dev, oracle, dtype):
    # iterative_solver_test.cpp:463-495
    ws = capi.Workspace()
    for n in (10, 1000, 100000):
        m = oracle.CSR(n, n, np.arange(n + 1, dtype=np.int64), np.arange(n, dtype=np.int64), np.full(n, 2, dtype))
        a = capi.to_device_csr(m)
        x = torch.zeros(n, dtype=torch.float32 if dtype == np.float32 else torch.float64, device=dev)
        rep = capi.pcg(a, t(np.full(n, 2, dtype), dev), x, ws)
        assert rep.converged and rep.iterations == 1
        hx = host(x)
        assert abs(hx[0] - 1) < 2e-4 and abs(hx[-1] - 1) < 2e-4
        assert ws.capacity_rows == n


def test_disabled_jacobi(dev, ws, oracle):
    # iterative_solver_test.cpp:497-519
    m = oracle.coo_to_csr(3, 3, [0, 1, 2], [0, 1, 2], np.array([2.0, 4.0, 8.0]))
    a = capi.to_device_csr(m)
    x = torch.zeros(3, dtype=torch.float64, device=dev)
    rep = capi.pcg(a, t([2.0, 8.0, 24.0], dev), x, ws, max_iterations=1, relative_tolerance=0.0, use_jacobi=False)
    assert not rep.converged and rep.iterations == 1 and rep.final_residual > 0


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_breakdown_leaves_solution_untouched(dev, ws, oracle, dtype):
    # iterative_solver_test.cpp:521-560
    m = oracle.coo_to_csr(2, 2, [0, 1], [0, 1], np.array([-1, 1], dtype))
    a = capi.to_device_csr(m)
    x = torch.zeros(2, dtype=torch.float32 if dtype == np.float32 else torch.float64, device=dev)
    with pytest.raises(capi.PlamatrixError) as e:
        capi.cg(a, t(np.ones(2, dtype), dev), x, ws)
    assert kinds(e) == "runtime_error"
    assert list(host(x)) == [0, 0]
    tiny = np.array([np.finfo(dtype).smallest_subnormal], dtype)
    m1 = oracle.CSR(1, 1, np.array([0, 1], np.int64), np.array([0], np.int64), tiny)
    x1 = torch.zeros(1, dtype=x.dtype, device=dev)
    with pytest.raises(capi.PlamatrixError):
        capi.cg(capi.to_device_csr(m1), t(np.ones(1, dtype), dev), x1, ws)
    assert host(x1)[0] == 0


def test_jacobi_rejects_bad_diagonal(dev, ws, oracle):
    for vals, cols in (([1.0, 1.0], [1, 0]), ([1.0, 0.0], [0, 1])):
        m = oracle.coo_to_csr(2, 2, [0, 1], cols, np.array(vals))
        x = torch.zeros(2, dtype=torch.float64, device=dev)
        with pytest.raises(capi.PlamatrixError) as e:
            capi.pcg(capi.to_device_csr(m), t([1.0, 1.0], dev), x, ws)
        assert kinds(e) == "runtime_error"


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_fixed_iterations_closed_form(dev, ws, oracle, dtype):
    # iterative_solver_test.cpp:782-820
    m = poisson1d(oracle, 8, dtype)
    a = capi.to_device_csr(m)
    x = torch.zeros(8, dtype=torch.float32 if dtype == np.float32 else torch.float64, device=dev)
    b = t(np.ones(8, dtype), dev)
    st0 = capi.solve_fixed_async(a, b, x, 0, True, ws)
    torch.cuda.synchronize()
    r0 = st0.finalize()
    assert r0.iterations == 0 and r0.initial_residual == r0.final_residual
    st = capi.solve_fixed_async(a, b, x, 4, True, ws)
    torch.cuda.synchronize()
    rep = st.finalize()
    assert rep.iterations == 4 and rep.converged and np.isfinite(rep.final_residual)
    expected = np.array([(i + 1) * (8 - i) / 2 for i in range(8)])
    tol = 2e-4 if dtype == np.float32 else 1e-10
    assert np.all(np.abs(host(x) - expected) <= tol * expected)


def test_c0_poisson_fp64_parity(dev, ws, oracle):
    """2-D Poisson 256^2 from shuffled COO with split duplicates: device
    COO->CSR then Jacobi-PCG to rel 1e-8 (max 2000) against the CPU solve."""
    coo = generators.poisson2d_coo(256)
    ref = oracle.coo_to_csr(coo.rows, coo.cols, coo.row_indices, coo.col_indices, coo.values)
    b = generators.rhs_uniform(coo.rows, seed=1)
    cpu_x, cpu = oracle.pcg(ref, b, relative_tolerance=1e-8, max_iterations=2000)
    a = capi.coo_to_csr(coo.rows, coo.cols, t(coo.row_indices, dev), t(coo.col_indices, dev), t(coo.values, dev), ws)
    x = torch.zeros(coo.rows, dtype=torch.float64, device=dev)
    rep = capi.pcg(a, t(b, dev), x, ws, relative_tolerance=1e-8, max_iterations=2000)
    assert rep.converged and cpu.converged
    assert abs(rep.iterations - cpu.iterations) <= 2
    gx = host(x)
    assert np.linalg.norm(gx - cpu_x) / np.linalg.norm(cpu_x) <= 1e-6
    relres = capi.true_relres(a, t(b, dev), x, ws)
    assert relres <= 1.5e-8
    assert abs(relres - oracle.true_relres(ref, b, gx)) <= 1e-12


def test_solve_is_deterministic_and_graph_equals_direct(dev, ws, oracle):
    h = generators.diffusion3d_csr(24, seed=5)
    a = capi.to_device_csr(h, compact=True)
    b = t(generators.rhs_normal(h.rows, seed=6), dev)
    out = []
    for graph in (1, 1, 0):
        x = torch.zeros(h.rows, dtype=torch.float32, device=dev)
        rep = capi.pcg(a, b, x, ws, capi.policy(use_graph=graph, refresh_interval=50), relative_tolerance=1e-5,
                       max_iterations=3000)
        out.append((rep.iterations, host(x).copy()))
    assert out[0][0] == out[1][0] == out[2][0]
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][1], out[2][1])


def test_fp32_pcg_matches_cpu_within_1e4(dev, ws, oracle):
    h = generators.diffusion3d_csr(32, seed=12)
    b = generators.rhs_normal(h.rows, seed=13)
    ref = oracle.CSR(h.rows, h.cols, h.row_offsets, h.col_indices, h.values)
    cpu_x, cpu = oracle.pcg(ref, b, relative_tolerance=1e-6, max_iterations=5000)
    a = capi.to_device_csr(h, compact=True)
    x = torch.zeros(h.rows, dtype=torch.float32, device=dev)
    rep = capi.pcg(a, t(b, dev), x, ws, capi.policy(refresh_interval=50), relative_tolerance=1e-6, max_iterations=5000)
    assert rep.converged
    gx = host(x).astype(np.float64)
    assert np.linalg.norm(gx - cpu_x) / np.linalg.norm(cpu_x) <= 1e-4
    assert oracle.true_relres(ref, b, gx) <= 5e-6


def test_refinement_reaches_1e8(dev, ws, oracle):
    h = generators.diffusion3d_csr(32, seed=22)
    b = generators.rhs_normal(h.rows, seed=23).astype(np.float64)
    a = capi.to_device_csr(h, compact=True)
    x = torch.zeros(h.rows, dtype=torch.float64, device=dev)
    rep = capi.pcg_refine(a, t(b, dev), x, ws, target=1e-8, max_outer=20, pol=capi.policy(refresh_interval=50),
                          relative_tolerance=1e-5, max_iterations=5000)
    assert rep.converged and rep.final_relres <= 1e-8
    ref = oracle.CSR(h.rows, h.cols, h.row_offsets, h.col_indices, h.values)
    assert oracle.true_relres(ref, b, host(x)) <= 1.2e-8
    cpu_x, _ = oracle.pcg(oracle.CSR(h.rows, h.cols, h.row_offsets, h.col_indices, h.values.astype(np.float64)), b,
                          relative_tolerance=1e-10, max_iterations=10000)
    assert np.linalg.norm(host(x) - cpu_x) / np.linalg.norm(cpu_x) <= 1e-6


def test_autotuner_picks_a_valid_width(dev, ws):
    h = generators.diffusion3d_csr(32)
    a = capi.to_device_csr(h, compact=True)
    x = torch.ones(h.rows, dtype=torch.float32, device=dev)
    y = torch.empty_like(x)
    lanes, ms = capi.autotune_spmv(a, x, y, repeats=5)
    assert lanes in (0, 1, 2, 4, 8, 16, 32, 64) and ms > 0
    assert capi.lib().pm_tuned_lanes(capi.ctypes.byref(a.descriptor())) == lanes


# ---- SpMM -----------------------------------------------------------------------------

@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_spmm_kat(dev, ws, oracle, dtype):
    # sparse_ops_test.cpp:704-737 (column-major B and output)
    ref = oracle.coo_to_csr(3, 3, [1, 0, 2, 0], [1, 2, 0, 0], np.array([4, 2, -1, 3], dtype))
    a = capi.to_device_csr(ref)
    b = t(np.array([[1, 2, 3], [10, 20, 30]], dtype), dev)  # column j of B is row j here
    y = torch.full((2, 3), -100, dtype=b.dtype, device=dev)
    capi.spmm(a, b, y, ws)
    assert host(y).tolist() == [[9, 8, -1], [90, 80, -10]]
    with pytest.raises(capi.PlamatrixError) as e:
        capi.spmm(a, b, b, ws)
    assert kinds(e) == "invalid_argument"
    empty = capi.to_device_csr(oracle.coo_to_csr(0, 3, [], [], np.zeros(0, dtype)))
    capi.spmm(empty, b, torch.empty((2, 0), dtype=b.dtype, device=dev), ws)


@pytest.mark.parametrize("k", [1, 3, 8, 13])
@pytest.mark.parametrize("compact", [False, True])
def test_spmm_random_against_oracle(dev, ws, oracle, k, compact):
    rng = np.random.default_rng(k)
    rows, cols = 3000, 2500
    counts = rng.integers(0, 90, rows)
    counts[::7] = 0
    r = np.repeat(np.arange(rows), counts)
    c = rng.integers(0, cols, len(r))
    v = rng.standard_normal(len(r))
    ref = oracle.coo_to_csr(rows, cols, r, c, v)
    b = rng.standard_normal((cols, k))
    exact = oracle.spmm(ref, b)
    a = capi.to_device_csr(ref, compact=compact)
    y = torch.empty((k, rows), dtype=torch.float64, device=dev)
    capi.spmm(a, t(np.ascontiguousarray(b.T), dev), y, ws)
    np.testing.assert_allclose(host(y).T, exact, rtol=1e-12, atol=1e-12)
    ref32 = oracle.CSR(rows, cols, ref.row_offsets, ref.col_indices, ref.values.astype(np.float32))
    b32 = b.astype(np.float32)
    a32 = capi.to_device_csr(ref32, compact=compact)
    y32 = torch.empty((k, rows), dtype=torch.float32, device=dev)
    capi.spmm(a32, t(np.ascontiguousarray(b32.T), dev), y32, ws)
    abs_ref = oracle.CSR(rows, cols, ref.row_offsets, ref.col_indices, np.abs(ref32.values.astype(np.float64)))
    exact32 = oracle.spmm(oracle.CSR(rows, cols, ref.row_offsets, ref.col_indices, ref32.values.astype(np.float64)),
                          b32.astype(np.float64))
    scale = oracle.spmm(abs_ref, np.abs(b32.astype(np.float64)))
    assert np.all(np.abs(host(y32).T - exact32) <= 4e-6 * scale + 1e-30)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_spmm_columns_equal_spmv_including_long_rows(dev, ws, oracle, dtype):
    """Each column of A B is bit-for-bit the SpMV of that column (same
    left-to-right FMA order), on short, long and empty rows."""
    rng = np.random.default_rng(7)
    rows, cols = 700, 4000
    counts = rng.integers(0, 12, rows)
    counts[100:180] = 300
    counts[::11] = 0
    r = np.repeat(np.arange(rows), counts)
    c = rng.integers(0, cols, len(r))
    ref = oracle.coo_to_csr(rows, cols, r, c, rng.standard_normal(len(r)).astype(dtype))
    a = capi.to_device_csr(ref, compact=True)
    b = t(rng.standard_normal((5, cols)).astype(dtype), dev)
    y = torch.empty((5, rows), dtype=b.dtype, device=dev)
    capi.spmm(a, b, y, ws)
    for j in range(5):
        yj = torch.empty(rows, dtype=b.dtype, device=dev)
        capi.spmv(a, b[j].contiguous(), yj, ws, lanes=0)
        assert torch.equal(y[j], yj), j


def test_spmm_diffusion3d_full_size_eight_columns(dev, ws):
    """128^3 diffusion, 8 right-hand sides (the probe count of the learned
    preconditioner): every column equals its SpMV bit for bit."""
    h = generators.diffusion3d_csr(128)
    a = capi.to_device_csr(h, compact=True)
    b = torch.randn((8, h.rows), dtype=torch.float32, device=dev)
    y = torch.empty_like(b)
    capi.spmm(a, b, y, ws)
    yj = torch.empty(h.rows, dtype=torch.float32, device=dev)
    for j in range(8):
        capi.spmv(a, b[j].contiguous(), yj, ws)
        assert torch.equal(y[j], yj), j


@pytest.mark.parametrize("dims", [(40, 40, 40), (48, 48, 48)])
def test_single_reduction_keeps_iteration_count_on_ill_conditioned_refresh(dev, monkeypatch, dims):
    # FP32 Jacobi-PCG with a residual refresh every 50 iterations on a badly
    # conditioned diffusion box (log-normal sigma 3, z-couplings x 0.01).  The
    # textbook single-reduction denominator (delta - beta gamma / alpha_prev)
    # stalls here: a CPU restatement reaches only relres 0.07 after 20,000
    # iterations, where the three-kernel form converges in about 1,660.  The
    # expanded denominator must converge in the three-kernel form's count.
    # (48^3 = 110,592 rows takes the diagonal layout, 40^3 the CSR kernels.)
    h = generators.diffusion3d_box_csr(*dims, seed=3, sigma=3.0, z_scale=0.01, dtype=np.float32)
    a = capi.to_device_csr(h, compact=True)
    b = torch.from_numpy(np.random.default_rng(1).standard_normal(h.rows).astype(np.float32)).to(dev)
    pol = capi.policy(refresh_interval=50)
    its = {}
    for form in ("1", "0"):
        monkeypatch.setenv("PLAMATRIX_AMD_SINGLE_REDUCTION", form)
        ws = capi.Workspace()
        x = torch.zeros_like(b)
        rep = capi.pcg(a, b, x, ws, pol, relative_tolerance=1e-3, max_iterations=20000)
        assert rep.converged, (form, rep)
        its[form] = rep.iterations
        res = capi.true_relres(a, b.double(), x.double(), ws)
        assert res <= 2e-3, (form, res)
    assert its["1"] <= 1.05 * its["0"], its
