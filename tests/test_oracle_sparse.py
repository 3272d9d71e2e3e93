"""Pins the oracle's sparse functions (oracle/sparse.c).

Each check is one the reference's own tests hold (cited), a closed form, a
brute-force enumeration, or scipy.  A plausible slip in the restatement (fold
order, sort stability, tolerance rule, preconditioner) fails at least one.
"""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def poisson1d(oracle, n, scale=1.0, dtype=np.float64):
    """iterative_solver_test.cpp:45-70"""
    rows, cols, vals = [], [], []
    for r in range(n):
        if r > 0:
            rows.append(r); cols.append(r - 1); vals.append(-scale)
        rows.append(r); cols.append(r); vals.append(2 * scale)
        if r + 1 < n:
            rows.append(r); cols.append(r + 1); vals.append(-scale)
    return oracle.coo_to_csr(n, n, rows, cols, np.array(vals, dtype=dtype))


# ---- COO -> CSR ----------------------------------------------------------

def test_coo_sorts_and_combines_duplicates_in_insertion_order(oracle):
    # sparse_ops_test.cpp:22-45 (1e16 catastrophic cancellation)
    csr = oracle.coo_to_csr(4, 5, [2, 0, 2, 1, 2, 0, 2], [4, 3, 1, 2, 1, 3, 1],
                            np.array([5.0, 2.0, 1.0e16, 4.0, -1.0e16, -2.0, 1.0]))
    assert csr.nnz == 4
    assert list(csr.row_offsets) == [0, 1, 2, 4, 4]
    assert list(csr.col_indices) == [3, 2, 1, 4]
    assert list(csr.values) == [0.0, 4.0, 1.0, 5.0]


def test_coo_float_32_cancellation(oracle):
    # sparse_ops_test.cpp:320-369 with 1e8 for float
    csr = oracle.coo_to_csr(4, 5, [2, 0, 2, 1, 2, 0, 2], [4, 3, 1, 2, 1, 3, 1],
                            np.array([5, 2, 1e8, 4, -1e8, -2, 1], dtype=np.float32))
    assert list(csr.values) == [0.0, 4.0, 1.0, 5.0]
    assert csr.values.dtype == np.float32


def test_coo_reuses_float_and_combines(oracle):
    # sparse_ops_test.cpp:49-71
    csr = oracle.coo_to_csr(2, 2, [1, 0, 1], [0, 1, 0], np.array([2, 3, 4], np.float32))
    assert list(csr.row_offsets) == [0, 1, 2]
    assert list(csr.col_indices) == [1, 0]
    assert list(csr.values) == [3.0, 6.0]


def test_coo_empty(oracle):
    # sparse_ops_test.cpp:73-84
    csr = oracle.coo_to_csr(0, 5, [], [], np.array([], np.float32))
    assert csr.rows == 0 and csr.nnz == 0 and list(csr.row_offsets) == [0]


def test_coo_rejects_invalid(oracle):
    # sparse_ops_test.cpp:86-118
    with pytest.raises(oracle.OracleError) as e:
        oracle.coo_to_csr(-1, 1, [], [], np.array([], np.float32))
    assert e.value.kind == "invalid_argument"
    with pytest.raises(oracle.OracleError) as e:
        oracle.coo_to_csr(1, -1, [], [], np.array([], np.float32))
    assert e.value.kind == "invalid_argument"
    with pytest.raises(oracle.OracleError) as e:
        oracle.coo_to_csr(1, 1, [0], [], np.array([1.0], np.float32))
    assert e.value.kind == "invalid_argument"
    for r, c in ([-1], [0]), ([0], [-1]), ([1], [0]), ([0], [1]):
        with pytest.raises(oracle.OracleError) as e:
            oracle.coo_to_csr(1, 1, r, c, np.array([1.0], np.float32))
        assert e.value.kind == "out_of_range"
    with pytest.raises(oracle.OracleError) as e:
        oracle.coo_to_csr(2, 2, [2 ** 31], [0], np.array([1.0], np.float32))
    assert e.value.kind == "out_of_range"
    assert "triplet 0" in str(e.value)


def test_coo_matrix_to_csr(oracle):
    # sparse_ops_test.cpp:276-293 (COOMatrix::toCsr)
    csr = oracle.coo_to_csr(2, 3, [1, 0, 1], [2, 1, 2], np.array([1.5, 2.0, 2.5]))
    assert list(csr.row_offsets) == [0, 1, 2]
    assert list(csr.col_indices) == [1, 2]
    assert list(csr.values) == [2.0, 4.0]


def test_coo_async_case_values(oracle):
    # sparse_ops_test.cpp:371-430 expected CSR
    csr = oracle.coo_to_csr(3, 3, [1, 0, 1, 2, 1], [0, 2, 0, 1, 0], np.array([2, 3, 4, 5, -1.0]))
    assert list(csr.row_offsets) == [0, 1, 2, 3]
    assert list(csr.col_indices) == [2, 0, 1]
    assert list(csr.values) == [3.0, 5.0, 5.0]


def test_coo_fold_order_brute_force(oracle):
    """Every ordering of up to 6 duplicate triplets: the result must be the left
    fold in input order from the first value (not from 0, not reordered)."""
    base = np.array([1e16, 1.0, -1e16, 3.0, -0.0, 2.0 ** -60])
    for k in range(1, 7):
        for perm in itertools.permutations(range(k)):
            vals = base[list(perm)]
            csr = oracle.coo_to_csr(1, 1, [0] * k, [0] * k, vals)
            expected = vals[0]
            for v in vals[1:]:
                expected = expected + v
            assert csr.nnz == 1
            assert np.array_equal(csr.values.view(np.int64), np.array([expected]).view(np.int64)), (perm, csr.values)


def test_coo_negative_zero_kept_when_alone(oracle):
    # the CPU fold starts from the first value, so a lone -0.0 stays -0.0
    csr = oracle.coo_to_csr(1, 2, [0], [1], np.array([-0.0]))
    assert np.signbit(csr.values[0])


def test_coo_random_matches_scipy_structure(oracle):
    rng = np.random.default_rng(5)
    n, m, t = 50, 40, 800
    r = rng.integers(0, n, t)
    c = rng.integers(0, m, t)
    v = rng.integers(-4, 5, t).astype(np.float64)  # small integers: sums exact in any order
    csr = oracle.coo_to_csr(n, m, r, c, v)
    ref = sp.coo_matrix((v, (r, c)), shape=(n, m)).tocsr()
    ref.sum_duplicates()
    ref.sort_indices()
    # scipy drops nothing on sum_duplicates; explicit zeros are kept by both
    assert np.array_equal(csr.row_offsets, ref.indptr)
    assert np.array_equal(csr.col_indices, ref.indices)
    assert np.array_equal(csr.values, ref.data)


def test_coo_matches_reference_build_bitwise(ref_lib):
    rng = np.random.default_rng(11)
    for dtype in (np.float64, np.float32):
        n, m, t = 300, 250, 6000
        r = rng.integers(0, n, t)
        c = rng.integers(0, m, t)
        v = (rng.standard_normal(t) * 10.0 ** rng.integers(-8, 9, t)).astype(dtype)
        ours = ref_lib.coo_to_csr(n, m, r, c, v)
        theirs = ref_lib.coo_to_csr(n, m, r, c, v, use_ref=True)
        assert np.array_equal(ours.row_offsets, theirs.row_offsets)
        assert np.array_equal(ours.col_indices, theirs.col_indices)
        assert np.array_equal(ours.values.view(np.uint8), theirs.values.view(np.uint8))


# ---- SpMV ---------------------------------------------------------------

def test_spmv_kat(oracle):
    # sparse_ops_test.cpp:120-149
    csr = oracle.coo_to_csr(3, 4, [2, 0, 1, 0], [3, 2, 1, 0], np.array([5, 2, -1, 3], np.float32))
    y = oracle.spmv(csr, np.array([1, 2, 3, 4], np.float32))
    assert list(y) == [9.0, -2.0, 20.0]


def test_spmv_empty_rows(oracle):
    # sparse_ops_test.cpp:152-173
    csr = oracle.coo_to_csr(3, 2, [1], [0], np.array([2.0]))
    assert list(oracle.spmv(csr, np.array([3.0, 4.0]))) == [0.0, 6.0, 0.0]


def test_spmv_row_sums_and_reference(oracle):
    rng = np.random.default_rng(3)
    a = sp.random(200, 150, density=0.05, random_state=4, format="coo")
    csr = oracle.coo_to_csr(200, 150, a.row, a.col, a.data)
    ones = oracle.spmv(csr, np.ones(150))
    np.testing.assert_allclose(ones, np.asarray(a.sum(axis=1)).ravel(), rtol=1e-14, atol=1e-14)
    x = rng.standard_normal(150)
    np.testing.assert_allclose(oracle.spmv(csr, x), a @ x, rtol=1e-12, atol=1e-12)


def test_spmv_matches_reference_bitwise(ref_lib):
    rng = np.random.default_rng(9)
    for dtype in (np.float32, np.float64):
        a = sp.random(400, 300, density=0.03, random_state=8, format="coo")
        csr = ref_lib.coo_to_csr(400, 300, a.row, a.col, a.data.astype(dtype))
        x = rng.standard_normal(300).astype(dtype)
        assert np.array_equal(ref_lib.spmv(csr, x).view(np.uint8), ref_lib.spmv(csr, x, use_ref=True).view(np.uint8))


def test_spmm_kat_column_major(oracle):
    # sparse_ops_test.cpp:196-233
    for dtype in (np.float32, np.float64):
        csr = oracle.coo_to_csr(3, 3, [1, 0, 2, 0], [1, 2, 0, 0], np.array([4.0, 2.0, -1.0, 3.0], dtype))
        b = np.array([[1.0, 10.0], [2.0, 20.0], [3.0, 30.0]], dtype)
        y = oracle.spmm(csr, b)
        assert y.tolist() == [[9.0, 90.0], [8.0, 80.0], [-1.0, -10.0]]


def test_spmm_empty_and_aliasing(oracle):
    # sparse_ops_test.cpp:237-265
    csr = oracle.coo_to_csr(0, 3, [], [], np.zeros(0, np.float32))
    assert oracle.spmm(csr, np.zeros((3, 2), np.float32)).shape == (0, 2)
    two = oracle.coo_to_csr(2, 2, [0, 1], [0, 1], np.array([2.0, 3.0]))
    b = np.zeros((2, 2), order="F")
    with pytest.raises(oracle.OracleError) as e:
        oracle._check(oracle.lib().or_spmm_f64(oracle._i64(2), oracle._i64(2), oracle._i64(2),
                                               oracle._ptr(two.row_offsets), oracle._ptr(two.col_indices),
                                               oracle._ptr(two.values), oracle._ptr(b), oracle._ptr(b)))
    assert "alias" in str(e.value)


def test_spmm_columns_are_spmv(oracle):
    # the definition: column j of A B is A times column j of B, bit for bit
    rng = np.random.default_rng(12)
    a = sp.random(300, 200, density=0.04, random_state=13, format="coo")
    for dtype in (np.float32, np.float64):
        csr = oracle.coo_to_csr(300, 200, a.row, a.col, a.data.astype(dtype))
        b = rng.standard_normal((200, 5)).astype(dtype)
        y = oracle.spmm(csr, b)
        for j in range(5):
            assert np.array_equal(y[:, j], oracle.spmv(csr, np.ascontiguousarray(b[:, j])))
        np.testing.assert_allclose(y.astype(np.float64), a @ b.astype(np.float64),
                                   rtol=1e-4 if dtype == np.float32 else 1e-12, atol=1e-4 if dtype == np.float32 else 1e-12)


def test_spmm_matches_reference_bitwise(ref_lib):
    rng = np.random.default_rng(21)
    for dtype in (np.float32, np.float64):
        a = sp.random(250, 180, density=0.05, random_state=22, format="coo")
        csr = ref_lib.coo_to_csr(250, 180, a.row, a.col, a.data.astype(dtype))
        b = rng.standard_normal((180, 3)).astype(dtype)
        mine, theirs = ref_lib.spmm(csr, b), ref_lib.spmm(csr, b, use_ref=True)
        assert np.array_equal(mine.ravel(order="F").view(np.uint8), theirs.ravel(order="F").view(np.uint8))


# ---- CG / PCG ------------------------------------------------------------

@pytest.mark.parametrize("dtype,tol", [(np.float32, 2e-4), (np.float64, 1e-10)])
def test_pcg_diagonal_one_iteration(oracle, dtype, tol):
    # iterative_solver_test.cpp:103-123
    csr = oracle.coo_to_csr(3, 3, [0, 1, 2], [0, 1, 2], np.array([2, 4, 8], dtype))
    x, rep = oracle.pcg(csr, np.array([2, 8, 24], dtype))
    assert rep.converged and rep.iterations == 1
    np.testing.assert_allclose(x, [1, 2, 3], atol=tol)
    assert rep.final_residual <= rep.initial_residual


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_cg_pcg_poisson_across_scales(oracle, dtype):
    # iterative_solver_test.cpp:125-155
    tol = 2e-4 if dtype == np.float32 else 1e-10
    for scale in (1e-3, 1.0, 1e3):
        csr = poisson1d(oracle, 8, scale, dtype)
        expected = np.arange(1, 9, dtype=dtype)
        rhs = oracle.spmv(csr, expected)
        rel = 1e-5 if dtype == np.float32 else 1e-12
        xc, rc = oracle.cg(csr, rhs, relative_tolerance=rel)
        xp, rp = oracle.pcg(csr, rhs, relative_tolerance=rel)
        assert rc.converged and rp.converged
        for i in range(8):
            assert abs(xc[i] - expected[i]) <= tol * (i + 1)
            assert abs(xp[i] - expected[i]) <= tol * (i + 1)


def test_zero_iteration_cases(oracle):
    # iterative_solver_test.cpp:157-176
    csr = poisson1d(oracle, 4)
    x, rep = oracle.cg(csr, np.zeros(4))
    assert rep.converged and rep.iterations == 0 and rep.initial_residual == 0 and rep.final_residual == 0
    two = np.full(4, 2.0)
    x, rep = oracle.pcg(csr, oracle.spmv(csr, two), x0=two)
    assert rep.converged and rep.iterations == 0


def test_rejections(oracle):
    # iterative_solver_test.cpp:178-214, 253-286
    nonsq = oracle.coo_to_csr(2, 3, [0], [0], np.array([1.0]))
    with pytest.raises(oracle.OracleError) as e:
        oracle.cg(nonsq, np.zeros(2), np.zeros(3))
    assert e.value.kind == "invalid_argument"
    sq = poisson1d(oracle, 2)
    for opts in (dict(max_iterations=-1), dict(relative_tolerance=-1.0), dict(relative_tolerance=1e100),
                 dict(relative_tolerance=float("nan")), dict(absolute_tolerance=float("inf"))):
        with pytest.raises(oracle.OracleError) as e:
            oracle.cg(sq, np.zeros(2), **opts)
        assert e.value.kind == "invalid_argument"
    for vals, cols in (([1.0, 1.0], [1, 0]), ([1.0, 0.0], [0, 1]), ([-1.0, 1.0], [0, 1])):
        m = oracle.coo_to_csr(2, 2, [0, 1], cols, np.array(vals))
        with pytest.raises(oracle.OracleError) as e:
            oracle.pcg(m, np.ones(2))
        assert e.value.kind == "runtime_error"
    bad = oracle.CSR(2, 2, np.array([0, 2, 1], np.int64), np.zeros(2, np.int64), np.ones(2))
    with pytest.raises(oracle.OracleError) as e:
        oracle.cg(bad, np.ones(2))
    assert e.value.kind == "invalid_argument"
    badc = oracle.CSR(2, 2, np.array([0, 1, 2], np.int64), np.array([0, 2], np.int64), np.ones(2))
    with pytest.raises(oracle.OracleError) as e:
        oracle.cg(badc, np.ones(2))
    assert e.value.kind == "invalid_argument"


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_jacobi_small_scale_and_duplicate_diagonal(oracle, dtype):
    # iterative_solver_test.cpp:216-251
    small = dtype(1e-7) if dtype == np.float32 else 1e-20
    m = oracle.coo_to_csr(2, 2, [0, 1], [0, 1], np.array([small, 2 * small], dtype))
    x, rep = oracle.pcg(m, np.array([small, 2 * small], dtype))
    assert rep.converged and rep.iterations == 1
    tol = 2e-4 if dtype == np.float32 else 1e-10
    np.testing.assert_allclose(x, [1, 1], atol=tol)
    dup = oracle.CSR(2, 2, np.array([0, 2, 4], np.int64), np.array([0, 0, 1, 1], np.int64),
                     np.array([1, 1, 1, 3], dtype))
    x, rep = oracle.pcg(dup, np.array([2, 4], dtype))
    assert rep.converged and rep.iterations == 1


def test_absolute_tolerance_zero_iterations_unpreconditioned(oracle):
    # iterative_solver_test.cpp:288-314
    m = poisson1d(oracle, 4)
    x, rep = oracle.cg(m, np.ones(4), max_iterations=0, relative_tolerance=0.0)
    assert not rep.converged and rep.iterations == 0
    x, rep = oracle.cg(m, np.ones(4), x0=x, max_iterations=10, relative_tolerance=0.0, absolute_tolerance=2.0)
    assert rep.converged and rep.iterations == 0
    x, rep = oracle.pcg(m, np.ones(4), x0=x, relative_tolerance=1e-12, use_jacobi=False)
    assert rep.converged


def test_nonconvergence_reporting(oracle):
    # iterative_solver_test.cpp:316-335
    m = poisson1d(oracle, 16)
    x, rep = oracle.cg(m, np.ones(16), max_iterations=1, relative_tolerance=0.0)
    assert not rep.converged and rep.iterations == 1
    assert np.isfinite(rep.initial_residual) and np.isfinite(rep.final_residual)
    with pytest.raises(oracle.OracleError) as e:
        oracle.cg(m, np.ones(16), max_iterations=1, relative_tolerance=0.0, require_convergence=True)
    assert e.value.kind == "runtime_error"


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_fixed_four_iterations_closed_form(oracle, dtype):
    # iterative_solver_test.cpp:782-820: poisson1d(8), rhs 1 -> exact in 4
    # Jacobi-PCG steps, x_i = (i+1)(8-i)/2
    m = poisson1d(oracle, 8, dtype=dtype)
    x, rep = oracle.pcg(m, np.ones(8, dtype), max_iterations=4, relative_tolerance=0.0)
    expected = np.array([(i + 1) * (8 - i) / 2 for i in range(8)])
    tol = 2e-4 if dtype == np.float32 else 1e-10
    np.testing.assert_allclose(x, expected, rtol=tol)
    assert rep.iterations == 4
    assert rep.final_residual < 1e-5 * rep.initial_residual


def test_pcg_matches_spsolve(oracle):
    rng = np.random.default_rng(21)
    n = 60
    a = sp.random(n, n, density=0.1, random_state=22)
    a = (a + a.T) * 0.5 + sp.diags(np.abs(a).sum(axis=1).A1 + 1.0)
    a = a.tocoo()
    csr = oracle.coo_to_csr(n, n, a.row, a.col, a.data)
    b = rng.standard_normal(n)
    x, rep = oracle.pcg(csr, b, relative_tolerance=1e-13)
    np.testing.assert_allclose(x, spla.spsolve(a.tocsc(), b), rtol=1e-9, atol=1e-11)
    # unit-diagonal matrix: Jacobi-PCG and CG produce the same iterates
    d = 1.0 / np.sqrt(a.diagonal())
    s = (sp.diags(d) @ a @ sp.diags(d)).tocoo()
    scsr = oracle.coo_to_csr(n, n, s.row, s.col, s.data)
    xs1, r1 = oracle.pcg(scsr, b, relative_tolerance=1e-10)
    xs2, r2 = oracle.cg(scsr, b, relative_tolerance=1e-10)
    assert r1.iterations == r2.iterations
    np.testing.assert_allclose(xs1, xs2, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_block_pcg_two_blocks_one_iteration(oracle, dtype):
    # iterative_solver_test.cpp:78-101, 394-424 (analytic inverse blocks)
    m = oracle.coo_to_csr(4, 4, [0, 0, 1, 1, 2, 2, 3, 3], [0, 1, 0, 1, 2, 3, 2, 3],
                          np.array([4, 1, 1, 3, 2, 0.5, 0.5, 1.5], dtype))
    inv = np.array([3 / 11, -1 / 11, -1 / 11, 4 / 11, 1.5 / 2.75, -0.5 / 2.75, -0.5 / 2.75, 2 / 2.75], dtype)
    expected = np.array([1, 2, -1, 3], dtype)
    rhs = oracle.spmv(m, expected)
    rel = 1e-5 if dtype == np.float32 else 1e-12
    x, rep = oracle.block_pcg(m, rhs, inv, 2, relative_tolerance=rel)
    assert rep.converged and rep.iterations == 1
    np.testing.assert_allclose(x, expected, atol=2e-4 if dtype == np.float32 else 1e-10)


def test_block_pcg_block_one_equals_jacobi(oracle):
    m = poisson1d(oracle, 20)
    inv = np.full(20, 0.5)
    rhs = np.linspace(-1, 1, 20)
    xb, rb = oracle.block_pcg(m, rhs, inv, 1, relative_tolerance=1e-12)
    xj, rj = oracle.pcg(m, rhs, relative_tolerance=1e-12)
    assert rb.iterations == rj.iterations
    assert np.array_equal(xb, xj)
    with pytest.raises(oracle.OracleError):
        oracle.block_pcg(m, rhs, inv, 3)


def test_solve_matches_reference_bitwise(ref_lib):
    rng = np.random.default_rng(31)
    n = 120
    a = sp.random(n, n, density=0.05, random_state=32)
    a = ((a + a.T) * 0.5 + sp.diags(np.abs(a).sum(axis=1).A1 + 0.5)).tocoo()
    for dtype in (np.float64, np.float32):
        csr = ref_lib.coo_to_csr(n, n, a.row, a.col, a.data.astype(dtype))
        b = rng.standard_normal(n).astype(dtype)
        for pre in (ref_lib.cg, ref_lib.pcg):
            x1, r1 = pre(csr, b, relative_tolerance=1e-9)
            x2, r2 = pre(csr, b, relative_tolerance=1e-9, use_ref=True)
            assert r1 == r2
            assert np.array_equal(x1.view(np.uint8), x2.view(np.uint8))


def test_true_relres(oracle):
    m = poisson1d(oracle, 10)
    x = np.arange(10.0)
    b = oracle.spmv(m, x)
    assert oracle.true_relres(m, b, x) == 0.0
    assert abs(oracle.true_relres(m, b, x + 1e-3) - np.linalg.norm(oracle.spmv(m, np.full(10, 1e-3))) /
               np.linalg.norm(b)) < 1e-15
