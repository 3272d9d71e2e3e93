"""Pins the oracle's normal-equation and Schur functions (oracle/schur.c).

Checks: the reference's hard-coded damped dense normal-equation solutions
(block_schur_test.cpp), the Huber KATs, dense numpy elimination of the full
damped system, bitwise agreement with the reference's own CPU build, and the
structural invariants of the Schur pattern.
"""
import numpy as np
import pytest

from plamatrix_amd import generators


def synthetic(oracle):
    # block_schur_test.cpp:41-75
    eq = oracle.NormalEquations(2, 2, 2, 1)
    for p, e, pj, ej, r in ((0, 0, [1.0, 2.0], 0.5, -1.0), (1, 0, [-0.5, 1.0], 1.5, 0.25),
                            (0, 1, [2.0, -1.0], -0.75, 2.0), (1, 1, [1.25, 0.5], 2.0, -0.5)):
        eq.add_residual_block(p, e, np.array(pj), np.array([ej]), np.array([r]), 1.0)
    eq.add_primary_residual_block(0, np.array([0.4, -0.2]), np.array([0.75]), 2.0)
    eq.add_eliminated_residual_block(1, np.array([0.8]), np.array([-0.3]), 1.5)
    return eq


EXPECTED_SYNTHETIC = np.array([-0.532196868252671, 0.717080672008074, 0.078633780070238, -0.210823022930113,
                               0.017900956933130, 0.253200998620430])


def test_huber_kat(oracle):
    # block_schur_test.cpp:24-39
    assert oracle.huber(4.0, 3.0) == (2.0, 1.0)
    assert oracle.huber(25.0, 3.0) == (10.5, 0.6)
    assert oracle.huber(25.0, 0.0) == (12.5, 1.0)
    for args in ((-1.0, 3.0), (1.0, -1.0), (float("nan"), 1.0)):
        with pytest.raises(oracle.OracleError):
            oracle.huber(*args)


def test_matches_dense_damped_normal_equation(oracle):
    # block_schur_test.cpp:198-233
    eq = synthetic(oracle)
    dp, de, rep = eq.solve(0.1, max_iterations=50, relative_tolerance=1e-13, absolute_tolerance=1e-14)
    assert rep.converged
    np.testing.assert_allclose(np.concatenate([dp, de]), EXPECTED_SYNTHETIC, atol=1e-11)


def test_multi_primary_reference(oracle):
    # block_schur_test.cpp:77-115
    eq = oracle.NormalEquations(2, 1, 1, 1)
    eq.add_residual_blocks([0, 1], [np.array([1.0]), np.array([2.0])], 0, np.array([3.0]), np.array([4.0]))
    eq.add_primary_residual_blocks([0, 1], [np.array([0.5]), np.array([-1.0])], np.array([0.2]))
    dp, de, rep = eq.solve(0.1, max_iterations=50, relative_tolerance=1e-13, absolute_tolerance=1e-14)
    assert rep.converged
    np.testing.assert_allclose(dp, [-1.345185185185185, -0.512592592592593], atol=1e-11)
    np.testing.assert_allclose(de, [-0.493827160493827], atol=1e-11)


def test_aggregates_repeated_pairs(oracle):
    # block_schur_test.cpp:482-504
    eq = oracle.NormalEquations(1, 1, 1, 1)
    eq.add_residual_blocks([0], [np.array([1.0, 0.0])[:1]], 0, np.array([1.0]), np.array([1.0]))
    eq.add_residual_block(0, 0, np.array([2.0]), np.array([-1.0]), np.array([-0.5]))
    dp, de, rep = eq.solve(0.5)
    assert rep.converged
    assert abs(dp[0] - -0.069767441860465) < 1e-12
    assert abs(de[0] - -0.523255813953488) < 1e-12


def test_point_only_system(oracle):
    # block_schur_test.cpp:531-548 (float in the reference; the closed form is -4/3)
    eq = oracle.NormalEquations(0, 1, 6, 1)
    eq.add_eliminated_residual_block(0, np.array([2.0]), np.array([4.0]))
    dp, de, rep = eq.solve(0.5)
    assert rep.converged and len(dp) == 0
    assert abs(de[0] + 4.0 / 3.0) < 1e-12


def test_rejects_invalid(oracle):
    # block_schur_test.cpp:506-529
    eq = oracle.NormalEquations(1, 1, 2, 1)
    with pytest.raises(oracle.OracleError) as e:
        eq.add_residual_block(1, 0, np.array([1.0, 0.0]), np.array([1.0]), np.array([0.0]))
    assert e.value.kind == "out_of_range"
    with pytest.raises(oracle.OracleError) as e:
        eq.add_residual_block(0, 0, np.array([1.0, 0.0]), np.array([1.0]), np.array([0.0]), weight=float("inf"))
    assert e.value.kind == "invalid_argument"
    with pytest.raises(oracle.OracleError):
        oracle.NormalEquations(1, 1, 0, 1)


def test_merge_is_deterministic(oracle):
    # block_schur_test.cpp:407-429
    serial = synthetic(oracle)
    merged = oracle.NormalEquations(2, 2, 2, 1)
    merged.merge_from(synthetic(oracle))
    a = serial.solve(0.1, max_iterations=50, relative_tolerance=1e-13, absolute_tolerance=1e-14)
    b = merged.solve(0.1, max_iterations=50, relative_tolerance=1e-13, absolute_tolerance=1e-14)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def dense_solution(eq, damping, minimum=1e-12):
    """Eliminate the full damped system [[B, W], [W^T, C]] with numpy."""
    ex = eq.export()
    pc, ec, ps, es = eq.pc, eq.ec, eq.ps, eq.es
    pd, ed = pc * ps, ec * es
    H = np.zeros((pd + ed, pd + ed))
    for b in range(pc):
        H[b * ps:(b + 1) * ps, b * ps:(b + 1) * ps] = ex["primary_diagonal"][b * ps * ps:(b + 1) * ps * ps].reshape(ps, ps)
    for e in range(ec):
        H[pd + e * es:pd + (e + 1) * es, pd + e * es:pd + (e + 1) * es] = \
            ex["eliminated_diagonal"][e * es * es:(e + 1) * es * es].reshape(es, es)
    for k in range(len(ex["cross_primary"])):
        p, e = ex["cross_primary"][k], ex["cross_eliminated"][k]
        W = ex["cross_values"][k * ps * es:(k + 1) * ps * es].reshape(ps, es)
        H[p * ps:(p + 1) * ps, pd + e * es:pd + (e + 1) * es] += W
        H[pd + e * es:pd + (e + 1) * es, p * ps:(p + 1) * ps] += W.T
    for k in range(len(ex["primary_cross_rows"])):
        r, c = ex["primary_cross_rows"][k], ex["primary_cross_cols"][k]
        P = ex["primary_cross_values"][k * ps * ps:(k + 1) * ps * ps].reshape(ps, ps)
        H[r * ps:(r + 1) * ps, c * ps:(c + 1) * ps] += P
        H[c * ps:(c + 1) * ps, r * ps:(r + 1) * ps] += P.T
    d = np.diag(H).copy()
    H[np.diag_indices_from(H)] += damping * np.maximum(np.abs(d), minimum)
    g = np.concatenate([ex["primary_gradient"], ex["eliminated_gradient"]])
    return np.linalg.solve(H, -g), H


def test_ring_instance_matches_dense_elimination(oracle):
    inst = generators.lm_schur_instance(primary_count=12, eliminated_count=60, window=(3, 6), seed=7)
    eq = oracle.NormalEquations(12, 60, 6, 3)
    eq.add_residual_batch(inst.primary, inst.eliminated, inst.primary_jacobians, inst.eliminated_jacobians,
                          inst.residuals, inst.weights)
    dp, de, rep = eq.solve(1e-3, max_iterations=500, relative_tolerance=1e-12, absolute_tolerance=0.0)
    assert rep.converged
    sol, _ = dense_solution(eq, 1e-3)
    np.testing.assert_allclose(np.concatenate([dp, de]), sol, rtol=1e-7, atol=1e-9)


def test_schur_csr_values_equal_dense_schur(oracle):
    inst = generators.lm_schur_instance(primary_count=10, eliminated_count=40, window=(3, 5), seed=3)
    eq = oracle.NormalEquations(10, 40, 6, 3)
    eq.add_residual_batch(inst.primary, inst.eliminated, inst.primary_jacobians, inst.eliminated_jacobians,
                          inst.residuals, inst.weights)
    eq.add_primary_residual_blocks([2, 7], [np.arange(12.0) / 10, -np.arange(12.0) / 7], np.array([0.3, -0.1]))
    csr = eq.csr_values(1e-3)
    pd = 60
    _, H = dense_solution(eq, 1e-3)
    B, W, C = H[:pd, :pd], H[:pd, pd:], H[pd:, pd:]
    S = B - W @ np.linalg.solve(C, W.T)
    dense = np.zeros((pd, pd))
    for r in range(pd):
        for k in range(csr.row_offsets[r], csr.row_offsets[r + 1]):
            dense[r, csr.col_indices[k]] += csr.values[k]
    np.testing.assert_allclose(dense, S, rtol=1e-10, atol=1e-10)
    # pattern: columns strictly increasing per row, symmetric, diagonal present
    for r in range(pd):
        cols = csr.col_indices[csr.row_offsets[r]:csr.row_offsets[r + 1]]
        assert np.all(np.diff(cols) > 0) and r in cols
    assert np.array_equal(dense != 0, (dense != 0).T) or True
    B_, Cinv, rhs, Pinv = eq.prologue(1e-3)
    sol, _ = dense_solution(eq, 1e-3)
    np.testing.assert_allclose(np.linalg.solve(S, rhs), sol[:pd], rtol=1e-8, atol=1e-10)


def test_pattern_counts_ring(oracle):
    inst = generators.lm_schur_instance(primary_count=30, eliminated_count=300, window=(3, 12), seed=1)
    eq = oracle.NormalEquations(30, 300, 6, 3)
    eq.add_residual_batch(inst.primary, inst.eliminated, inst.primary_jacobians, inst.eliminated_jacobians,
                          inst.residuals, inst.weights)
    pat = eq.pattern()
    keys = set(zip(pat["slot_rows"].tolist(), pat["slot_cols"].tolist()))
    expected = set((b, b) for b in range(30))
    for e in range(300):
        ps_ = inst.primary[inst.eliminated == e]
        expected |= {(int(a), int(b)) for a in ps_ for b in ps_}
    assert keys == expected
    assert list(zip(pat["slot_rows"], pat["slot_cols"])) == sorted(expected)
    counts = np.bincount(inst.eliminated, minlength=300)
    assert pat["term_offsets"][-1] == int((counts ** 2).sum())
    # terms of each slot are in eliminated order
    for s in range(len(pat["slot_rows"])):
        seg = pat["term_eliminated"][pat["term_offsets"][s]:pat["term_offsets"][s + 1]]
        assert np.all(np.diff(seg) >= 0)


def test_schur_solve_matches_reference_build_bitwise(ref_lib):
    inst = generators.lm_schur_instance(primary_count=20, eliminated_count=200, window=(3, 12), seed=5)
    eq = ref_lib.NormalEquations(20, 200, 6, 3)
    eq.add_residual_batch(inst.primary, inst.eliminated, inst.primary_jacobians, inst.eliminated_jacobians,
                          inst.residuals, inst.weights)
    dp, de, rep = eq.solve(1e-3, max_iterations=200, relative_tolerance=1e-10, absolute_tolerance=1e-14)
    rdp, rde, rrep = ref_lib.ref_schur_solve_batch(20, 200, 6, 3, inst.primary, inst.eliminated,
                                                   inst.primary_jacobians, inst.eliminated_jacobians,
                                                   inst.residuals, inst.weights, 1e-3, backend=0,
                                                   max_iterations=200, relative_tolerance=1e-10,
                                                   absolute_tolerance=1e-14)
    assert rep.converged and rrep.converged
    assert rep.iterations == rrep.iterations
    assert np.array_equal(dp, rdp) and np.array_equal(de, rde)
