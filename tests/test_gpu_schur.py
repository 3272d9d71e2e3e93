"""GPU parity of the damped Schur-complement path through the C ABI.

Normal equations assembled on the device from a residual batch must equal the
oracle's BlockNormalEquations (oracle/schur.c) bit for bit: the device folds
each group in input order without contraction, as addResidualBlocks does.
The reduced system's pattern must equal the oracle's std::map-ordered pattern
exactly, its values and right-hand side must match the oracle's CPU
accumulation (bitwise where the operation order is the same), and the solved
steps must agree with the oracle's CPU solve within relative 1e-6 (FP64),
with the FP64 residual of the reduced system stated for every solve.
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


@pytest.fixture
def sw(dev):
    w = capi.SchurWorkspace()
    yield w
    torch.cuda.synchronize()


def t(a, dev, dtype=None):
    arr = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype=dtype))
    return torch.from_numpy(arr).to(dev)


def shuffled_instance(pc, ec, window, seed, **kw):
    inst = generators.lm_schur_instance(primary_count=pc, eliminated_count=ec, window=window, seed=seed, **kw)
    order = np.random.default_rng(seed + 1000).permutation(len(inst.primary))
    for name in ("primary", "eliminated", "primary_jacobians", "eliminated_jacobians", "residuals", "weights"):
        setattr(inst, name, np.ascontiguousarray(getattr(inst, name)[order]))
    return inst


def squared_norm(r):
    s = 0.0
    for v in r:
        s += float(v) * float(v)
    return s


def device_batch(inst, dev):
    return dict(primary=t(inst.primary, dev, np.int64), eliminated=t(inst.eliminated, dev, np.int64),
                primary_jacobians=t(inst.primary_jacobians, dev, np.float64),
                eliminated_jacobians=t(inst.eliminated_jacobians, dev, np.float64),
                residuals=t(inst.residuals, dev, np.float64))


def oracle_equations(oracle, inst, weights):
    eq = oracle.NormalEquations(inst.primary_count, inst.eliminated_count, inst.primary_size, inst.eliminated_size)
    eq.add_residual_batch(inst.primary, inst.eliminated, inst.primary_jacobians, inst.eliminated_jacobians,
                          inst.residuals, weights)
    return eq


def reduced_relres(sw, dp):
    a, rhs_ptr, _ = capi.schur_system(sw)
    n = a.rows
    ro = capi.device_array(a.row_offsets, n + 1, "int64")
    ci = capi.device_array(a.col_indices, a.nnz, "int64")
    vals = capi.device_array(a.values, a.nnz)
    rhs = capi.device_array(rhs_ptr, n)
    import scipy.sparse as sp

    S = sp.csr_matrix((vals, ci, ro), shape=(n, n))
    x = dp.detach().cpu().numpy()
    return np.linalg.norm(rhs - S @ x) / np.linalg.norm(rhs), S, rhs


# ---- normal equations ---------------------------------------------------------------

@pytest.mark.parametrize("huber", [False, True])
def test_normal_equations_bitwise(dev, sw, oracle, huber):
    inst = shuffled_instance(16, 300, (3, 8), seed=11)
    b = device_batch(inst, dev)
    if huber:
        weights = np.array([oracle.huber(squared_norm(r), inst.huber_delta)[1] for r in inst.residuals])
        ne = capi.assemble_normal_equations(sw, 16, 300, 6, 3, **b, weights=None, huber_delta=inst.huber_delta)
    else:
        weights = inst.weights
        ne = capi.assemble_normal_equations(sw, 16, 300, 6, 3, **b, weights=t(weights, dev, np.float64))
    got = capi.normal_equations_arrays(ne)
    ref = oracle_equations(oracle, inst, weights).export()
    for key in ("primary_diagonal", "eliminated_diagonal", "primary_gradient", "eliminated_gradient",
                "cross_primary", "cross_eliminated", "cross_values"):
        assert got[key].shape == ref[key].shape, key
        assert np.array_equal(got[key].view(np.uint64), ref[key].view(np.uint64)), key


def test_normal_equations_repeated_pairs_and_reuse(dev, sw, oracle):
    # repeated (primary, eliminated) pairs aggregate into one cross block
    # (block_schur_test.cpp:482-504); a second batch with the same linkage
    # reuses the grouping
    rng = np.random.default_rng(5)
    n = 400
    primary = rng.integers(0, 5, n)
    eliminated = rng.integers(0, 20, n)
    jp = rng.standard_normal((n, 2 * 4))
    je = rng.standard_normal((n, 2 * 2))
    r = rng.standard_normal((n, 2))
    w = rng.random(n) + 0.5

    class I:
        pass

    inst = I()
    inst.primary_count, inst.eliminated_count, inst.primary_size, inst.eliminated_size = 5, 20, 4, 2
    inst.primary, inst.eliminated, inst.primary_jacobians, inst.eliminated_jacobians = primary, eliminated, jp, je
    inst.residuals = r
    ref = oracle_equations(oracle, inst, w).export()
    for scale in (1.0, 2.0):
        ne = capi.assemble_normal_equations(sw, 5, 20, 4, 2, t(primary, dev, np.int64), t(eliminated, dev, np.int64),
                                            t(jp * scale, dev, np.float64), t(je, dev, np.float64),
                                            t(r, dev, np.float64), weights=t(w, dev, np.float64))
        got = capi.normal_equations_arrays(ne)
        if scale == 1.0:
            for key in ("primary_diagonal", "eliminated_diagonal", "cross_primary", "cross_values"):
                assert np.array_equal(got[key].view(np.uint64), ref[key].view(np.uint64)), key
    assert len(got["cross_primary"]) == len(ref["cross_primary"]) < n


def test_normal_equations_errors(dev, sw):
    b = dict(primary=t([0, 3], dev, np.int64), eliminated=t([0, 0], dev, np.int64),
             primary_jacobians=t(np.ones((2, 2)), dev, np.float64), eliminated_jacobians=t(np.ones((2, 1)), dev,
                                                                                             np.float64),
             residuals=t(np.ones((2, 1)), dev, np.float64))
    with pytest.raises(capi.PlamatrixError) as e:
        capi.assemble_normal_equations(sw, 2, 1, 2, 1, **b)
    assert e.value.kind == "out_of_range"
    b["primary"] = t([0, 1], dev, np.int64)
    b["residuals"] = t([[1.0], [float("nan")]], dev, np.float64)
    with pytest.raises(capi.PlamatrixError) as e:
        capi.assemble_normal_equations(sw, 2, 1, 2, 1, **b)
    assert e.value.kind == "invalid_argument"
    b["residuals"] = t(np.ones((2, 1)), dev, np.float64)
    with pytest.raises(capi.PlamatrixError) as e:
        capi.assemble_normal_equations(sw, 2, 1, 2, 1, **b, weights=t([1.0, -1.0], dev, np.float64))
    assert e.value.kind == "invalid_argument"


# ---- Schur pattern, values and solve ------------------------------------------------

@pytest.mark.parametrize("kernel", ["slots", "segments"])
def test_pattern_and_values_match_oracle(dev, sw, ws, oracle, monkeypatch, kernel):
    # slots: the CPU accumulation order, bit for bit; segments (the default
    # for 6 x 6 / 3 x 3 blocks): the terms summed in segments of 16, a fixed
    # order that equals the CPU's to rounding
    monkeypatch.setenv("PLAMATRIX_AMD_SCHUR_ASSEMBLY", kernel)
    inst = shuffled_instance(24, 400, (3, 9), seed=21)
    b = device_batch(inst, dev)
    ne = capi.assemble_normal_equations(sw, 24, 400, 6, 3, **b, weights=t(inst.weights, dev, np.float64))
    dp = torch.zeros(24 * 6, dtype=torch.float64, device=dev)
    de = torch.zeros(400 * 3, dtype=torch.float64, device=dev)
    rep = capi.schur_solve(sw, ne, inst.damping, dp, de, ws, max_iterations=500, relative_tolerance=1e-12,
                           absolute_tolerance=0.0)
    assert rep.converged and rep.assembly_on_device and not rep.pattern_reused
    eq = oracle_equations(oracle, inst, inst.weights)
    pat = eq.pattern()
    a, rhs_ptr, pre_ptr = capi.schur_system(sw)
    n = 24 * 6
    assert a.nnz == len(pat["col_indices"])
    assert np.array_equal(capi.device_array(a.row_offsets, n + 1, "int64"), pat["row_offsets"])
    assert np.array_equal(capi.device_array(a.col_indices, a.nnz, "int64"), pat["col_indices"])
    assert np.array_equal(capi.device_array(a.row_offsets32, n + 1, "int32"), pat["row_offsets"])
    ref_values = eq.csr_values(inst.damping).values
    got_values = capi.device_array(a.values, a.nnz)
    if kernel == "slots":
        # same terms in the same (eliminated, left, right) order and association
        assert np.array_equal(got_values.view(np.uint64), ref_values.view(np.uint64))
    else:
        assert np.max(np.abs(got_values - ref_values)) <= 1e-13 * np.max(np.abs(ref_values))
    B, Cinv, rhs, Pinv = eq.prologue(inst.damping)
    assert np.array_equal(capi.device_array(rhs_ptr, n).view(np.uint64), rhs.view(np.uint64))
    # the preconditioner comes from the assembled diagonal block (terms in
    # eliminated order) rather than creation order: equal to rounding
    np.testing.assert_allclose(capi.device_array(pre_ptr, 24 * 36), Pinv, rtol=1e-9, atol=1e-12)
    relres, _, _ = reduced_relres(sw, dp)
    assert relres < 1e-10


@pytest.mark.parametrize("pc,ec,window,seed", [(12, 60, (3, 6), 7), (40, 2000, (3, 12), 3)])
def test_solve_matches_oracle(dev, sw, ws, oracle, pc, ec, window, seed):
    inst = shuffled_instance(pc, ec, window, seed)
    b = device_batch(inst, dev)
    ne = capi.assemble_normal_equations(sw, pc, ec, 6, 3, **b, weights=t(inst.weights, dev, np.float64))
    dp = torch.zeros(pc * 6, dtype=torch.float64, device=dev)
    de = torch.zeros(ec * 3, dtype=torch.float64, device=dev)
    rep = capi.schur_solve(sw, ne, inst.damping, dp, de, ws, max_iterations=1000, relative_tolerance=1e-10,
                           absolute_tolerance=0.0)
    assert rep.converged and rep.message == "converged"
    eq = oracle_equations(oracle, inst, inst.weights)
    rdp, rde, rrep = eq.solve(inst.damping, max_iterations=1000, relative_tolerance=1e-10, absolute_tolerance=0.0)
    assert rrep.converged
    got = np.concatenate([dp.cpu().numpy(), de.cpu().numpy()])
    ref = np.concatenate([rdp, rde])
    assert np.linalg.norm(got - ref) <= 1e-6 * np.linalg.norm(ref)
    relres, _, _ = reduced_relres(sw, dp)
    assert relres <= 1e-10 * 1.01
    assert abs(rep.iterations - rrep.iterations) <= 3


def test_pattern_reuse_and_rebuild(dev, sw, ws, oracle):
    # block_schur_test.cpp:160-196: same topology reuses the pattern
    inst = shuffled_instance(10, 80, (3, 6), seed=4)
    b = device_batch(inst, dev)
    dp = torch.zeros(60, dtype=torch.float64, device=dev)
    de = torch.zeros(240, dtype=torch.float64, device=dev)
    ne = capi.assemble_normal_equations(sw, 10, 80, 6, 3, **b, weights=t(inst.weights, dev, np.float64))
    r1 = capi.schur_solve(sw, ne, 1e-3, dp, de, ws)
    first = dp.clone()
    r2 = capi.schur_solve(sw, ne, 1e-3, dp, de, ws)
    assert not r1.pattern_reused and r2.pattern_reused
    assert sw.pattern_build_count == 1
    assert torch.equal(first, dp)
    # different damping: same pattern, new values
    r3 = capi.schur_solve(sw, ne, 1e-1, dp, de, ws)
    assert r3.pattern_reused and sw.pattern_build_count == 1
    eq = oracle_equations(oracle, inst, inst.weights)
    rdp, _, _ = eq.solve(1e-1)
    assert np.linalg.norm(dp.cpu().numpy() - rdp) <= 1e-6 * np.linalg.norm(rdp)
    # a changed topology rebuilds
    inst2 = shuffled_instance(10, 80, (3, 6), seed=5)
    ne2 = capi.assemble_normal_equations(sw, 10, 80, 6, 3, **device_batch(inst2, dev),
                                         weights=t(inst2.weights, dev, np.float64))
    r4 = capi.schur_solve(sw, ne2, 1e-3, dp, de, ws)
    assert not r4.pattern_reused and sw.pattern_build_count == 2


def test_host_equations_with_primary_cross(dev, sw, ws, oracle):
    # multi-primary residuals produce primary cross blocks (block_schur_test.cpp:77-115)
    eq = oracle.NormalEquations(2, 1, 1, 1)
    eq.add_residual_blocks([0, 1], [np.array([1.0]), np.array([2.0])], 0, np.array([3.0]), np.array([4.0]))
    eq.add_primary_residual_blocks([0, 1], [np.array([0.5]), np.array([-1.0])], np.array([0.2]))
    ne = capi.upload_normal_equations(2, 1, 1, 1, eq.export(), device=dev)
    dp = torch.zeros(2, dtype=torch.float64, device=dev)
    de = torch.zeros(1, dtype=torch.float64, device=dev)
    rep = capi.schur_solve(sw, ne, 0.1, dp, de, ws, max_iterations=50, relative_tolerance=1e-13,
                           absolute_tolerance=1e-14)
    assert rep.converged
    np.testing.assert_allclose(dp.cpu().numpy(), [-1.345185185185185, -0.512592592592593], atol=1e-11)
    np.testing.assert_allclose(de.cpu().numpy(), [-0.493827160493827], atol=1e-11)


def test_synthetic_reference_system(dev, sw, ws, oracle):
    # block_schur_test.cpp:41-75, 198-233 (primary-only and eliminated-only residuals too)
    eq = oracle.NormalEquations(2, 2, 2, 1)
    for p, e, pj, ej, r in ((0, 0, [1.0, 2.0], 0.5, -1.0), (1, 0, [-0.5, 1.0], 1.5, 0.25),
                            (0, 1, [2.0, -1.0], -0.75, 2.0), (1, 1, [1.25, 0.5], 2.0, -0.5)):
        eq.add_residual_block(p, e, np.array(pj), np.array([ej]), np.array([r]), 1.0)
    eq.add_primary_residual_block(0, np.array([0.4, -0.2]), np.array([0.75]), 2.0)
    eq.add_eliminated_residual_block(1, np.array([0.8]), np.array([-0.3]), 1.5)
    ne = capi.upload_normal_equations(2, 2, 2, 1, eq.export(), device=dev)
    dp = torch.zeros(4, dtype=torch.float64, device=dev)
    de = torch.zeros(2, dtype=torch.float64, device=dev)
    rep = capi.schur_solve(sw, ne, 0.1, dp, de, ws, max_iterations=50, relative_tolerance=1e-13,
                           absolute_tolerance=1e-14)
    assert rep.converged
    expected = [-0.532196868252671, 0.717080672008074, 0.078633780070238, -0.210823022930113,
                0.017900956933130, 0.253200998620430]
    np.testing.assert_allclose(np.concatenate([dp.cpu().numpy(), de.cpu().numpy()]), expected, atol=1e-11)


def test_point_only_system(dev, sw, ws, oracle):
    # block_schur_test.cpp:531-548: no primary blocks
    eq = oracle.NormalEquations(0, 1, 6, 1)
    eq.add_eliminated_residual_block(0, np.array([2.0]), np.array([4.0]))
    ne = capi.upload_normal_equations(0, 1, 6, 1, eq.export(), device=dev)
    dp = torch.zeros(1, dtype=torch.float64, device=dev)
    de = torch.zeros(1, dtype=torch.float64, device=dev)
    rep = capi.schur_solve(sw, ne, 0.5, dp, de, ws)
    assert rep.converged
    assert abs(de.cpu().numpy()[0] + 4.0 / 3.0) < 1e-12


def test_not_positive_definite_and_options(dev, sw, ws, oracle):
    eq = oracle.NormalEquations(1, 1, 1, 1)
    eq.add_residual_block(0, 0, np.array([1.0]), np.array([0.0]), np.array([1.0]))
    ne = capi.upload_normal_equations(1, 1, 1, 1, eq.export(), device=dev)
    dp = torch.zeros(1, dtype=torch.float64, device=dev)
    de = torch.zeros(1, dtype=torch.float64, device=dev)
    # C = 0, damped with the minimum diagonal 1e-12 * 0 -> stays non-positive at damping 0
    rep = capi.schur_solve(sw, ne, 0.0, dp, de, ws)
    assert not rep.converged and rep.status == 1
    assert rep.message == "eliminated block is not positive definite"
    with pytest.raises(capi.PlamatrixError) as e:
        capi.schur_solve(sw, ne, -1.0, dp, de, ws)
    assert e.value.kind == "invalid_argument"
    with pytest.raises(capi.PlamatrixError) as e:
        capi.schur_solve(sw, ne, 0.1, dp, dp, ws)
    assert e.value.kind == "invalid_argument"


def test_iteration_limit_reported(dev, sw, ws, oracle):
    inst = shuffled_instance(20, 200, (3, 8), seed=9)
    ne = capi.assemble_normal_equations(sw, 20, 200, 6, 3, **device_batch(inst, dev),
                                        weights=t(inst.weights, dev, np.float64))
    dp = torch.zeros(120, dtype=torch.float64, device=dev)
    de = torch.full((600,), 7.0, dtype=torch.float64, device=dev)
    rep = capi.schur_solve(sw, ne, inst.damping, dp, de, ws, max_iterations=2, relative_tolerance=1e-14,
                           absolute_tolerance=0.0)
    assert not rep.converged and rep.status == 4 and rep.iterations == 2
    assert rep.message == "PCG iteration limit reached"
    assert torch.count_nonzero(de).item() == 0  # eliminated step stays zero without convergence


def test_full_size_lm_system(dev, sw, ws):
    # 2,000 primaries x 200,000 secondaries, 1.2M residuals: size-independent checks
    inst = generators.lm_schur_instance()
    b = device_batch(inst, dev)
    ne = capi.assemble_normal_equations(sw, 2000, 200000, 6, 3, **b, weights=None, huber_delta=inst.huber_delta)
    assert ne.cross_count() == len(inst.primary)  # every (primary, secondary) pair is seen once
    dp = torch.zeros(12000, dtype=torch.float64, device=dev)
    de = torch.zeros(600000, dtype=torch.float64, device=dev)
    rep = capi.schur_solve(sw, ne, inst.damping, dp, de, ws, max_iterations=2000, relative_tolerance=1e-6,
                           absolute_tolerance=0.0)
    assert rep.converged, rep.message
    relres, S, rhs = reduced_relres(sw, dp)
    assert relres <= 1e-6 * 1.01
    slots, values, terms = sw.pattern_size()
    assert values == slots * 36 and S.nnz == values
    # symmetry of the assembled reduced system
    assert abs(S - S.T).max() <= 1e-9 * abs(S).max()


@pytest.mark.parametrize("size", ["small", "full"])
def test_segmented_assembly_matches_slot_assembly(dev, ws, monkeypatch, size):
    # the segmented kernels (terms in segments of 16, one wavefront each,
    # then the segment sums in order; slots below the block diagonal take
    # their mirror's sums transposed) against the slot kernel (the CPU order):
    # equal to rounding, symmetric bit for bit, and bitwise the same from run
    # to run, at a ring with long diagonal slots and at the full LM size
    if size == "small":
        inst = shuffled_instance(40, 3000, (3, 12), seed=5)
        pc, ec = 40, 3000
    else:
        inst = generators.lm_schur_instance()
        pc, ec = 2000, 200000
    b = device_batch(inst, dev)
    got = {}
    for kernel in ("segments", "slots", "segments-again"):
        monkeypatch.setenv("PLAMATRIX_AMD_SCHUR_ASSEMBLY", kernel.split("-")[0])
        sw = capi.SchurWorkspace()
        ne = capi.assemble_normal_equations(sw, pc, ec, 6, 3, **b, weights=None, huber_delta=inst.huber_delta)
        rep = capi.schur_assemble(sw, ne, inst.damping)
        assert rep.status == 0, rep.message
        a, rhs_ptr, pre_ptr = capi.schur_system(sw)
        got[kernel] = (capi.device_array(a.values, a.nnz), capi.device_array(pre_ptr, pc * 36))
        if kernel == "segments":
            import scipy.sparse as sp

            ro = capi.device_array(a.row_offsets, a.rows + 1, "int64")
            ci = capi.device_array(a.col_indices, a.nnz, "int64")
            S = sp.csr_matrix((got[kernel][0], ci, ro), shape=(a.rows, a.rows))
            assert (S != S.T).nnz == 0
        del sw, ne
    seg, slot = got["segments"], got["slots"]
    assert np.max(np.abs(seg[0] - slot[0])) <= 1e-13 * np.max(np.abs(slot[0]))
    np.testing.assert_allclose(seg[1], slot[1], rtol=1e-9, atol=1e-12 * np.max(np.abs(slot[1])))
    again = got["segments-again"]
    assert np.array_equal(seg[0].view(np.uint64), again[0].view(np.uint64))
    assert np.array_equal(seg[1].view(np.uint64), again[1].view(np.uint64))


@pytest.mark.parametrize("ps,es", [(9, 3), (12, 10), (16, 16)])
def test_large_blocks_match_oracle_and_lapack(dev, sw, ws, oracle, ps, es):
    # blocks above 8 x 8 (9-parameter cameras; the reference test at
    # block_schur_test.cpp:341-405 uses 9 x 3): slot entries are assembled 64
    # at a time; both device backends against the oracle's CPU solve and the
    # dense solution of the assembled reduced system
    rng = np.random.default_rng(ps * 100 + es)
    pc, ec = 10, 40
    eq = oracle.NormalEquations(pc, ec, ps, es)
    for p in range(pc):
        eq.add_primary_residual_block(p, np.eye(ps).ravel(), np.zeros(ps), 0.25)
    for e in range(ec):
        for obs in range(4):
            p = (e * 7 + obs * 5) % pc
            eq.add_residual_block(p, e, rng.standard_normal(2 * ps) * 0.1, rng.standard_normal(2 * es) * 0.2,
                                  rng.standard_normal(2) * 0.01, 1.0)
    ne = capi.upload_normal_equations(pc, ec, ps, es, eq.export(), device=dev)
    rdp, rde, rrep = eq.solve(0.1, max_iterations=500, relative_tolerance=1e-12, absolute_tolerance=0.0)
    assert rrep.converged
    ref = np.concatenate([rdp, rde])
    for dense in (False, True):
        dp = torch.zeros(pc * ps, dtype=torch.float64, device=dev)
        de = torch.zeros(ec * es, dtype=torch.float64, device=dev)
        rep = capi.schur_solve(sw, ne, 0.1, dp, de, ws, max_iterations=500, relative_tolerance=1e-12,
                               absolute_tolerance=0.0, dense=dense)
        assert rep.converged, rep.message
        got = np.concatenate([dp.cpu().numpy(), de.cpu().numpy()])
        assert np.linalg.norm(got - ref) <= 1e-8 * np.linalg.norm(ref)
        relres, S, rhs = reduced_relres(sw, dp)
        assert relres <= 1e-11
        x = np.linalg.solve(S.toarray(), rhs)
        assert np.linalg.norm(dp.cpu().numpy() - x) <= 1e-9 * np.linalg.norm(x)


# ---- primary-only residuals (addPrimaryTerms) on the device -------------------------


def primary_batch(pc, count, k, rs, ps, seed):
    """count residuals over k distinct primaries each (consecutive ring
    neighbours, some in reversed order, some repeated), N(0,1) Jacobians."""
    rng = np.random.default_rng(seed)
    ids = np.empty((count, k), np.int64)
    for i in range(count):
        start = int(rng.integers(0, pc))
        blocks = [(start + 2 * j) % pc for j in range(k)]
        if i % 3 == 1:
            blocks = blocks[::-1]
        ids[i] = blocks
    jac = rng.standard_normal((count, k, rs, ps))
    res = rng.standard_normal((count, rs)) * 0.5
    res[::7] *= 20.0  # outliers for the Huber weight
    weights = rng.uniform(0.2, 2.0, count)
    return ids, jac, res, weights


@pytest.mark.parametrize("huber", [False, True])
def test_primary_terms_bitwise(dev, sw, oracle, huber):
    inst = shuffled_instance(16, 300, (3, 8), seed=12)
    ne = capi.assemble_normal_equations(sw, 16, 300, 6, 3, **device_batch(inst, dev),
                                        weights=t(inst.weights, dev, np.float64))
    eq = oracle_equations(oracle, inst, inst.weights)
    for k, count, seed in ((2, 40, 3), (3, 12, 4)):
        ids, jac, res, weights = primary_batch(16, count, k, 2, 6, seed)
        if huber:
            w_ref = np.array([oracle.huber(squared_norm(r), 1.0)[1] for r in res])
            ne = capi.add_primary_terms(sw, ne, t(ids, dev), t(jac, dev), t(res, dev), weights=None, huber_delta=1.0)
        else:
            w_ref = weights
            ne = capi.add_primary_terms(sw, ne, t(ids, dev), t(jac, dev), t(res, dev), weights=t(weights, dev))
        for i in range(count):
            eq.add_primary_residual_blocks(ids[i], [jac[i, a] for a in range(k)], res[i], w_ref[i])
    got = capi.normal_equations_arrays(ne)
    ref = eq.export()
    assert len(ref["primary_cross_rows"]) > 0
    for key in ("primary_diagonal", "eliminated_diagonal", "primary_gradient", "eliminated_gradient",
                "cross_primary", "cross_eliminated", "cross_values", "primary_cross_rows", "primary_cross_cols",
                "primary_cross_values"):
        assert got[key].shape == ref[key].shape, key
        assert np.array_equal(got[key].view(np.uint64), ref[key].view(np.uint64)), key


def test_primary_terms_solve_matches_oracle(dev, sw, ws, oracle):
    # the primary cross blocks enter the reduced system: the device solve of
    # the augmented equations agrees with the oracle's CPU solve
    inst = shuffled_instance(20, 400, (3, 8), seed=13)
    ne = capi.assemble_normal_equations(sw, 20, 400, 6, 3, **device_batch(inst, dev),
                                        weights=t(inst.weights, dev, np.float64))
    ids, jac, res, weights = primary_batch(20, 30, 2, 6, 6, seed=5)
    ne = capi.add_primary_terms(sw, ne, t(ids, dev), t(jac, dev), t(res, dev), weights=t(weights, dev))
    eq = oracle_equations(oracle, inst, inst.weights)
    for i in range(len(ids)):
        eq.add_primary_residual_blocks(ids[i], [jac[i, 0], jac[i, 1]], res[i], weights[i])
    dp = torch.zeros(20 * 6, dtype=torch.float64, device=dev)
    de = torch.zeros(400 * 3, dtype=torch.float64, device=dev)
    rep = capi.schur_solve(sw, ne, inst.damping, dp, de, ws, max_iterations=1000, relative_tolerance=1e-10,
                           absolute_tolerance=0.0)
    assert rep.converged
    rdp, rde, rrep = eq.solve(inst.damping, max_iterations=1000, relative_tolerance=1e-10, absolute_tolerance=0.0)
    assert rrep.converged
    got = np.concatenate([dp.cpu().numpy(), de.cpu().numpy()])
    refv = np.concatenate([rdp, rde])
    assert np.linalg.norm(got - refv) <= 1e-6 * np.linalg.norm(refv)
    relres, _, _ = reduced_relres(sw, dp)
    assert relres <= 1e-10 * 1.01


def test_primary_terms_errors(dev, sw):
    inst = shuffled_instance(6, 20, (3, 4), seed=14)
    ne = capi.assemble_normal_equations(sw, 6, 20, 6, 3, **device_batch(inst, dev),
                                        weights=t(inst.weights, dev, np.float64))
    before = capi.normal_equations_arrays(ne)
    ids, jac, res, weights = primary_batch(6, 5, 2, 2, 6, seed=6)

    def call(i, j, r, w):
        return capi.add_primary_terms(sw, ne, t(i, dev), t(j, dev), t(r, dev), weights=t(w, dev))

    bad = ids.copy()
    bad[3, 1] = 6
    with pytest.raises(capi.PlamatrixError) as e:
        call(bad, jac, res, weights)
    assert e.value.kind == "out_of_range" and "residual 3" in str(e.value)
    bad = ids.copy()
    bad[2, 1] = bad[2, 0]
    with pytest.raises(capi.PlamatrixError) as e:
        call(bad, jac, res, weights)
    assert e.value.kind == "invalid_argument" and "unique" in str(e.value)
    badj = jac.copy()
    badj[1, 0, 1, 2] = np.nan
    with pytest.raises(capi.PlamatrixError) as e:
        call(ids, badj, res, weights)
    assert e.value.kind == "invalid_argument" and "finite" in str(e.value) and "residual 1" in str(e.value)
    badw = weights.copy()
    badw[4] = -1.0
    with pytest.raises(capi.PlamatrixError) as e:
        call(ids, jac, res, badw)
    assert e.value.kind == "invalid_argument" and "weight" in str(e.value)
    # the lowest offending residual decides: a NaN in residual 1 before a bad id in residual 3
    bad = ids.copy()
    bad[3, 1] = 6
    with pytest.raises(capi.PlamatrixError) as e:
        call(bad, badj, res, weights)
    assert "finite" in str(e.value)
    # the input equations are untouched
    after = capi.normal_equations_arrays(ne)
    for key in before:
        assert np.array_equal(before[key], after[key]), key
