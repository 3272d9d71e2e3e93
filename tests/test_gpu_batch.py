"""GPU parity of the batched PCG through the C ABI.

Each system of a batch must behave as if solved alone: its solution agrees
with the oracle's CPU PCG on that system within relative 1e-6 (FP64) / 1e-4
(FP32), its iteration count with the single-system GPU solve, and its FP64
true residual is recomputed; stop conditions (convergence, iteration limit,
bad diagonal, zero right-hand side) are per system.
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


def mixed_systems(count, seed, lo=200, hi=6000, dtype=np.float64):
    sizes = generators.batch_sizes(count, seed=seed, lo=lo, hi=hi)
    systems = []
    for i, n in enumerate(sizes):
        if i % 2 == 0:
            systems.append(generators.laplacian2d_csr(*generators.grid2d_dims(int(n)), dtype=dtype))
        else:
            systems.append(generators.diffusion3d_box_csr(*generators.grid3d_dims(int(n)), seed=100 + i,
                                                          dtype=dtype))
    return systems


def pack(systems, dev, dtype, seed=1):
    big, begin, end = generators.block_diagonal(systems, dtype=dtype)
    rng = np.random.default_rng(seed)
    b = np.zeros(big.rows, dtype=dtype)
    for s, b0 in zip(systems, begin):
        b[b0:b0 + s.rows] = rng.uniform(-1.0, 1.0, s.rows)
    a = capi.to_device_csr(big, compact=True)
    return a, begin, end, big, b


def relres(sys, b, x):
    import scipy.sparse as sp

    A = sp.csr_matrix((sys.values.astype(np.float64), sys.col_indices, sys.row_offsets), shape=(sys.rows, sys.rows))
    b = b.astype(np.float64)
    return np.linalg.norm(b - A @ x.astype(np.float64)) / np.linalg.norm(b)


@pytest.mark.parametrize("dtype,rtol", [(np.float64, 1e-6), (np.float32, 1e-4)])
def test_batch_matches_oracle_per_system(dev, ws, oracle, dtype, rtol):
    systems = mixed_systems(12, seed=7, dtype=dtype)
    a, begin, end, big, b = pack(systems, dev, dtype)
    bt = torch.from_numpy(b).to(dev)
    x = torch.zeros(big.rows, dtype=bt.dtype, device=dev)
    tol = 1e-8 if dtype == np.float64 else 1e-5
    reps = capi.batch_pcg(a, begin, end, bt, x, ws, relative_tolerance=tol, max_iterations=3000)
    xs = x.cpu().numpy()
    for s, b0, e0, rep in zip(systems, begin, end, reps):
        assert rep.converged and rep.status == 1, rep
        ref_sys = oracle.CSR(s.rows, s.rows, s.row_offsets, s.col_indices, s.values)
        xr, rr = oracle.pcg(ref_sys, b[b0:e0], relative_tolerance=tol, max_iterations=3000)
        assert rr.converged
        got = xs[b0:e0].astype(np.float64)
        assert np.linalg.norm(got - xr) <= rtol * np.linalg.norm(xr)
        if dtype == np.float64:
            assert abs(rep.iterations - rr.iterations) <= 2
        assert relres(s, b[b0:e0], got) <= tol * 10


def test_batch_equals_single_solves(dev, ws):
    systems = mixed_systems(6, seed=11, dtype=np.float64)
    a, begin, end, big, b = pack(systems, dev, np.float64)
    bt = torch.from_numpy(b).to(dev)
    x = torch.zeros(big.rows, dtype=torch.float64, device=dev)
    reps = capi.batch_pcg(a, begin, end, bt, x, ws, relative_tolerance=1e-10, max_iterations=5000)
    for s, b0, e0, rep in zip(systems, begin, end, reps):
        single = capi.to_device_csr(s)
        xs = torch.zeros(s.rows, dtype=torch.float64, device=dev)
        r1 = capi.pcg(single, bt[b0:e0].contiguous(), xs, ws, relative_tolerance=1e-10, max_iterations=5000)
        # same recurrence; only the dot-product partition differs
        assert abs(rep.iterations - r1.iterations) <= 1
        xb, x1 = x[b0:e0].cpu().numpy(), xs.cpu().numpy()
        assert np.linalg.norm(xb - x1) <= 1e-6 * np.linalg.norm(x1)


def test_batch_per_system_stops(dev, ws):
    lap = generators.laplacian2d_csr(20, 20)
    bad = generators.laplacian2d_csr(10, 10)
    bad.values = bad.values.copy()
    assert bad.col_indices[0] == 0  # row 0 starts with its diagonal
    bad.values[0] = 0.0
    systems = [lap, generators.laplacian2d_csr(30, 30), bad, generators.laplacian2d_csr(8, 8)]
    a, begin, end, big, b = pack(systems, dev, np.float64)
    b[begin[3]:end[3]] = 0.0  # zero right-hand side: converged at iteration 0
    bt = torch.from_numpy(b).to(dev)
    x = torch.zeros(big.rows, dtype=torch.float64, device=dev)
    reps = capi.batch_pcg(a, begin, end, bt, x, ws, relative_tolerance=1e-10, max_iterations=25)
    assert reps[0].status == 2 and reps[0].iterations == 25  # 20x20 needs more than 25
    assert reps[1].status == 2
    assert reps[2].status == 6 and reps[2].bad_row == begin[2]
    assert reps[3].status == 1 and reps[3].iterations == 0
    assert torch.count_nonzero(x[begin[2]:end[2]]).item() == 0  # never updated
    reps = capi.batch_pcg(a, begin[:2], end[:2], bt, x.zero_(), ws, relative_tolerance=1e-10, max_iterations=2000)
    assert all(r.converged for r in reps)


def test_batch_cg_refresh_and_gaps(dev, ws):
    systems = mixed_systems(5, seed=3, dtype=np.float32)
    a, begin, end, big, b = pack(systems, dev, np.float32)
    bt = torch.from_numpy(b).to(dev)
    x = torch.zeros(big.rows, dtype=torch.float32, device=dev)
    pol = capi.policy(refresh_interval=10, chunk_iterations=10)
    reps = capi.batch_pcg(a, begin, end, bt, x, ws, pol=pol, relative_tolerance=1e-5, max_iterations=4000,
                          use_jacobi=False)
    for s, b0, e0, rep in zip(systems, begin, end, reps):
        assert rep.converged
        assert relres(s, b[b0:e0], x[b0:e0].cpu().numpy()) <= 1e-4
    # a subset of the systems (others are ignored rows)
    x.zero_()
    reps = capi.batch_pcg(a, begin[1::2], end[1::2], bt, x, ws, relative_tolerance=1e-5, max_iterations=4000)
    assert all(r.converged for r in reps)
    assert torch.count_nonzero(x[begin[0]:end[0]]).item() == 0


def test_batch_rejects_bad_layout(dev, ws):
    systems = mixed_systems(2, seed=1)
    a, begin, end, big, b = pack(systems, dev, np.float64)
    bt = torch.from_numpy(b).to(dev)
    x = torch.zeros(big.rows, dtype=torch.float64, device=dev)
    with pytest.raises(capi.PlamatrixError) as e:
        capi.batch_pcg(a, begin + 1, end, bt, x, ws)
    assert e.value.kind == "invalid_argument"
    with pytest.raises(capi.PlamatrixError) as e:
        capi.batch_pcg(a, begin[::-1].copy(), end[::-1].copy(), bt, x, ws)
    assert e.value.kind == "invalid_argument"


def test_batch_refinement_reaches_fp64_target(dev, ws, oracle):
    # FP32 matrices and inner solves, FP64 x and residuals: every system must
    # reach an FP64 true relative residual below the target the FP32
    # recurrence alone cannot reach, and agree with the oracle's FP64 solve
    systems = mixed_systems(9, seed=21, dtype=np.float32)
    systems.append(generators.laplacian2d_csr(16, 16, dtype=np.float32))
    a, begin, end, big, b = pack(systems, dev, np.float32)
    b64 = b.astype(np.float64)
    b64[begin[-1]:end[-1]] = 0.0  # zero right-hand side: converged before any inner solve
    bt = torch.from_numpy(b64).to(dev)
    x = torch.zeros(big.rows, dtype=torch.float64, device=dev)
    target = 1e-10
    reps = capi.batch_pcg_refine(a, begin, end, bt, x, ws, target=target, max_outer=20, relative_tolerance=1e-3,
                                 max_iterations=5000)
    xs = x.cpu().numpy()
    for s, b0, e0, rep in zip(systems[:-1], begin, end, reps):
        assert rep.converged and rep.final_relres <= target, rep
        assert rep.outer_iterations >= 2 and rep.inner_iterations > 0
        got = xs[b0:e0]
        true = relres(s, b64[b0:e0], got)
        assert true <= target * 1.01
        assert abs(true - rep.final_relres) <= 5e-2 * rep.final_relres + 1e-14
        ref_sys = oracle.CSR(s.rows, s.rows, s.row_offsets, s.col_indices, s.values.astype(np.float64))
        xr, rr = oracle.pcg(ref_sys, b64[b0:e0], relative_tolerance=1e-12, max_iterations=20000)
        assert rr.converged
        assert np.linalg.norm(got - xr) <= 1e-6 * np.linalg.norm(xr)
    last = reps[-1]
    assert last.converged and last.outer_iterations == 0 and last.inner_iterations == 0
    assert torch.count_nonzero(x[begin[-1]:end[-1]]).item() == 0
    # a warm start at the solution takes no outer step
    reps2 = capi.batch_pcg_refine(a, begin, end, bt, x, ws, target=1e-8, max_outer=20, relative_tolerance=1e-3)
    assert all(r.converged and r.outer_iterations == 0 for r in reps2)


def test_batch_refinement_outer_limit(dev, ws):
    systems = mixed_systems(3, seed=5, dtype=np.float32)
    a, begin, end, big, b = pack(systems, dev, np.float32)
    bt = torch.from_numpy(b.astype(np.float64)).to(dev)
    x = torch.zeros(big.rows, dtype=torch.float64, device=dev)
    reps = capi.batch_pcg_refine(a, begin, end, bt, x, ws, target=1e-12, max_outer=1, relative_tolerance=1e-2,
                                 max_iterations=2000)
    for rep in reps:
        assert not rep.converged and rep.outer_iterations == 1
        assert rep.final_relres < rep.initial_relres
    with pytest.raises(capi.PlamatrixError) as e:
        capi.batch_pcg_refine(a, begin, end, bt, x, ws, target=-1.0)
    assert e.value.kind == "invalid_argument"


# ---- block-Jacobi systems in a batch -------------------------------------------------


def schur_members(dev, counts, seed0=7000):
    """Reduced LM Schur systems (6x6 block CSR, device-assembled), as host CSR."""
    from plamatrix_amd import workloads

    out = []
    for i, pc in enumerate(counts):
        inst = generators.lm_schur_instance(primary_count=pc, eliminated_count=3 * pc, seed=seed0 + i)
        rows, ro, ci, va = workloads.reduced_schur(capi, torch, inst, dev)
        out.append(generators.CSRHost(rows, rows, ro.cpu().numpy(), ci.cpu().numpy(), va.cpu().numpy()))
    return out


def host_block_inverses(sys, bs, dtype):
    import scipy.sparse as sp

    m = sp.csr_matrix((sys.values.astype(np.float64), sys.col_indices, sys.row_offsets), shape=(sys.rows, sys.rows))
    nb = sys.rows // bs
    inv = np.empty((nb, bs, bs))
    for k in range(nb):
        inv[k] = np.linalg.inv(m[k * bs:(k + 1) * bs, k * bs:(k + 1) * bs].toarray())
    return inv.astype(dtype).reshape(-1)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batch_block_jacobi_matches_single_solves(dev, ws, dtype):
    # Schur members with 6x6 block-Jacobi beside Jacobi Laplacian / diffusion
    # members: each system behaves as the single-system solve of its kind
    # (blockPcg with the same block inverses, or Jacobi-PCG)
    schur = schur_members(dev, [40, 90])
    systems = [schur[0], generators.laplacian2d_csr(17, 19, dtype=np.float64), schur[1],
               generators.diffusion3d_box_csr(9, 8, 7, seed=5, dtype=np.float64)]
    systems = [generators.CSRHost(s.rows, s.cols, s.row_offsets, s.col_indices, s.values.astype(dtype))
               for s in systems]
    blocks = [6, 1, 6, 1]
    a, begin, end, big, b = pack(systems, dev, dtype)
    bt = torch.from_numpy(b).to(dev)
    x = torch.zeros(big.rows, dtype=bt.dtype, device=dev)
    rel = 1e-10 if dtype == np.float64 else 1e-5
    reps = capi.batch_block_pcg(a, begin, end, blocks, bt, x, ws, relative_tolerance=rel, max_iterations=5000)
    for s, bs, b0, e0, rep in zip(systems, blocks, begin, end, reps):
        assert rep.converged, rep
        single = capi.to_device_csr(s, compact=True)
        xs = torch.zeros(s.rows, dtype=bt.dtype, device=dev)
        rhs = bt[b0:e0].contiguous()
        if bs > 1:
            inv = torch.from_numpy(host_block_inverses(s, bs, dtype)).to(dev)
            r1 = capi.block_pcg(single, rhs, xs, inv, bs, ws, relative_tolerance=rel, max_iterations=5000)
        else:
            r1 = capi.pcg(single, rhs, xs, ws, relative_tolerance=rel, max_iterations=5000)
        assert r1.converged
        # same recurrence; the inverses (device Gauss-Jordan vs host LAPACK)
        # and the dot-product partition differ in rounding only
        assert abs(rep.iterations - r1.iterations) <= 2, (bs, rep.iterations, r1.iterations)
        xb, x1 = x[b0:e0].cpu().numpy().astype(np.float64), xs.cpu().numpy().astype(np.float64)
        tol = 1e-7 if dtype == np.float64 else 1e-3
        assert np.linalg.norm(xb - x1) <= tol * np.linalg.norm(x1)
        assert relres(s, b[b0:e0], xb) <= (2e-10 if dtype == np.float64 else 5e-5)
    # block-Jacobi needs fewer iterations than Jacobi on the Schur members
    xj = torch.zeros_like(x)
    repj = capi.batch_pcg(a, begin, end, bt, xj, ws, relative_tolerance=rel, max_iterations=5000)
    assert reps[0].iterations < repj[0].iterations and reps[2].iterations < repj[2].iterations
    # all block sizes 1 is the Jacobi batch, bit for bit
    x1 = torch.zeros_like(x)
    reps1 = capi.batch_block_pcg(a, begin, end, [1, 1, 1, 1], bt, x1, ws, relative_tolerance=rel,
                                 max_iterations=5000)
    assert [r.iterations for r in reps1] == [r.iterations for r in repj]
    assert torch.equal(x1, xj)


def test_batch_block_jacobi_rejects_bad_block_and_layout(dev, ws):
    lap = generators.laplacian2d_csr(12, 12)
    bad = generators.laplacian2d_csr(10, 10)
    bad.values = bad.values.copy()
    # rows 4, 5 form block 2 of size 2; a negative diagonal breaks its SPD pivots
    k = bad.row_offsets[4] + int(np.nonzero(bad.col_indices[bad.row_offsets[4]:bad.row_offsets[5]] == 4)[0][0])
    bad.values[k] = -4.0
    systems = [lap, bad]
    a, begin, end, big, b = pack(systems, dev, np.float64)
    bt = torch.from_numpy(b).to(dev)
    x = torch.zeros(big.rows, dtype=torch.float64, device=dev)
    reps = capi.batch_block_pcg(a, begin, end, [2, 2], bt, x, ws, relative_tolerance=1e-10, max_iterations=2000)
    assert reps[0].converged
    assert reps[1].status == 6 and reps[1].bad_row == begin[1] + 4
    assert torch.count_nonzero(x[begin[1]:end[1]]).item() == 0
    with pytest.raises(capi.PlamatrixError) as e:  # 144 rows are not a multiple of 5
        capi.batch_block_pcg(a, begin, end, [5, 2], bt, x, ws)
    assert e.value.kind == "invalid_argument"
    with pytest.raises(capi.PlamatrixError) as e:
        capi.batch_block_pcg(a, begin, end, [2, 9], bt, x, ws)
    assert e.value.kind == "invalid_argument"


def test_batch_block_jacobi_refinement_reaches_target(dev, ws, oracle):
    # FP32 inner solves with 6x6 block-Jacobi on the Schur members, FP64 x:
    # every system reaches the FP64 target and the oracle's FP64 solution
    schur = schur_members(dev, [50, 120], seed0=7100)
    systems = [schur[0], generators.laplacian2d_csr(20, 20, dtype=np.float64), schur[1]]
    systems = [generators.CSRHost(s.rows, s.cols, s.row_offsets, s.col_indices, s.values.astype(np.float32))
               for s in systems]
    a, begin, end, big, b = pack(systems, dev, np.float32)
    b64 = b.astype(np.float64)
    bt = torch.from_numpy(b64).to(dev)
    x = torch.zeros(big.rows, dtype=torch.float64, device=dev)
    reps = capi.batch_block_pcg_refine(a, begin, end, [6, 1, 6], bt, x, ws, target=1e-9, max_outer=20,
                                       relative_tolerance=1e-3, max_iterations=5000)
    xs = x.cpu().numpy()
    for s, b0, e0, rep in zip(systems, begin, end, reps):
        assert rep.converged and rep.final_relres <= 1e-9, rep
        got = xs[b0:e0]
        assert relres(s, b64[b0:e0], got) <= 1.01e-9
        ref_sys = oracle.CSR(s.rows, s.rows, s.row_offsets, s.col_indices, s.values.astype(np.float64))
        xr, rr = oracle.pcg(ref_sys, b64[b0:e0], relative_tolerance=1e-12, max_iterations=20000)
        assert rr.converged
        assert np.linalg.norm(got - xr) <= 1e-6 * np.linalg.norm(xr)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batch_single_reduction_matches_classic(dev, monkeypatch, dtype):
    # Jacobi, CG and block-Jacobi batches run the single-reduction form
    # (step, SpMV of u with four partials, per-system reduce); against the
    # direction / update form on the same batch, with residual refreshes:
    # every system ends in the same state, within 2 % (at least 2) of its
    # iterations, at a true residual within the solve's tolerance
    schur = schur_members(dev, [40, 90])
    systems = mixed_systems(6, seed=31, dtype=np.float64) + schur
    systems = [generators.CSRHost(s.rows, s.cols, s.row_offsets, s.col_indices, s.values.astype(dtype))
               for s in systems]
    blocks = [1] * 6 + [6, 6]
    a, begin, end, big, b = pack(systems, dev, dtype)
    bt = torch.from_numpy(b).to(dev)
    tol = 1e-9 if dtype == np.float64 else 1e-5
    pol = capi.policy(refresh_interval=20)
    for kind in ("cg", "jacobi", "block"):
        out = {}
        for flag in ("0", "1"):
            monkeypatch.setenv("PLAMATRIX_AMD_BATCH_SINGLE_REDUCTION", flag)
            w = capi.Workspace()
            x = torch.zeros(big.rows, dtype=bt.dtype, device=dev)
            if kind == "block":
                reps = capi.batch_block_pcg(a, begin, end, blocks, bt, x, w, pol=pol, relative_tolerance=tol,
                                            max_iterations=4000)
            else:
                reps = capi.batch_pcg(a, begin, end, bt, x, w, pol=pol, use_jacobi=kind == "jacobi",
                                      relative_tolerance=tol, max_iterations=4000)
            torch.cuda.synchronize()
            out[flag] = (reps, x.cpu().numpy())
        (r0, x0), (r1, x1) = out["0"], out["1"]
        for s, b0, e0, c, d in zip(systems, begin, end, r0, r1):
            assert c.converged and d.converged and c.status == d.status, (kind, c, d)
            assert abs(c.iterations - d.iterations) <= max(2, 0.02 * c.iterations), (kind, s.rows, c, d)
            for xs in (x0, x1):
                assert relres(s, b[b0:e0], xs[b0:e0]) <= (10 if dtype == np.float64 else 40) * tol, (kind, s.rows)



def test_batch_refinement_final_pass_aims_below_target(dev, ws):
    # inner tolerance 1e-3 and target 1e-6: a second pass would land a hair
    # either side of the target, so a pass that can end the refinement aims at
    # 0.9 x target / relres; every system then ends in two passes, at most
    # 0.95 of the target (0.9 and the inner solve's small recursive-residual
    # gap), with the true FP64 residual
    systems = mixed_systems(16, seed=41, dtype=np.float32)
    a, begin, end, big, b = pack(systems, dev, np.float32)
    b64 = b.astype(np.float64)
    bt = torch.from_numpy(b64).to(dev)
    x = torch.zeros(big.rows, dtype=torch.float64, device=dev)
    reps = capi.batch_pcg_refine(a, begin, end, bt, x, ws, target=1e-6, max_outer=10, relative_tolerance=1e-3,
                                 max_iterations=5000)
    xs = x.cpu().numpy()
    for s, b0, e0, rep in zip(systems, begin, end, reps):
        assert rep.converged and rep.outer_iterations == 2, (s.rows, rep)
        assert rep.final_relres <= 0.95e-6, (s.rows, rep)
        assert relres(s, b64[b0:e0], xs[b0:e0]) <= 0.96e-6


def test_refinement_final_pass_aims_below_target(dev, ws):
    # the single-system refinement follows the same rule
    h = generators.laplacian2d_csr(300, 300, dtype=np.float32)
    a = capi.to_device_csr(h, compact=True)
    rng = np.random.default_rng(43)
    b64 = rng.standard_normal(h.rows)
    bt = torch.from_numpy(b64).to(dev)
    x = torch.zeros(h.rows, dtype=torch.float64, device=dev)
    rep = capi.pcg_refine(a, bt, x, ws, target=1e-6, max_outer=10, relative_tolerance=1e-3, max_iterations=20000)
    assert rep.converged and rep.outer_iterations == 2 and rep.final_relres <= 0.95e-6, rep
    assert relres(h, b64, x.cpu().numpy()) <= 0.96e-6
