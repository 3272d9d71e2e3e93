"""GPU checks of the batched PCG with the learned factored preconditioner
(pm_batch_pcg_factored / pm_batch_pcg_refine_factored) and of the per-system
choice between the learned factor and Jacobi (pm_fsai_batch_select).

- a graph built on the block-diagonal batch gives every system the G of a
  graph built on that system alone, bit for bit (the network is local);
- each system of a batched factored solve agrees with the oracle's FP64 PCG
  with z = G^T G r (oracle/fsai.py factored_pcg) within relative 1e-6 (FP64) /
  1e-4 (FP32), iterations within 2 in FP64, FP64 true residual recomputed;
- systems flagged for Jacobi run exactly the batched Jacobi-PCG (same bits),
  and the other systems of a mixed batch are untouched by the flags;
- the selector's per-system losses agree with a host estimate of the same
  probe loss, the untrained network (G = D^-1/2) ties with Jacobi and
  falls back, and the trained network is kept on diffusion boxes;
- batched FP64 refinement with the learned factors reaches a true relres of
  1e-6 on every system.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import fsai  # noqa: E402
from plamatrix_amd import capi, generators  # noqa: E402

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "plamatrix_amd",
                       "fsai_weights.npy")


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


@pytest.fixture(scope="module")
def net(dev):
    n = capi.FsaiNet(seed=0)
    n.parameters = np.load(WEIGHTS)
    return n


def systems_mix(count, seed, dtype):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        if i % 3 == 2:
            nx, ny = (int(v) for v in rng.integers(12, 40, 2))
            out.append(generators.laplacian2d_csr(nx, ny, dtype=dtype))
        else:
            nx, ny, nz = (int(v) for v in rng.integers(6, 16, 3))
            out.append(generators.diffusion3d_box_csr(nx, ny, nz, seed=300 + i, dtype=dtype))
    return out


def pack(systems, dev, dtype, seed=1):
    big, begin, end = generators.block_diagonal(systems, dtype=dtype, pad_identity=True)
    rng = np.random.default_rng(seed)
    b = np.zeros(big.rows, dtype=np.float64)
    for s, b0 in zip(systems, begin):
        b[b0:b0 + s.rows] = rng.standard_normal(s.rows)
    return capi.to_device_csr(big, compact=True), begin, end, big, b


def csr_block(ro, ci, vals, b0, e0):
    """Rows [b0, e0) of a host CSR, columns shifted by -b0."""
    k0, k1 = ro[b0], ro[e0]
    return ro[b0:e0 + 1] - k0, ci[k0:k1] - b0, vals[k0:k1]


def relres(sys, b, x):
    import scipy.sparse as sp

    A = sp.csr_matrix((sys.values.astype(np.float64), sys.col_indices, sys.row_offsets), shape=(sys.rows, sys.rows))
    return float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))


def test_batched_graph_equals_per_system_graphs(dev, net):
    systems = systems_mix(7, seed=11, dtype=np.float32)
    a, begin, end, big, _ = pack(systems, dev, np.float32)
    graph = capi.FsaiGraph(a)
    capi.fsai_forward(net, graph)
    g, gt = graph.factors()
    gro, gci, gv = capi.csr_arrays(g)
    tro, tci, tv = capi.csr_arrays(gt)
    for s, b0, e0 in zip(systems, begin, end):
        one = capi.to_device_csr(s, compact=True)
        g1 = capi.FsaiGraph(one)
        capi.fsai_forward(net, g1)
        f, ft = g1.factors()
        for (ro, ci, v), d in (((gro, gci, gv), f), ((tro, tci, tv), ft)):
            bro, bci, bv = csr_block(ro, ci, v, b0, e0)
            sro, sci, sv = capi.csr_arrays(d)
            assert np.array_equal(bro, sro) and np.array_equal(bci, sci)
            assert np.array_equal(bv.view(np.uint32), sv.view(np.uint32))
        del g1


@pytest.mark.parametrize("dtype,rtol", [(np.float64, 1e-6), (np.float32, 1e-4)])
def test_batch_factored_matches_oracle(dev, ws, net, dtype, rtol):
    systems = systems_mix(9, seed=5, dtype=dtype)
    a, begin, end, big, b = pack(systems, dev, dtype)
    graph = capi.FsaiGraph(a)
    capi.fsai_forward(net, graph)
    g, _ = graph.factors()
    gro, gci, gv = capi.csr_arrays(g)
    bt = torch.from_numpy(b.astype(dtype)).to(dev)
    x = torch.zeros_like(bt)
    tol = 1e-8 if dtype == np.float64 else 1e-5
    reps = capi.batch_pcg_factored(a, begin, end, bt, x, graph, ws, relative_tolerance=tol, max_iterations=3000)
    xs = x.cpu().numpy().astype(np.float64)
    for s, b0, e0, rep in zip(systems, begin, end, reps):
        assert rep.converged and rep.status == 1, rep
        lro, lci, lv = csr_block(gro, gci, gv, b0, e0)
        lrows = np.repeat(np.arange(s.rows), np.diff(lro))
        bs = b[b0:e0].astype(dtype).astype(np.float64)
        xr, it = fsai.factored_pcg(s.row_offsets, s.col_indices, s.values, lrows, lci, lv, bs,
                                   relative_tolerance=tol, max_iterations=3000)
        got = xs[b0:e0]
        assert np.linalg.norm(got - xr) <= rtol * np.linalg.norm(xr), (s.rows, rep)
        if dtype == np.float64:
            assert abs(rep.iterations - it) <= 2, (rep.iterations, it)
        res = relres(s, bs, got)
        assert res <= (2.0 if dtype == np.float64 else 20.0) * tol, res
        print(f"n={s.rows} iterations gpu={rep.iterations} oracle={it} true relres={res:.2e}")


def test_jacobi_flags_run_batched_jacobi_bitwise(dev, ws, net, monkeypatch):
    # the factored batch keeps the direction / update form: its Jacobi-flagged
    # systems match the Jacobi batch in that form bit for bit
    monkeypatch.setenv("PLAMATRIX_AMD_BATCH_SINGLE_REDUCTION", "0")
    dtype = np.float32
    systems = systems_mix(8, seed=21, dtype=dtype)
    a, begin, end, big, b = pack(systems, dev, dtype)
    graph = capi.FsaiGraph(a)
    capi.fsai_forward(net, graph)
    bt = torch.from_numpy(b.astype(dtype)).to(dev)
    pol = capi.policy(refresh_interval=8)
    opts = dict(relative_tolerance=1e-5, max_iterations=2000)
    xj = torch.zeros_like(bt)
    rj = capi.batch_pcg(a, begin, end, bt, xj, ws, pol=pol, use_jacobi=True, **opts)
    xf = torch.zeros_like(bt)
    rf = capi.batch_pcg_factored(a, begin, end, bt, xf, graph, ws, pol=pol, **opts)
    flags = torch.tensor([i % 2 for i in range(len(systems))], dtype=torch.uint8, device=dev)
    xm = torch.zeros_like(bt)
    rm = capi.batch_pcg_factored(a, begin, end, bt, xm, graph, ws, jacobi=flags, pol=pol, **opts)
    allj = torch.ones(len(systems), dtype=torch.uint8, device=dev)
    xa = torch.zeros_like(bt)
    ra = capi.batch_pcg_factored(a, begin, end, bt, xa, graph, ws, jacobi=allj, pol=pol, **opts)
    xj, xf, xm, xa = (t.cpu().numpy().view(np.uint32) for t in (xj, xf, xm, xa))
    for i, (b0, e0) in enumerate(zip(begin, end)):
        assert np.array_equal(xa[b0:e0], xj[b0:e0]) and ra[i].iterations == rj[i].iterations
        ref, rrep = (xj, rj[i]) if i % 2 else (xf, rf[i])
        assert np.array_equal(xm[b0:e0], ref[b0:e0]), i
        assert rm[i].iterations == rrep.iterations and rm[i].converged
    # the learned factor needs fewer iterations than Jacobi on the diffusion boxes
    diffusion = [i for i in range(len(systems)) if i % 3 != 2]
    assert sum(rf[i].iterations for i in diffusion) < sum(rj[i].iterations for i in diffusion)


def host_losses(a_sys, lro, lci, lv, probes, seed):
    import scipy.sparse as sp

    n = a_sys.rows
    A = sp.csr_matrix((a_sys.values.astype(np.float64), a_sys.col_indices, a_sys.row_offsets), shape=(n, n))
    G = sp.csr_matrix((lv.astype(np.float64), lci, lro), shape=(n, n))
    d = 1.0 / np.sqrt(A.diagonal())
    z = np.random.default_rng(seed).standard_normal((n, probes))
    learned = float(np.sum((z - G @ (A @ (G.T @ z))) ** 2))
    jacobi = float(np.sum((z - d[:, None] * (A @ (d[:, None] * z))) ** 2))
    return learned, jacobi


def test_batch_select(dev, net):
    systems = systems_mix(9, seed=8, dtype=np.float32)
    a, begin, end, big, _ = pack(systems, dev, np.float32)
    fresh = capi.FsaiNet(seed=3)
    graph = capi.FsaiGraph(a)
    capi.fsai_forward(fresh, graph)
    flags, losses = capi.fsai_batch_select(graph, begin, end, probes=8, seed=5)
    # G = D^-1/2: the learned loss is Jacobi's, so every system falls back
    np.testing.assert_allclose(losses[:, 0], losses[:, 1], rtol=1e-5)
    assert flags.cpu().numpy().tolist() == [1] * len(systems)
    capi.fsai_forward(net, graph)
    flags, losses = capi.fsai_batch_select(graph, begin, end, probes=8, seed=5)
    flags = flags.cpu().numpy()
    g, _ = graph.factors()
    gro, gci, gv = capi.csr_arrays(g)
    for i, (s, b0, e0) in enumerate(zip(systems, begin, end)):
        lro, lci, lv = csr_block(gro, gci, gv, b0, e0)
        hl, hj = host_losses(s, lro, lci, lv, 8, seed=100 + i)
        # same loss, other probes: within the sampling error of 8 n terms
        assert abs(losses[i, 0] - hl) <= 0.15 * hl and abs(losses[i, 1] - hj) <= 0.15 * hj, (losses[i], hl, hj)
        assert flags[i] == (0 if losses[i, 0] < losses[i, 1] else 1)
    # trained on diffusion boxes: kept there
    assert all(flags[i] == 0 for i in range(len(systems)) if i % 3 != 2)


def test_batch_refine_factored_reaches_target(dev, ws, net):
    systems = systems_mix(12, seed=31, dtype=np.float32)
    a, begin, end, big, b = pack(systems, dev, np.float32)
    graph = capi.FsaiGraph(a)
    capi.fsai_forward(net, graph)
    flags, _ = capi.fsai_batch_select(graph, begin, end, probes=8, seed=1)
    b64 = torch.from_numpy(b).to(dev)
    x = torch.zeros_like(b64)
    reps = capi.batch_pcg_refine_factored(a, begin, end, b64, x, graph, ws, jacobi=flags, target=1e-6,
                                          max_outer=10, pol=capi.policy(refresh_interval=50),
                                          relative_tolerance=1e-3, max_iterations=5000)
    xs = x.cpu().numpy()
    for s, b0, e0, rep in zip(systems, begin, end, reps):
        res = relres(s, b[b0:e0], xs[b0:e0])
        assert rep.converged and rep.final_relres <= 1e-6, rep
        assert res <= 1.05e-6, res


def test_split_and_fused_step_are_bitwise(dev, ws, net, monkeypatch):
    # the step x += alpha p, r -= alpha q folded into the factor products
    # (short factor rows) or run as its own pass (long rows): same bits,
    # on a batch that holds reduced Schur systems
    from plamatrix_amd import train_fsai, workloads

    hosts = [train_fsai.make_system("diffusion", 2_000_100 + i) for i in range(2)]
    hosts += [train_fsai.make_system("schur", 2_050_100 + i) for i in range(2)]
    a, begin, end, lens, _ = workloads.pack_members(capi, torch, [("host", h) for h in hosts], dev,
                                                    pad_identity=True)
    graph = capi.FsaiGraph(a)
    capi.fsai_forward(net, graph)
    rng = np.random.default_rng(3)
    b = np.zeros(a.rows, dtype=np.float32)
    for b0, ln in zip(begin, lens):
        b[b0:b0 + ln] = rng.standard_normal(ln)
    bt = torch.from_numpy(b).to(dev)
    out = {}
    for fuse in ("0", "1"):
        monkeypatch.setenv("PLAMATRIX_AMD_FUSE_STEP", fuse)
        x = torch.zeros_like(bt)
        reps = capi.batch_pcg_factored(a, begin, end, bt, x, graph, ws, pol=capi.policy(refresh_interval=10),
                                       relative_tolerance=1e-5, max_iterations=2000)
        out[fuse] = (x.cpu().numpy().view(np.uint32), [r.iterations for r in reps])
        assert all(r.converged for r in reps)
    assert out["0"][1] == out["1"][1]
    assert np.array_equal(out["0"][0], out["1"][0])
