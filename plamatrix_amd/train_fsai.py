"""Train the learned factorized preconditioner on the GPU and evaluate it
against Jacobi-PCG.

Training systems come from the method's own generators with seeds disjoint
from evaluation (train 10^6 + i, eval 2 * 10^6 + i): 3-D variable-coefficient
diffusion boxes (the 128^3 generator, scaled down) and reduced LM Schur systems
(scaled down, assembled by the device Schur pipeline), alternating, each used
for one step by default (20,000 systems).  Each step draws 8 Gaussian probes
and takes one Adam step on
(1/P) sum_p ||(I - G A G^T) z_p||^2 / ||z_p||^2 with the library's own BF16
MFMA forward and backward kernels.

    python -m plamatrix_amd.train_fsai --systems 20000 --steps 20000 --out plamatrix_amd/fsai_weights.npy
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

DEFAULT_WEIGHTS = os.path.join(HERE, "fsai_weights.npy")


def diffusion_system(seed: int, lo: int = 16, hi: int = 34):
    from plamatrix_amd import generators

    rng = np.random.default_rng(seed)
    nx, ny, nz = (int(v) for v in rng.integers(lo, hi, 3))
    return generators.diffusion3d_box_csr(nx, ny, nz, seed=seed, dtype=np.float32)


def schur_system(seed: int, primaries: int = 300, ratio: int = 20, instance=None):
    """Reduced Schur system of a scaled-down LM instance, FP32 CSR on the host,
    assembled on the device."""
    import torch

    from plamatrix_amd import capi, generators

    inst = instance if instance is not None else generators.lm_schur_instance(
        primary_count=primaries, eliminated_count=primaries * ratio, seed=seed)
    dev = torch.device("cuda:0")
    t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)
    sw = capi.SchurWorkspace()
    ne = capi.assemble_normal_equations(sw, inst.primary_count, inst.eliminated_count, 6, 3,
                                        t(inst.primary, np.int64), t(inst.eliminated, np.int64),
                                        t(inst.primary_jacobians, np.float64),
                                        t(inst.eliminated_jacobians, np.float64), t(inst.residuals, np.float64),
                                        weights=None, huber_delta=inst.huber_delta)
    rep = capi.schur_assemble(sw, ne, inst.damping)
    if rep.status != 0:
        raise RuntimeError(rep.message)
    a, _, _ = capi.schur_system(sw)
    n = a.rows
    ro = capi.device_array(a.row_offsets, n + 1, "int64")
    ci = capi.device_array(a.col_indices, a.nnz, "int64")
    vals = capi.device_array(a.values, a.nnz).astype(np.float32)
    return generators.CSRHost(n, n, ro, ci, vals)


def make_system(kind: str, seed: int):
    return diffusion_system(seed) if kind == "diffusion" else schur_system(seed)


def evaluate(net, systems, rel=1e-6, max_iterations=20000, inner_rel=1e-3):
    """Jacobi-PCG vs learned-G PCG on each system: FP32 inner solves with FP64
    refinement to an FP64 true relative residual of `rel`; iterations (inner,
    summed over outer steps), wall time and the true residual.  The learned
    time includes the graph build and the network inference."""
    import torch

    from plamatrix_amd import capi, generators

    dev = torch.device("cuda:0")
    ws = capi.Workspace()
    pol = capi.policy(refresh_interval=50)
    inner = dict(relative_tolerance=inner_rel, max_iterations=max_iterations)
    rows = []
    for h in systems:
        a = capi.to_device_csr(h, compact=True)
        b = torch.from_numpy(generators.rhs_normal(h.rows, seed=h.rows).astype(np.float64)).to(dev)
        out = {}
        for name in ("jacobi", "learned"):
            x = torch.zeros_like(b)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if name == "jacobi":
                rep = capi.pcg_refine(a, b, x, ws, target=rel, max_outer=10, pol=pol, **inner)
            else:
                graph = capi.FsaiGraph(a)
                capi.fsai_forward(net, graph)
                rep = capi.pcg_refine_factored(a, b, x, graph, ws, target=rel, max_outer=10, pol=pol, **inner)
            torch.cuda.synchronize()
            out[name] = dict(iterations=rep.inner_iterations, outer=rep.outer_iterations, converged=rep.converged,
                             ms=(time.perf_counter() - t0) * 1e3, relres=capi.true_relres(a, b, x, ws))
        rows.append(out)
    return rows


def host_system(kind: str, seed: int):
    """Host-side part of a training system: a diffusion CSR, or an LM instance
    whose Schur system is assembled on the device by the consumer."""
    from plamatrix_amd import generators

    if kind == "diffusion":
        return diffusion_system(seed)
    return generators.lm_schur_instance(primary_count=300, eliminated_count=6000, seed=seed)


def device_system(kind: str, seed: int, host):
    if kind == "diffusion":
        return host
    return schur_system(seed, instance=host)


def train_seed(kind: str, i: int) -> int:
    return 1_000_000 + i + (50_000 if kind == "schur" else 0)


def train(steps: int, seed: int, lr: float, kinds: list[str], systems: int, log_every: int = 100):
    """One Adam step per training system: system i (kinds alternating, seeds
    10^6 + i, Schur seeds offset by 50,000) is generated on a host thread
    ahead of use, moved to the device, and used for `steps / systems` steps."""
    import queue
    import threading

    from plamatrix_amd import capi

    net = capi.FsaiNet(seed=seed)
    order = [(kinds[i % len(kinds)], i // len(kinds)) for i in range(systems)]
    ahead: queue.Queue = queue.Queue(maxsize=16)

    def produce():
        for kind, i in order:
            ahead.put((kind, i, host_system(kind, train_seed(kind, i))))
        ahead.put(None)

    threading.Thread(target=produce, daemon=True).start()
    per_system = max(1, steps // max(systems, 1))
    history = []
    t0 = time.perf_counter()
    running = []
    step = 0
    total = per_system * systems
    while True:
        item = ahead.get()
        if item is None:
            break
        kind, i, host = item
        a = capi.to_device_csr(device_system(kind, train_seed(kind, i), host), compact=True)
        graph = capi.FsaiGraph(a)
        for _ in range(per_system):
            # cosine decay to lr / 10
            rate = lr * (0.1 + 0.9 * 0.5 * (1.0 + math.cos(math.pi * step / max(total, 1))))
            running.append(capi.fsai_train_step(net, graph, probes=8, seed=step + 1, lr=rate))
            step += 1
            if step % log_every == 0 or step == total:
                history.append(dict(step=step, loss=float(np.mean(running)), seconds=time.perf_counter() - t0))
                print(json.dumps(history[-1]), flush=True)
                running = []
        del graph, a
    return net, history


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--systems", type=int, default=20000, help="training systems (kinds alternate)")
    ap.add_argument("--kinds", default="diffusion,schur")
    ap.add_argument("--eval", type=int, default=16, help="held-out systems per kind for the summary")
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    args = ap.parse_args()
    kinds = [k for k in args.kinds.split(",") if k]
    net, history = train(args.steps, args.seed, args.lr, kinds, args.systems, log_every=500)
    np.save(args.out, net.parameters)
    summary = {}
    for kind in kinds:
        held = [make_system(kind, 2_000_000 + i + (50_000 if kind == "schur" else 0)) for i in range(args.eval)]
        res = evaluate(net, held)
        summary[kind] = dict(
            jacobi_iterations=float(np.mean([r["jacobi"]["iterations"] for r in res])),
            learned_iterations=float(np.mean([r["learned"]["iterations"] for r in res])),
            jacobi_ms=float(np.mean([r["jacobi"]["ms"] for r in res])),
            learned_ms=float(np.mean([r["learned"]["ms"] for r in res])),
            learned_better=int(sum(r["learned"]["iterations"] < r["jacobi"]["iterations"] for r in res)),
            systems=len(res))
    print(json.dumps(dict(weights=args.out, final_loss=history[-1]["loss"], eval=summary)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
