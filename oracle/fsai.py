"""CPU restatement of the learned factorized preconditioner (test
infrastructure only; the network has no counterpart in the reference).

Restates, in float64 torch on the CPU, what plamatrix_amd/csrc/learned.hip
computes: node features, normalised couplings, the lower pattern, the
message-passing network with the same BF16 rounding points (inputs, weights,
activations rounded to bfloat16 with round-to-nearest-even, gradients passed
straight through the rounding), the decoder, the probe generator and the
probe loss ``(1/P) sum_p ||(I - G A G^T) z_p||^2 / ||z_p||^2``.  Gradients
come from torch autograd, which the tests pin against central finite
differences of this same loss.
"""
from __future__ import annotations

import numpy as np

HID, FEAT, LAYERS, MAX_PROBES = 64, 32, 3, 8
OFF_WIN = 0
OFF_BIN = OFF_WIN + FEAT * HID
OFF_WL = OFF_BIN + HID
OFF_BL = OFF_WL + LAYERS * 2 * HID * HID
OFF_WOFF = OFF_BL + LAYERS * HID
OFF_BOFF = OFF_WOFF + HID
OFF_ALPHA = OFF_BOFF + 1
OFF_WDIAG = OFF_ALPHA + 1
OFF_BDIAG = OFF_WDIAG + HID
PARAMS = OFF_BDIAG + 1


def graph(ro, ci, vals):
    """dinv, couplings s (CSR order, 0 on the diagonal), features, lower pattern."""
    ro = np.asarray(ro, np.int64)
    ci = np.asarray(ci, np.int64)
    v32 = np.asarray(vals).astype(np.float32)
    vd = np.asarray(vals).astype(np.float64)
    n = len(ro) - 1
    rows = np.repeat(np.arange(n), np.diff(ro))
    diag = np.zeros(n)
    np.add.at(diag, rows[ci == rows], vd[ci == rows])
    dinv = (1.0 / np.sqrt(diag)).astype(np.float32)
    off = ci != rows
    s = np.zeros(len(ci), np.float32)
    s[off] = (v32[off] * dinv[rows[off]]) * dinv[ci[off]]
    feats = np.zeros((n, FEAT), np.float32)
    for i in range(n):
        k = slice(ro[i], ro[i + 1])
        mask = off[k]
        si = s[k][mask].astype(np.float32)
        deg = int(mask.sum())
        tot = np.float32(0)
        sabs = np.float32(0)
        for x in si:
            tot = np.float32(tot + x)
            sabs = np.float32(sabs + abs(x))
        feats[i, 0] = 1.0
        feats[i, 1] = np.float32(deg) / np.float32(8)
        feats[i, 2] = tot
        feats[i, 3] = sabs
        feats[i, 4] = si.max() if deg else 0.0
        feats[i, 5] = si.min() if deg else 0.0
        feats[i, 6] = np.abs(si).max() if deg else 0.0
        feats[i, 7] = tot / sabs if sabs > 0 else 0.0
        feats[i, 8] = np.float32(np.log(diag[i] if diag[i] > 0 else 1.0)) / np.float32(8)
        feats[i, 9] = sabs * sabs
    lower = ci <= rows
    return dict(n=n, ro=ro, ci=ci, rows=rows, vals=vd, dinv=dinv, s=s, feats=feats,
                l_rows=rows[lower], l_cols=ci[lower], l_src=np.nonzero(lower)[0])


def _mix64(x):
    x = x ^ (x >> np.uint64(30))
    x = x * np.uint64(0xBF58476D1CE4E5B9)
    x = x ^ (x >> np.uint64(27))
    x = x * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def probes(n, count, seed):
    """z[i][p] of the device generator (index t = i * count + p)."""
    with np.errstate(over="ignore"):
        t = np.arange(n * count, dtype=np.uint64)
        h1 = _mix64(np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + t * np.uint64(2) + np.uint64(1))
        h2 = _mix64(h1 ^ np.uint64(0xD1B54A32D192ED03))
    u1 = ((h1 >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)
    u2 = (h2 >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    z = np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)
    return z.astype(np.float32).reshape(n, count)


def _torch():
    import torch

    return torch


def _round(x, emulate):
    torch = _torch()
    if not emulate:
        return x
    return x + (x.to(torch.bfloat16).to(torch.float64) - x).detach()


def forward(gr, params, emulate_bf16=True):
    """G values on the lower pattern (torch float64, differentiable in params)."""
    torch = _torch()
    p = params
    n = gr["n"]
    X = _round(torch.tensor(gr["feats"], dtype=torch.float64), emulate_bf16)
    W_in = _round(p[OFF_WIN:OFF_BIN].reshape(FEAT, HID), emulate_bf16)
    H = _round(torch.relu(X @ W_in + p[OFF_BIN:OFF_WL]), emulate_bf16)
    rows = torch.tensor(gr["rows"])
    cols = torch.tensor(gr["ci"])
    s = torch.tensor(gr["s"], dtype=torch.float64)
    S = torch.zeros((n, n), dtype=torch.float64)
    S[rows, cols] = s  # diagonal entries carry 0
    for layer in range(LAYERS):
        M = _round(S @ H, emulate_bf16)
        base = OFF_WL + layer * 2 * HID * HID
        W = _round(p[base:base + 2 * HID * HID].reshape(2 * HID, HID), emulate_bf16)
        b = p[OFF_BL + layer * HID:OFF_BL + (layer + 1) * HID]
        H = _round(torch.relu(torch.cat([H, M], dim=1) @ W + b), emulate_bf16)
    lr = torch.tensor(gr["l_rows"])
    lc = torch.tensor(gr["l_cols"])
    dinv = torch.tensor(gr["dinv"], dtype=torch.float64)
    sl = s[torch.tensor(gr["l_src"])]
    w_off, w_diag = p[OFF_WOFF:OFF_BOFF], p[OFF_WDIAG:OFF_BDIAG]
    is_diag = lr == lc
    out_off = (H[lr] * H[lc] * w_off).sum(dim=1) + p[OFF_BOFF] + p[OFF_ALPHA] * sl
    out_diag = (H[lr] * w_diag).sum(dim=1) + p[OFF_BDIAG]
    out = torch.where(is_diag, out_diag, out_off)
    g = torch.where(is_diag, torch.nn.functional.softplus(out_diag) * dinv[lr], out_off * dinv[lc])
    return g, out


def loss(gr, params, probe_count, seed, emulate_bf16=True):
    torch = _torch()
    n = gr["n"]
    g, _ = forward(gr, params, emulate_bf16)
    G = torch.zeros((n, n), dtype=torch.float64)
    G[torch.tensor(gr["l_rows"]), torch.tensor(gr["l_cols"])] = g
    A = torch.zeros((n, n), dtype=torch.float64)
    A[torch.tensor(gr["rows"]), torch.tensor(gr["ci"])] = torch.tensor(gr["vals"], dtype=torch.float64)
    Z = torch.tensor(probes(n, probe_count, seed), dtype=torch.float64)
    R = Z - G @ (A @ (G.T @ Z))
    return ((R * R).sum(dim=0) / (probe_count * (Z * Z).sum(dim=0))).sum()


def loss_and_grad(ro, ci, vals, params, probe_count=8, seed=0, emulate_bf16=True):
    torch = _torch()
    gr = graph(ro, ci, vals)
    p = torch.tensor(np.asarray(params, np.float64), requires_grad=True)
    value = loss(gr, p, probe_count, seed, emulate_bf16)
    value.backward()
    return float(value.detach()), p.grad.numpy().copy()


def factors(ro, ci, vals, params, emulate_bf16=True):
    """Lower pattern (rows, cols) and G values."""
    torch = _torch()
    gr = graph(ro, ci, vals)
    with torch.no_grad():
        g, _ = forward(gr, torch.tensor(np.asarray(params, np.float64)), emulate_bf16)
    return gr["l_rows"], gr["l_cols"], g.numpy()


def factored_pcg(ro, ci, vals, g_rows, g_cols, g_vals, rhs, relative_tolerance=1e-8, max_iterations=1000):
    """FP64 PCG with z = G^T G r and the solver's stop rule max(abs, rel ||r0||)
    on the recursive residual (iterative_solver_cpu.cpp:218-270 with the
    preconditioner application replaced)."""
    import scipy.sparse as sp

    n = len(ro) - 1
    A = sp.csr_matrix((np.asarray(vals, np.float64), np.asarray(ci), np.asarray(ro)), shape=(n, n))
    G = sp.csr_matrix((np.asarray(g_vals, np.float64), (g_rows, g_cols)), shape=(n, n))
    b = np.asarray(rhs, np.float64)
    x = np.zeros(n)
    r = b - A @ x
    rr0 = float(r @ r)
    tol = max(0.0, relative_tolerance * np.sqrt(rr0))
    if np.sqrt(rr0) <= tol:
        return x, 0
    z = G.T @ (G @ r)
    p = z.copy()
    rz = float(r @ z)
    for it in range(1, max_iterations + 1):
        q = A @ p
        alpha = rz / float(p @ q)
        x += alpha * p
        r -= alpha * q
        if np.sqrt(float(r @ r)) <= tol:
            return x, it
        z = G.T @ (G @ r)
        rz_new = float(r @ z)
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x, max_iterations
