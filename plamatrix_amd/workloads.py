"""Device-resident batches of generated SPD systems, shared by the benchmark
and the GPU tests.

batch_systems()  uneven systems with sizes log-uniform in [lo, hi] (seed),
                 cycling 2-D Laplacian / 3-D variable-coefficient diffusion /
                 reduced LM Schur system (primaries = n / 6, 3 secondaries per
                 primary, assembled on the device by the Schur pipeline),
                 packed as the diagonal blocks of one FP32 CSR with 32-bit
                 indices, each starting at a multiple of 64 rows.
reduced_schur()  the reduced Schur system of one LM instance, assembled on
                 the device, as device int64 / FP64 CSR arrays.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

KINDS = ("laplace2d", "diffusion3d", "schur")


def reduced_schur(capi, torch, inst, dev):
    """(rows, row_offsets int64, col_indices int64, values FP64) of the reduced
    Schur system of LM instance `inst`, assembled on the device."""
    t = lambda arr, dt: torch.from_numpy(np.ascontiguousarray(arr, dtype=dt)).to(dev)
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
    rows = a.rows
    ro = torch.empty(rows + 1, dtype=torch.int64, device=dev)
    ci = torch.empty(a.nnz, dtype=torch.int64, device=dev)
    va = torch.empty(a.nnz, dtype=torch.float64, device=dev)
    capi.lib_memcpy(ro.data_ptr(), a.row_offsets, ro.numel() * 8)
    capi.lib_memcpy(ci.data_ptr(), a.col_indices, ci.numel() * 8)
    capi.lib_memcpy(va.data_ptr(), a.values, va.numel() * 8)
    torch.cuda.synchronize()
    del sw, ne
    return rows, ro, ci, va


@dataclass
class Batch:
    a: object                 # DeviceCSR of the block-diagonal matrix (FP32, int32 mirrors)
    begin: np.ndarray         # row_begin per system
    end: np.ndarray           # row_end per system
    kinds: list
    lens: list
    nnzs: list
    members: list = field(repr=False)  # ("host", CSRHost) or ("device", (rows, ro, ci, values))

    def host_csr(self, i, dtype=np.float64):
        """(row_offsets, col_indices, values) of system i on the host, with the
        values the batch solves (rounded to FP32), widened to `dtype`."""
        where, m = self.members[i]
        if where == "host":
            return m.row_offsets, m.col_indices, m.values.astype(np.float32).astype(dtype)
        _, ro, ci, va = m
        return ro.cpu().numpy(), ci.cpu().numpy(), va.cpu().numpy().astype(np.float32).astype(dtype)


def pack_members(capi, torch, members, dev, pad_identity=False):
    """Block-diagonal device CSR (FP32, int32) of host / device members."""
    lens = [m[1].rows if m[0] == "host" else m[1][0] for m in members]
    begin, total = capi.batch_layout(lens)
    nnzs = [int(m[1].nnz) if m[0] == "host" else int(m[1][2].numel()) for m in members]
    counts = torch.zeros(total, dtype=torch.int64, device=dev)
    pad = torch.ones(total, dtype=torch.bool, device=dev)
    for (where, m), b0, ln in zip(members, begin, lens):
        ro = torch.from_numpy(m.row_offsets).to(dev) if where == "host" else m[1]
        counts[b0:b0 + ln] = ro[1:] - ro[:-1]
        pad[b0:b0 + ln] = False
    if pad_identity:
        counts[pad] = 1
    ro_all = torch.zeros(total + 1, dtype=torch.int64, device=dev)
    ro_all[1:] = torch.cumsum(counts, 0)
    nnz_total = int(ro_all[-1].item())
    ci_all = torch.empty(nnz_total, dtype=torch.int32, device=dev)
    va_all = torch.empty(nnz_total, dtype=torch.float32, device=dev)
    for (where, m), b0, ln, z in zip(members, begin, lens, nnzs):
        if where == "host":
            ci = torch.from_numpy(m.col_indices).to(dev)
            va = torch.from_numpy(m.values.astype(np.float32)).to(dev)
        else:
            _, _, ci, va = m
            va = va.to(torch.float32)
        k = int(ro_all[b0].item())
        ci_all[k:k + z] = (ci + int(b0)).to(torch.int32)
        va_all[k:k + z] = va
    if pad_identity:
        rows = torch.nonzero(pad).flatten()
        ci_all[ro_all[rows]] = rows.to(torch.int32)
        va_all[ro_all[rows]] = 1.0
    a = capi.DeviceCSR(total, total, None, None, va_all, ro_all.to(torch.int32), ci_all)
    end = begin + np.asarray(lens, dtype=np.int64)
    return a, begin, end, lens, nnzs


def batch_systems(capi, generators, torch, count=512, seed=4, lo=1000, hi=200000, pad_identity=False):
    """The uneven batch: sizes log-uniform in [lo, hi] (seed), members cycling
    2-D Laplacian / 3-D diffusion (seed 100 + i) / reduced LM Schur system
    (seed 5000 + i)."""
    dev = torch.device("cuda:0")
    sizes = generators.batch_sizes(count, seed=seed, lo=lo, hi=hi)
    members = []
    for i, n in enumerate(sizes):
        kind = KINDS[i % 3]
        if kind == "laplace2d":
            members.append(("host", generators.laplacian2d_csr(*generators.grid2d_dims(int(n)), dtype=np.float32)))
        elif kind == "diffusion3d":
            members.append(("host", generators.diffusion3d_box_csr(*generators.grid3d_dims(int(n)), seed=100 + i,
                                                                   dtype=np.float32)))
        else:
            pc = max(int(n) // 6, 20)
            inst = generators.lm_schur_instance(primary_count=pc, eliminated_count=3 * pc, seed=5000 + i)
            members.append(("device", reduced_schur(capi, torch, inst, dev)))
    a, begin, end, lens, nnzs = pack_members(capi, torch, members, dev, pad_identity=pad_identity)
    return Batch(a, begin, end, [KINDS[i % 3] for i in range(count)], lens, nnzs, members)
