"""Host overloads of the C++ drop-in API (include/plamatrix/sparse/
host_sparse.h): cooToCsr on std::vector triplets, spmv on CPU matrices and
cg / pcg on CPU matrices.

- tests/cpp/test_host_api runs the reference's own CPU known-answer and
  exception cases (sparse_ops_test.cpp:21-292, iterative_solver_test.cpp:
  103-335, 782-820);
- its dump mode converts, multiplies and solves generated systems, and every
  result is compared bit for bit with the oracle (itself bitwise equal to the
  reference build): CSR arrays and values, y = A x, and the CG / Jacobi-PCG
  reports and solutions.
No GPU call is made."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from plamatrix_amd import generators

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_host_api")


@pytest.fixture(scope="module")
def binary():
    if not os.path.exists(BIN):
        pytest.skip("tests/cpp/test_host_api is not built (run __graft_entry__.build())")
    return BIN


def test_reference_cpu_cases(binary):
    proc = subprocess.run([binary], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0 and "ALL PASSED" in proc.stdout, proc.stdout + proc.stderr


def _dump(binary, tmp_path, rows, cols, r, c, v, x, b, max_iterations, rel):
    dt = v.dtype
    path_in, path_out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(path_in, "wb") as f:
        np.array([rows, cols, len(r), 1 if dt == np.float64 else 0], np.int64).tofile(f)
        np.asarray(r, np.int64).tofile(f)
        np.asarray(c, np.int64).tofile(f)
        np.asarray(v, dt).tofile(f)
        np.asarray(x, dt).tofile(f)
        np.asarray(b, dt).tofile(f)
        np.array([max_iterations], np.int64).tofile(f)
        np.array([rel], np.float64).tofile(f)
    proc = subprocess.run([binary, "dump", str(path_in), str(path_out)], capture_output=True, text=True, timeout=300)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    raw = open(path_out, "rb").read()
    pos = 0

    def take(n, t):
        nonlocal pos
        a = np.frombuffer(raw, dtype=t, count=n, offset=pos)
        pos += a.nbytes
        return a

    out = {}
    nnz = int(take(1, np.int64)[0])
    out["ro"], out["ci"], out["v"] = take(rows + 1, np.int64), take(nnz, np.int64), take(nnz, dt)
    out["y"] = take(rows, dt)
    if rows == cols:
        for name in ("cg", "pcg"):
            head = take(2, np.int64)
            res = take(2, np.float64)
            out[name] = (bool(head[0]), int(head[1]), float(res[0]), float(res[1]), take(rows, dt))
    return out


def shuffled_coo(rng, a, duplicate_fraction=0.3):
    """Triplets of CSR `a` in random order, some split into 2-4 parts (large
    parts that cancel, so the left fold in input order is order-dependent at
    rounding level and the matrix stays SPD)."""
    rows = np.repeat(np.arange(a.rows), np.diff(a.row_offsets))
    r, c, v = [], [], []
    for i in range(len(a.values)):
        parts = rng.integers(2, 5) if rng.random() < duplicate_fraction else 1
        noise = list(rng.standard_normal(parts - 1) * 10.0)
        vals = [a.values[i] - sum(noise)] + noise
        for p in vals:
            r.append(rows[i])
            c.append(a.col_indices[i])
            v.append(p)
    order = rng.permutation(len(v))
    return (np.array(r, np.int64)[order], np.array(c, np.int64)[order],
            np.array(v, dtype=a.values.dtype)[order])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_dump_matches_oracle_bitwise(binary, tmp_path, dtype, seed):
    rng = np.random.default_rng(100 + seed)
    a = generators.diffusion3d_box_csr(9, 8, 6 + seed, seed=seed, dtype=dtype)
    r, c, v = shuffled_coo(rng, a)
    n = a.rows
    x = rng.standard_normal(n).astype(dtype)
    b = generators.rhs_uniform(n, seed=seed).astype(dtype)
    got = _dump(binary, tmp_path, n, n, r, c, v, x, b, 400, 1e-6 if dtype == np.float32 else 1e-10)
    ref = oracle.coo_to_csr(n, n, r, c, v)
    assert np.array_equal(got["ro"], ref.row_offsets) and np.array_equal(got["ci"], ref.col_indices)
    assert got["v"].tobytes() == np.asarray(ref.values, dtype).tobytes()
    assert got["y"].tobytes() == oracle.spmv(ref, x).astype(dtype).tobytes()
    for name, solve in (("cg", oracle.cg), ("pcg", oracle.pcg)):
        conv, its, r0, rf, sol = got[name]
        xr, rep = solve(ref, b, relative_tolerance=1e-6 if dtype == np.float32 else 1e-10, max_iterations=400)
        assert (conv, its, r0, rf) == (rep.converged, rep.iterations, rep.initial_residual, rep.final_residual)
        assert sol.tobytes() == np.asarray(xr, dtype).tobytes()


def test_dump_rectangular_and_empty_rows(binary, tmp_path):
    rng = np.random.default_rng(7)
    rows, cols, count = 40, 25, 300
    r = rng.integers(0, rows // 2, count) * 2  # odd rows stay empty
    c = rng.integers(0, cols, count)
    v = rng.standard_normal(count)
    v[::7] = 0.0  # explicit zeros are kept
    x = rng.standard_normal(cols)
    got = _dump(binary, tmp_path, rows, cols, r, c, v, x, np.zeros(rows), 1, 1e-6)
    ref = oracle.coo_to_csr(rows, cols, r, c, v)
    assert np.array_equal(got["ro"], ref.row_offsets) and np.array_equal(got["ci"], ref.col_indices)
    assert got["v"].tobytes() == ref.values.tobytes()
    assert got["y"].tobytes() == oracle.spmv(ref, x).tobytes()
