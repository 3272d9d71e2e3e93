"""The C++ drop-in headers (include/plamatrix/) on the GPU: runs
tests/cpp/test_api, which checks the reference's known answers through
COOMatrix / CSRMatrix / cooToCsr / spmv / cg / pcg / blockPcg / the
fixed-iteration solves / solveDampedSchurComplement and the exception
classes they throw."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINARY = os.path.join(ROOT, "tests", "cpp", "test_api")


def test_cpp_api():
    assert os.path.exists(BINARY), "tests/cpp/test_api is not built; run __graft_entry__.build()"
    proc = subprocess.run([BINARY], capture_output=True, text=True, timeout=240)
    print(proc.stdout)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "ALL PASSED" in proc.stdout
