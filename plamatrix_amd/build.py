"""Builds the gfx950 HIP library ``plamatrix_amd/libplamatrix_amd.so`` in tree.

Every ``.hip`` file under ``plamatrix_amd/csrc`` is compiled with hipcc for
gfx950 (packed-FP32 instructions disabled) and linked into one shared library
exporting the C ABI declared in ``include/plamatrix_amd.h``.  The C++ API
test program ``tests/cpp/test_api`` (host code over the C++ headers in
``include/plamatrix``) is linked against it.
"""
from __future__ import annotations

import concurrent.futures
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBRARY = os.path.join(HERE, "libplamatrix_amd.so")
ARCH = os.environ.get("PLAMATRIX_AMD_ARCH", "gfx950")

COMMON_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + CSRC,
    "-Wall",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [hipcc(), *COMMON_FLAGS, "-c", src, "-o", obj]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{proc.stderr}")
    if proc.stderr.strip():
        noise = [l for l in proc.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if noise:
            sys.stderr.write("\n".join(noise) + "\n")
    return obj


def build(jobs: int | None = None) -> str:
    os.makedirs(BUILD, exist_ok=True)
    sources = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    jobs = jobs or min(len(sources), max(1, (os.cpu_count() or 4)))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        objects = list(pool.map(_compile, sources))
    tmp = LIBRARY + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objects]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stderr}")
    os.replace(tmp, LIBRARY)
    build_cpp_tests()
    return LIBRARY


CPP_TESTS = {
    # drop-in API over the device (run by a GPU test)
    os.path.join(ROOT, "tests", "cpp", "test_api.cpp"): os.path.join(ROOT, "tests", "cpp", "test_api"),
    # host overloads of the drop-in API (no GPU call; run by a CPU test)
    os.path.join(ROOT, "tests", "cpp", "test_host_api.cpp"): os.path.join(ROOT, "tests", "cpp", "test_host_api"),
}
CPP_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "test_api")


def build_cpp_tests() -> str:
    """Host C++ programs over the drop-in headers, linked to the library."""
    for src, binary in CPP_TESTS.items():
        cmd = [hipcc(), "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), src,
               "-L" + HERE, "-lplamatrix_amd", "-Wl,-rpath,$ORIGIN/../../plamatrix_amd", "-o", binary]
        proc = subprocess.run(cmd, capture_output=True, text=True)
        if proc.returncode != 0:
            raise RuntimeError(f"C++ API test build failed for {src}:\n{proc.stderr}")
    return CPP_TEST_BIN


if __name__ == "__main__":
    print(build())
