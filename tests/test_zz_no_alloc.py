"""The allocation-free SpMM (pm_spmm_noalloc) through the C ABI: same bits as
pm_spmm on two calls in a row (the second on a smaller matrix, so stale packed
columns of the first call would show), in FP32 and FP64 with 32- and 64-bit
indices; matrices above the scratch limit are rejected with PM_OVERFLOW."""
import ctypes

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


def spmm_noalloc(a, b, y):
    k = int(b.shape[0])
    d = a.descriptor()
    capi.check(capi.lib().pm_spmm_noalloc(ctypes.byref(d), capi._ptr(b), capi._i64(k), capi._ptr(y),
                                          capi.current_stream()))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("compact", [True, False])
def test_noalloc_spmm_matches_spmm(dev, dtype, compact):
    ws = capi.Workspace()
    rng = np.random.default_rng(3)
    cases = [(generators.diffusion3d_box_csr(40, 36, 30, seed=5, dtype=dtype), 11),
             (generators.laplacian2d_csr(50, 47, dtype=dtype), 6)]
    for h, k in cases:
        a = capi.to_device_csr(h, compact=compact)
        b = torch.from_numpy(rng.standard_normal((k, h.cols)).astype(dtype)).to(dev)
        y_ref = torch.full((k, h.rows), np.nan, dtype=b.dtype, device=dev)
        y = torch.full((k, h.rows), np.nan, dtype=b.dtype, device=dev)
        capi.spmm(a, b, y_ref, ws)
        spmm_noalloc(a, b, y)
        torch.cuda.synchronize()
        view = np.uint32 if dtype == np.float32 else np.uint64
        assert np.array_equal(y.cpu().numpy().view(view), y_ref.cpu().numpy().view(view)), (h.rows, k)


def test_noalloc_spmm_rejects_oversized(dev):
    n = (64 << 20) // (8 * 8) + 1  # FP64 columns one past the scratch limit
    ro = torch.zeros(2, dtype=torch.int64, device=dev)
    ci = torch.zeros(0, dtype=torch.int64, device=dev)
    va = torch.zeros(0, dtype=torch.float64, device=dev)
    a = capi.DeviceCSR(1, n, ro, ci, va)
    b = torch.zeros((1, n), dtype=torch.float64, device=dev)
    y = torch.zeros((1, 1), dtype=torch.float64, device=dev)
    with pytest.raises(capi.PlamatrixError) as e:
        spmm_noalloc(a, b, y)
    assert e.value.code == 3 and "scratch" in str(e.value)
