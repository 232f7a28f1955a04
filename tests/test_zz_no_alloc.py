"""flpr_xstats_noalloc (partials in the library's static device array)
against flpr_xstats (partials in a caller buffer), through the C-ABI, on two
calls in a row, the second with a smaller input; bit-for-bit equal, since both
run the same kernels with the same launch shape."""
import ctypes

import numpy as np
import pytest
import torch

from fastlpr_amd import _lib

pytestmark = pytest.mark.gpu


def _stats(fn, x, y, d, ny, n):
    out = torch.full((5 * d + 2 + ny,), np.nan, dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    lib = _lib.load()
    if fn == "flpr_xstats":
        part = torch.empty(512 * (5 * d + 2 + ny), dtype=torch.float64, device="cuda")
        rc = lib.flpr_xstats(x.data_ptr(), n, d, y.data_ptr() if ny else None, ny,
                             part.data_ptr(), out.data_ptr(), s)
    else:
        rc = lib.flpr_xstats_noalloc(x.data_ptr(), n, d, y.data_ptr() if ny else None, ny,
                                     out.data_ptr(), s)
    assert rc == 0, rc
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("d,ny", [(2, 1), (3, 32), (1, 0)])
def test_noalloc_matches_workspace_entry_point(d, ny):
    rs = np.random.default_rng(11 + d + ny)
    big, small = 300_000, 777
    xb = torch.as_tensor(rs.uniform(-2, 5, (big, d)), device="cuda")
    yb = torch.as_tensor(rs.standard_normal((big, max(ny, 1))), device="cuda")
    xb[5, 0] = float("nan")    # non-finite counts are part of the output
    # two calls in a row, the second smaller: the scratch keeps the first
    # call's partials, which the second must not see
    for n in (big, small):
        x = xb[:n].contiguous()
        y = yb[:n, :ny].contiguous() if ny else yb[:n, :1].contiguous()
        ref = _stats("flpr_xstats", x, y, d, ny, n)
        got = _stats("flpr_xstats_noalloc", x, y, d, ny, n)
        assert np.array_equal(ref, got, equal_nan=True), (n, ref, got)
        # and against the statistics computed directly
        xh = x.cpu().numpy()
        fin = np.isfinite(xh[:, 0])
        np.testing.assert_allclose(got[0], np.nanmin(xh[fin, 0]), rtol=0)
        np.testing.assert_allclose(got[4], fin.sum(), rtol=0)


def test_noalloc_rejects_sizes_beyond_scratch():
    x = torch.zeros((10, 2), dtype=torch.float64, device="cuda")
    y = torch.zeros((10, 33), dtype=torch.float64, device="cuda")
    out = torch.zeros((64,), dtype=torch.float64, device="cuda")
    lib = _lib.load()
    s = torch.cuda.current_stream().cuda_stream
    assert lib.flpr_xstats_noalloc(x.data_ptr(), 10, 2, y.data_ptr(), 33, out.data_ptr(), s) == -3
    assert lib.flpr_xstats_noalloc(x.data_ptr(), 10, 4, None, 0, out.data_ptr(), s) == -3
    assert lib.flpr_xstats_noalloc(None, 10, 2, None, 0, out.data_ptr(), s) == -1
    assert isinstance(lib.flpr_xstats_noalloc, ctypes._CFuncPtr)
