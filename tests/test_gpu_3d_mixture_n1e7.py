"""The 3-D mixture workload at full size: 3-D local linear regression and KDE with LCV on 1e7
points from a 4-component Gaussian mixture in [0,1]^3, 128^3 grid, 16
bandwidths in [0.01, 0.3].  The float64 oracle cannot run at this size, so the
float32 form is checked against the float64 form (which equals the float64
oracle to 1e-8 at the sizes the oracle runs, tests/test_gpu_parity.py) at 4096
grid nodes drawn with default_rng(1), plus the invariants of the fits."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mixture(n, seed=0):
    rng = np.random.default_rng(seed)
    means = rng.uniform(0.2, 0.8, (4, 3))
    sd = rng.uniform(0.05, 0.15, 4)
    comp = rng.integers(0, 4, n)
    x = means[comp] + sd[comp, None] * rng.standard_normal((n, 3))
    bad = np.any((x < 0) | (x > 1), axis=1)
    while bad.any():   # points outside the unit cube are drawn again
        k = int(bad.sum())
        c = rng.integers(0, 4, k)
        x[bad] = means[c] + sd[c, None] * rng.standard_normal((k, 3))
        bad = np.any((x < 0) | (x > 1), axis=1)
    y = np.sum((x - 0.5) ** 2, axis=1) + 0.1 * rng.standard_normal(n)
    return x, y


def _normwise(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_3d_mixture_n1e7_full_size():
    import torch
    import fastlpr_amd as fl
    n = 10_000_000
    x, y = _mixture(n)
    hh = fl.get_hlist(16, [0.01, 0.3])[:, 0]
    hl = np.column_stack([hh, hh, hh])
    dev = torch.device("cuda", 0)
    xd, yd = torch.as_tensor(x, device=dev), torch.as_tensor(y, device=dev)
    del x
    opt = {"order": 1, "N": [128, 128, 128]}
    r32 = fl.cv_fastlpr(xd, yd, hl, dict(opt))
    r64 = fl.cv_fastlpr(xd, yd, hl, dict(opt, precision="fp64"))
    nodes = np.random.default_rng(1).choice(128 ** 3, 4096, replace=False)
    g32 = r32.fpp_yhat.device_values().reshape(-1)[torch.as_tensor(nodes, device=dev)]
    g64 = r64.fpp_yhat.device_values().reshape(-1)[torch.as_tensor(nodes, device=dev)]
    assert r32.gcv_yhat["id1se"] == r64.gcv_yhat["id1se"]
    assert _normwise(g32.cpu().numpy(), g64.cpu().numpy()) <= 1e-5
    assert _normwise(r32.yhat.cpu().numpy(), r64.yhat.cpu().numpy()) <= 1e-5
    assert _normwise(r32.gcv_yhat["gcv_m"], r64.gcv_yhat["gcv_m"]) <= 1e-5
    # a linear fit reproduces the regression function's level: the fitted
    # values at the data track y - noise (mean squared residual ~ 0.01)
    resid = (yd - r32.yhat).double()
    assert abs(float((resid * resid).mean()) - 0.01) < 2e-3
    k32 = fl.cv_fastkde(xd, hl, {"N": [128, 128, 128]})
    k64 = fl.cv_fastkde(xd, hl, {"N": [128, 128, 128], "precision": "fp64"})
    assert k32.lcv["id1se"] == k64.lcv["id1se"]
    f32, f64 = k32.fhat.cpu().numpy(), k64.fhat.cpu().numpy()   # (128, 128, 128) grids
    assert _normwise(f32.reshape(-1)[nodes], f64.reshape(-1)[nodes]) <= 1e-5
    assert _normwise(k32.lcv["lcv_m"], k64.lcv["lcv_m"]) <= 1e-5
    # the density integrates to one on the raw grid and is non-negative
    dx = np.prod([float(g[1] - g[0]) for g in k32.grid])
    assert np.all(f32 >= 0)
    assert abs(float(f32.sum()) * dx - 1.0) < 1e-9
