"""fastlpr_predict and fastlpr_interval on GPU fits against the oracle."""
import warnings

import numpy as np
import pytest

from oracle import restate as R

pytestmark = pytest.mark.gpu


def _data(d, n=4000, seed=3):
    rs = np.random.default_rng(seed)
    x = rs.uniform(0, 1, (n, d))
    sd = 0.05 + 0.3 * x[:, 0]
    y = np.sin(2 * np.pi * x[:, 0]) + (np.cos(2 * np.pi * x[:, -1]) if d > 1 else 0) \
        + sd * rs.standard_normal(n)
    return x, y


@pytest.mark.parametrize("d,order,N", [(1, 1, 128), (1, 2, 101), (2, 1, [30, 28])])
@pytest.mark.parametrize("kind", ["confidence", "prediction"])
def test_interval_matches_oracle(d, order, N, kind):
    import fastlpr_amd as fl
    x, y = _data(d)
    h = np.column_stack([np.logspace(np.log10(0.03), np.log10(0.6), 6)] * d)
    opt = {"order": order, "N": N, "precision": "fp64"}
    mu = fl.cv_fastlpr(x, y, h, dict(opt))
    r2 = (y - mu.yhat) ** 2
    sg = fl.cv_fastlpr(x, r2, h, dict(opt, y_type_out="variance"))
    ci = fl.fastlpr_interval(mu, sg, 0.1, kind)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        om = R.cv_fastlpr(x, y, h, {"order": order, "N": N}, mode="f64")
        osg = R.cv_fastlpr(x, (y - om.yhat) ** 2, h,
                           {"order": order, "N": N, "y_type_out": "variance"}, mode="f64")
    exp = R.interval(om.grid_values, osg.grid_values, om.s0, om.h, order, 0.1, kind)
    assert ci.shape == exp.shape == tuple(np.atleast_1d(N)) + (2,)
    assert np.max(np.abs(ci - exp)) / np.max(np.abs(exp)) < 1e-8
    # a constant sigma grid gives the same widths as the formula on our own grid
    sig = np.full(mu.s0.shape, 0.04)
    c2 = fl.fastlpr_interval(mu, sig, 0.1, kind)
    e2 = R.interval(mu.fpp_yhat.values, sig, mu.s0, mu.h, order, 0.1, kind)
    np.testing.assert_allclose(c2, e2, rtol=1e-12, atol=1e-12)
    # like the reference (confidence.py:183-187), sigma must match the grid
    with pytest.raises(ValueError):
        fl.fastlpr_interval(mu, 0.04, 0.1, kind)


def test_interval_on_device_tensors():
    import torch
    import fastlpr_amd as fl
    x, y = _data(1)
    h = np.logspace(np.log10(0.03), np.log10(0.6), 6)[:, None]
    dev = torch.device("cuda", 0)
    mu = fl.cv_fastlpr(torch.as_tensor(x, device=dev), torch.as_tensor(y, device=dev), h,
                       {"order": 1, "N": 128})
    sig = torch.full((128,), 0.02, dtype=torch.float64, device=dev)
    ci = fl.fastlpr_interval(mu, sig, 0.05)
    assert isinstance(ci, torch.Tensor) and ci.is_cuda and ci.shape == (128, 2)
    mh = fl.cv_fastlpr(x, y, h, {"order": 1, "N": 128})
    np.testing.assert_allclose(ci.cpu().numpy(), fl.fastlpr_interval(mh, np.full(128, 0.02), 0.05),
                               rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("d", [1, 2])
def test_predict_matches_linear_interpolation_of_the_grid(d):
    from scipy.interpolate import RegularGridInterpolator
    import fastlpr_amd as fl
    x, y = _data(d)
    h = np.column_stack([np.logspace(np.log10(0.03), np.log10(0.6), 6)] * d)
    r = fl.cv_fastlpr(x, y, h, {"order": 1, "N": 64 if d == 1 else [40, 36]})
    rs = np.random.default_rng(9)
    xn = rs.uniform(-0.1, 1.1, (500, d))
    got = fl.fastlpr_predict(r, xn if d > 1 else xn[:, 0])
    ref = RegularGridInterpolator(tuple(r.grid), r.fpp_yhat.values, method="linear",
                                  bounds_error=False, fill_value=None)(xn)
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(fl.fastlpr_predict(r, x), r.yhat, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("name", ["ivl1d_o1", "ivl1d_o2", "ivl2d_o1"])
def test_interval_and_predict_match_reference_golden(name):
    """The reference's fastlpr_interval / fastlpr_predict on its own fits
    (tests/golden/<name>.npz) against this engine's fits of the same data: the
    same selected bandwidths, intervals within the float32 spread of the two
    fits, predictions to the grid-value tolerance of the regression fixtures."""
    import os
    import recipes
    import fastlpr_amd as fl
    g = np.load(os.path.join(os.path.dirname(recipes.__file__), f"{name}.npz"),
                allow_pickle=False)
    x, y, h, opt = recipes.make(name)
    mu = fl.cv_fastlpr(x, y, h, dict(opt))
    sg = fl.cv_fastlpr(x, (y - mu.yhat) ** 2, h, dict(opt, y_type_out="variance"))
    np.testing.assert_array_equal(np.atleast_1d(mu.h), g["h_mu"])
    np.testing.assert_array_equal(np.atleast_1d(sg.h), g["h_sigma"])
    for kind, key in (("confidence", "ci"), ("prediction", "pi")):
        ci = fl.fastlpr_interval(mu, sg, 0.05, kind)
        assert ci.shape == g[key].shape
        assert np.max(np.abs(ci - g[key])) / np.max(np.abs(g[key])) < 2e-4
    pred = fl.fastlpr_predict(mu, g["x_new"])
    assert np.max(np.abs(pred - g["pred"])) / np.max(np.abs(g["pred"])) < 5e-5
