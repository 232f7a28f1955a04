"""cv_fastlpr_pair: the heteroscedastic mean + variance fit with the second
pass reusing the first pass's probe sums equals two cv_fastlpr calls."""
import warnings

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _data(d, n, seed=5):
    rs = np.random.default_rng(seed)
    x = rs.uniform(0, 1, (n, d))
    sd = 0.05 + 0.2 * x[:, 0]
    y = np.sin(2 * np.pi * x[:, 0]) * (np.cos(2 * np.pi * x[:, 1]) if d > 1 else 1) \
        + sd * rs.standard_normal(n)
    return x, y


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("d,order,N,prec,tol", [
    (1, 1, 128, "fp64", 1e-10), (2, 2, [48, 40], "fp64", 1e-10),
    (2, 2, [48, 40], "fp32", 1e-5), (2, 1, [33, 35], "fp32", 1e-5)])
def test_pair_equals_two_calls(d, order, N, prec, tol):
    import fastlpr_amd as fl
    x, y = _data(d, 6000)
    hh = np.logspace(np.log10(0.01), np.log10(0.5), 9)
    h = np.column_stack([hh] * d)
    opt = {"order": order, "N": N, "precision": prec}
    mu, sg = fl.cv_fastlpr_pair(x, y, h, opt, {"dstd": 1.0})
    assert sg.metadata["dof_shared"] and not mu.metadata["dof_shared"]
    mu2 = fl.cv_fastlpr(x, y, h, dict(opt))
    sg2 = fl.cv_fastlpr(x, (y - mu2.yhat) ** 2, h, dict(opt, dstd=1.0, y_type_out="variance"))
    for a, b in ((mu, mu2), (sg, sg2)):
        assert a.gcv_yhat["id1se"] == b.gcv_yhat["id1se"]
        assert isinstance(a.yhat, np.ndarray)
        assert _rel(a.yhat, b.yhat) < tol
        assert _rel(a.gcv_yhat["gcv_m"], b.gcv_yhat["gcv_m"]) < tol
        assert _rel(a.gcv_yhat["dof"], b.gcv_yhat["dof"]) < tol
        assert _rel(a.fpp_yhat.values, b.fpp_yhat.values) < tol
    assert abs(sg.d - sg2.d) / abs(sg2.d) < tol


def test_pair_recomputes_when_variance_options_change_the_dof():
    import fastlpr_amd as fl
    x, y = _data(1, 3000)
    h = np.logspace(np.log10(0.02), np.log10(0.5), 6)[:, None]
    mu, sg = fl.cv_fastlpr_pair(x, y, h, {"order": 1, "N": 64}, {"order": 2})
    assert not sg.metadata["dof_shared"]
    sg2 = fl.cv_fastlpr(x, (y - mu.yhat) ** 2, h,
                        {"order": 2, "N": 64, "y_type_out": "variance"})
    assert sg.gcv_yhat["id1se"] == sg2.gcv_yhat["id1se"]
    assert _rel(sg.yhat, sg2.yhat) < 1e-5


def test_pair_small_2d_uses_cubic_and_matches_two_calls():
    """n < 5000 in 2-D: the default GCV/DOF interpolation is cubic, which the
    pair does not share; the result still equals two calls."""
    import fastlpr_amd as fl
    x, y = _data(2, 1500)
    h = np.column_stack([np.logspace(np.log10(0.04), np.log10(0.5), 5)] * 2)
    opt = {"order": 1, "N": [20, 22], "precision": "fp64"}
    mu, sg = fl.cv_fastlpr_pair(x, y, h, opt)
    assert not sg.metadata["dof_shared"]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        mu2 = fl.cv_fastlpr(x, y, h, dict(opt))
        sg2 = fl.cv_fastlpr(x, (y - mu2.yhat) ** 2, h, dict(opt, y_type_out="variance"))
    assert _rel(sg.yhat, sg2.yhat) < 1e-10


def test_reused_x_side_still_rejects_non_finite_y():
    """With the x side reused, y's non-finite check runs on the device and is
    reported when the results are read back: the same ValueError as cv_fastlpr."""
    import torch
    from fastlpr_amd import api
    x, y = _data(1, 3000)
    h = np.logspace(np.log10(0.02), np.log10(0.5), 6)[:, None]
    dev = torch.device("cuda", 0)
    xd = torch.as_tensor(x, device=dev)
    yd = torch.as_tensor(y, device=dev)
    opt = {"order": 1, "N": 64}
    finish, share = api._fit(xd, yd, h, opt, as_numpy=False, keep_x=True)
    bad = yd.clone()
    bad[17] = float("nan")
    finish2, _ = api._fit(xd, bad, h, opt, prior=share, as_numpy=False)
    finish()
    with pytest.raises(ValueError, match="NaN"):
        finish2()
    share.clear()
