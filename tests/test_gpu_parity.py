"""GPU parity: the HIP path (through the C-ABI) against the oracle and the reference.

* precision="fp64": float64 end to end; must equal the float64 oracle
  (same discrete estimator, FFT convolution there, banded MFMA here) to 1e-8
  relative on yhat / grid values / GCV, with identical selected index.
* precision="fp32" (the default): float32 MFMA convolution, float64 spread and
  solve; the gate is max|yhat - yhat*| / max|yhat*| <= 1e-5 against the float64
  oracle.
* golden fixtures: the reference's own outputs (float32 S/T/solve inside the
  reference), compared at float32-level tolerance (5e-5) with identical 1-SE index.
"""
import json
import os
import warnings

import numpy as np
import pytest

import recipes
from oracle import restate as R

pytestmark = pytest.mark.gpu

GOLDEN = os.path.dirname(recipes.__file__)


def _rel(a, b):
    """max |a - b| / max |b| over the finite entries; the NaN entries (a fit
    whose normal equations are 0/0 far from the data) must coincide."""
    a, b = np.asarray(a), np.asarray(b)
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return float("inf")
    if np.all(nb):
        return 0.0
    a, b = a[~nb], b[~nb]
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _cases():
    rng = np.random.default_rng(7)
    out = {}
    n = 3000
    x1 = rng.uniform(0, 1, (n, 1))
    y1 = np.sin(2 * np.pi * x1[:, 0]) + 0.2 * rng.standard_normal(n)
    h1 = np.logspace(np.log10(0.03), 0, 10)[:, None]
    for o in (0, 1, 2):
        out[f"1d_o{o}_m128"] = (x1, y1, h1, {"order": o, "N": 128})
        out[f"1d_o{o}_m101"] = (x1, y1, h1, {"order": o, "N": 101})
    out["1d_cplx_o1"] = (x1, (0.5 + 2j) * y1, h1, {"order": 1, "N": 64})
    r2 = (0.05 + 0.3 * x1[:, 0]) ** 2 * rng.standard_normal(n) ** 2
    out["1d_var_o1"] = (x1, r2, h1, {"order": 1, "N": 64, "y_type_out": "variance"})
    # explicit accuracies: S and the probes stay at the sample-size accuracy
    # (4 here), y is spread at the option's (0 = nearest-cell binning)
    for a in (0, 2, 6, 7, 12):
        out[f"1d_o1_acc{a}"] = (x1, y1, h1, {"order": 1, "N": 128, "accuracy": a})
    out["1d_o2_acc9_nodof"] = (x1, y1, h1, {"order": 2, "N": 101, "accuracy": 9,
                                            "calc_dof": False})
    out["1d_cplx_o1_acc8"] = (x1, (0.5 + 2j) * y1, h1, {"order": 1, "N": 64, "accuracy": 8})
    out["1d_o0_acc1_nodof"] = (x1, y1, h1, {"order": 0, "N": 101, "accuracy": 1,
                                            "calc_dof": False})
    out["1d_multi_o1"] = (x1, np.column_stack([y1, np.cos(3 * x1[:, 0])]), h1,
                          {"order": 1, "N": 64, "gcv_selection": "average"})
    n2 = 8000
    x2 = rng.uniform(0, 1, (n2, 2))
    y2 = np.sin(2 * np.pi * x2[:, 0]) * np.cos(2 * np.pi * x2[:, 1]) + 0.1 * rng.standard_normal(n2)
    hh = np.logspace(np.log10(0.02), np.log10(0.4), 7)
    h2 = np.column_stack([hh, hh])
    for o in (0, 1, 2):
        out[f"2d_o{o}_m40"] = (x2, y2, h2, {"order": o, "N": [40, 40]})
    out["2d_o2_m33x40"] = (x2, y2, h2, {"order": 2, "N": [33, 40]})
    out["2d_o1_nodof"] = (x2, y2, h2, {"order": 1, "N": [40, 40], "calc_dof": False})
    out["2d_o2_acc0"] = (x2, y2, h2, {"order": 2, "N": [40, 40], "accuracy": 0})
    out["2d_o1_acc3_nodof"] = (x2, y2, h2, {"order": 1, "N": [40, 40], "accuracy": 3,
                                            "calc_dof": False})
    out["2d_o2_acc7"] = (x2, y2, h2, {"order": 2, "N": [40, 33], "accuracy": 7})
    out["2d_o0_acc10_nodof"] = (x2, y2, h2, {"order": 0, "N": [40, 40], "accuracy": 10,
                                             "calc_dof": False})
    # fewer than 5000 points in 2-D / 3-D: GCV and DOF interpolate cubically
    n4 = 1500
    x4 = rng.uniform(0, 1, (n4, 2))
    y4 = np.sin(2 * np.pi * x4[:, 0]) * np.cos(2 * np.pi * x4[:, 1]) + 0.1 * rng.standard_normal(n4)
    h4 = np.column_stack([np.logspace(np.log10(0.04), np.log10(0.5), 6)] * 2)
    out["2d_o1_cubic"] = (x4, y4, h4, {"order": 1, "N": [24, 20]})
    out["2d_o2_cubic_gcvlin"] = (x4, y4, h4, {"order": 2, "N": [24, 20],
                                                "gcv_interp_method": "linear"})
    x5 = rng.uniform(0, 1, (n4, 3))
    y5 = np.sum((x5 - 0.5) ** 2, axis=1) + 0.1 * rng.standard_normal(n4)
    out["3d_o1_cubic"] = (x5, y5, np.column_stack([np.logspace(-1, np.log10(0.6), 4)] * 3),
                          {"order": 1, "N": [10, 9, 8]})
    n3 = 6000
    x3 = rng.uniform(0, 1, (n3, 3))
    y3 = np.sum((x3 - 0.5) ** 2, axis=1) + 0.1 * rng.standard_normal(n3)
    h3 = np.column_stack([np.logspace(-1, np.log10(0.5), 4)] * 3)
    out["3d_o1_m14"] = (x3, y3, h3, {"order": 1, "N": [14, 14, 14]})
    # accuracy > 6 in 3-D: the oversampled spread (flpr_spread_wide), y alone
    # (the ones channel and the probes stay at the sample-size accuracy) and,
    # without DOF, every channel; R = 3 at accuracy 12
    out["3d_o1_acc7"] = (x3, y3, h3, {"order": 1, "N": [12, 12, 12], "accuracy": 7})
    out["3d_o1_acc9_nodof"] = (x3, y3, h3, {"order": 1, "N": [12, 12, 12], "accuracy": 9,
                                            "calc_dof": False})
    out["3d_o0_acc12_nodof"] = (x3, y3, h3, {"order": 0, "N": [12, 12, 12], "accuracy": 12,
                                             "calc_dof": False})
    return out


CASES = _cases()


@pytest.fixture(scope="module")
def fl():
    import fastlpr_amd
    from fastlpr_amd import _lib
    _lib.load()
    return fastlpr_amd


def _oracle(x, y, h, opt):
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return R.cv_fastlpr(x, y, h, dict(opt), mode="f64")


@pytest.mark.parametrize("name", sorted(CASES))
def test_fp64_path_equals_fp64_oracle(fl, name):
    x, y, h, opt = CASES[name]
    o = _oracle(x, y, h, opt)
    r = fl.cv_fastlpr(x, y, h, dict(opt, precision="fp64"))
    assert r.gcv_yhat["id1se"] == o.gcv["id1se"]
    assert _rel(r.yhat, o.yhat) < 1e-8
    assert _rel(r.fpp_yhat.values, o.grid_values) < 1e-8
    assert _rel(r.s0, o.s0) < 1e-8
    if opt.get("calc_dof", True):
        assert _rel(r.gcv_yhat["gcv_m"], o.gcv["gcv_m"]) < 1e-8
        # the cubic DOF sums p'phat over n terms of both signs: the two exact
        # spline solves (per-axis here, one sparse LU in the oracle) differ in
        # rounding, which that cancellation lifts to 1e-8 of the small trace
        dtol = 1e-7 if "cubic" in name else 1e-8
        assert _rel(r.gcv_yhat["dof"], o.gcv["dof"]) < dtol
        assert _rel(r.gcv_yhat["gcv_sd"], o.gcv["gcv_sd"]) < 1e-6
    if o.d is not None:
        assert abs(r.d - o.d) / abs(o.d) < 1e-8


@pytest.mark.parametrize("name", sorted(CASES))
def test_fp32_path_within_1e5_of_fp64_oracle(fl, name):
    x, y, h, opt = CASES[name]
    o = _oracle(x, y, h, opt)
    r = fl.cv_fastlpr(x, y, h, dict(opt))
    assert r.gcv_yhat["id1se"] == o.gcv["id1se"]
    assert _rel(r.yhat, o.yhat) <= 1e-5
    if opt.get("calc_dof", True):
        assert _rel(r.gcv_yhat["gcv_m"], o.gcv["gcv_m"]) < 1e-5


@pytest.mark.parametrize("name", recipes.LPR_SMALL + recipes.LPR_CUBIC + recipes.LPR_ACC
                         + recipes.LPR_ACC_HI)
def test_gpu_matches_reference_golden(fl, name):
    g = np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False)
    x, y, h, opt = recipes.make(name)
    assert json.loads(str(g["options"])) == opt
    r = fl.cv_fastlpr(x, y, h, dict(opt))
    assert r.gcv_yhat["id1se"] == int(g["id1se"])
    assert r.gcv_yhat["idmin"] == int(g["idmin"])
    tol = 5e-5
    assert _rel(r.yhat, g["yhat"]) < tol
    assert _rel(r.fpp_yhat.values, g["grid_values"]) < tol
    # the reference's cubic GCV / DOF interpolation solves for its spline
    # coefficients iteratively to a 1e-5 residual; this engine solves exactly
    gtol = 2e-4 if name in recipes.LPR_CUBIC else 1e-5
    if opt.get("calc_dof", True):
        assert _rel(r.gcv_yhat["gcv_m"], g["gcv_m"]) < gtol
    assert _rel(r.s0, g["s0"]) < 1e-5


def test_gpu_matches_reference_1d_o1_n1e5_full_size(fl):
    g = np.load(os.path.join(GOLDEN, "large1d_o1_n1e5.npz"), allow_pickle=False)
    x, y, h, opt = recipes.make("large1d_o1_n1e5")
    r = fl.cv_fastlpr(x, y, h, dict(opt))
    assert r.gcv_yhat["id1se"] == int(g["id1se"])
    idx = g["yhat_idx"]
    assert _rel(r.yhat[idx], g["yhat_sub"]) < 5e-5
    assert _rel(r.gcv_yhat["gcv_m"], g["gcv_m"]) < 1e-5


@pytest.mark.parametrize("d", [1, 2, 3])
def test_spline_eval_equals_exact_not_a_knot_spline(fl, d):
    """flpr_spline_eval on coefficients from engine.spline_coefficients equals
    scipy's cubic RegularGridInterpolator solved directly, at random points, at
    grid nodes and at the ends."""
    import torch
    from scipy.interpolate import RegularGridInterpolator
    from scipy.sparse.linalg import spsolve
    from fastlpr_amd import engine as E
    rs = np.random.default_rng(40 + d)
    N = {1: (17,), 2: (9, 12), 3: (6, 7, 5)}[d]
    n = 3000
    x = rs.uniform(0.1, 2.0, (n, d))
    x[0] = x.min(axis=0)
    x[1] = x.max(axis=0)
    st = E.Stats(x.min(0), x.max(0), x.mean(0), x.std(0, ddof=1), 0, 0, None)
    geo = E.build_geometry(st, N, 4)
    grids = rs.standard_normal((5,) + N)
    dev = torch.device("cuda", 0)
    gd = torch.tensor(grids.reshape(5, -1), device=dev)
    pts = np.r_[x, np.stack(np.meshgrid(*geo.axes_raw, indexing="ij"), -1).reshape(-1, d)[::3]]
    out = E.spline_eval(geo, torch.tensor(pts, device=dev), E.spline_coefficients(geo, gd))
    got = out.cpu().numpy()
    for k in range(5):
        ref = RegularGridInterpolator(tuple(geo.axes_raw), grids[k], method="cubic",
                                      solver=spsolve)(pts)
        np.testing.assert_allclose(got[:, k], ref, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("name", ["2d_o2_m33x40", "2d_o1_m40", "1d_o2_m128", "3d_o1_m14"])
def test_block_scheduled_convolution_equals_oracle(fl, name, monkeypatch):
    """Every launch through flpr_conv's block table (conv_schedule, which the
    engine otherwise keeps for launches of 4096 tiles or more): float64 form
    against the float64 oracle at the usual 1e-8."""
    from fastlpr_amd import engine as E
    monkeypatch.setattr(E, "SCHED_MIN_TILES", 0)
    E._SCHED.clear()
    x, y, h, opt = CASES[name]
    o = _oracle(x, y, h, opt)
    r = fl.cv_fastlpr(x, y, h, dict(opt, precision="fp64"))
    assert r.gcv_yhat["id1se"] == o.gcv["id1se"]
    assert _rel(r.yhat, o.yhat) < 1e-8
    assert _rel(r.gcv_yhat["gcv_m"], o.gcv["gcv_m"]) < 1e-8
    r32 = fl.cv_fastlpr(x, y, h, dict(opt))
    assert _rel(r32.yhat, o.yhat) <= 1e-5
    E._SCHED.clear()
