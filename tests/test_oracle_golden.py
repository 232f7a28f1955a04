"""The oracle restatement against the reference's own outputs (golden fixtures).

The fixtures were written by tests/golden/make_golden.py, which ran the
reference cv_fastlpr / cv_fastkde on the seeded recipes in tests/golden/recipes.py.
The reference computes S, T and the solves in float32 (accuracy 4), so the
tolerance is float32-level: 2e-5 relative on yhat / grid values / GCV, and
exact equality of the selected indices.
"""
import json
import os
import warnings

import numpy as np
import pytest

import recipes
from oracle import restate as R

GOLDEN = os.path.dirname(recipes.__file__)


def _load(name):
    return np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False)


def _relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("name", recipes.LPR_SMALL + recipes.LPR_CUBIC + recipes.LPR_ACC
                         + recipes.LPR_ACC_HI)
def test_oracle_lpr_matches_reference(name):
    g = _load(name)
    x, y, h, opt = recipes.make(name)
    assert json.loads(str(g["options"])) == opt
    np.testing.assert_array_equal(g["x"], x)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        f = R.cv_fastlpr(x, y, h, dict(opt), mode="ref")
    assert int(g["id1se"]) == f.gcv["id1se"]
    assert int(g["idmin"]) == f.gcv["idmin"]
    tol = 3e-5 if name.startswith("lpr3d") else 2e-5
    assert _relmax(f.yhat, g["yhat"]) < tol
    assert _relmax(f.grid_values, g["grid_values"]) < tol
    if opt.get("calc_dof", True):
        assert _relmax(f.gcv["gcv_m"], g["gcv_m"]) < 2e-6
        assert _relmax(f.gcv["dof"], g["dof"]) < 1e-4
    else:
        assert np.all(np.isnan(g["gcv_m"])) and np.all(np.isnan(f.gcv["gcv_m"]))
    assert _relmax(f.s0, g["s0"]) < 2e-6
    if np.isfinite(float(g["d"])):
        assert abs(f.d - float(g["d"])) / abs(float(g["d"])) < 2e-5


def test_oracle_1d_o1_n1e5_full_size_matches_reference():
    g = _load("large1d_o1_n1e5")
    x, y, h, opt = recipes.make("large1d_o1_n1e5")
    import hashlib
    assert hashlib.sha256(np.ascontiguousarray(x).tobytes()).hexdigest() == str(g["x_hash"])
    f = R.cv_fastlpr(x, y, h, dict(opt), mode="ref")
    assert int(g["id1se"]) == f.gcv["id1se"]
    idx = g["yhat_idx"]
    assert _relmax(f.yhat[idx], g["yhat_sub"]) < 2e-5
    assert _relmax(f.gcv["gcv_m"], g["gcv_m"]) < 2e-6


@pytest.mark.parametrize("name", recipes.KDE + recipes.KDE_ACC_HI)
def test_oracle_kde_matches_reference(name):
    g = _load(name)
    x, _, h, opt = recipes.make(name)
    f = R.cv_fastkde(x, h, dict(opt), mode="ref")
    assert int(g["id1se"]) == f.lcv["id1se"]
    assert int(g["idmax"]) == f.lcv["idmax"]
    assert _relmax(f.fhat, g["fhat"]) < 2e-5
    assert _relmax(f.lcv["lcv_m"], g["lcv_m"]) < 2e-6


@pytest.mark.parametrize("name", recipes.IVL)
def test_oracle_interval_and_predict_match_reference(name):
    """fastlpr_interval's formula (confidence.py:60-216) and fastlpr_predict's
    linear interpolation (predict.py:15), applied to the reference's own mean
    and variance grids, reproduce the reference's intervals and predictions."""
    g = _load(name)
    x, y, h, opt = recipes.make(name)
    np.testing.assert_array_equal(g["x"], x)
    np.testing.assert_array_equal(g["x_new"], recipes.predict_points(name))
    for kind, key in (("confidence", "ci"), ("prediction", "pi")):
        got = R.interval(g["mu_grid"], g["sigma_grid"], g["s0"], g["h_mu"], opt["order"], 0.05,
                         kind)
        assert got.shape == g[key].shape
        assert _relmax(got, g[key]) < 1e-12
    axes = R.raw_grid(x, list(np.atleast_1d(opt["N"])) * (x.shape[1] if np.isscalar(opt["N"])
                                                          else 1))
    pred = R.interp_linear(axes, g["mu_grid"], g["x_new"])
    assert _relmax(pred, g["pred"]) < 1e-12
