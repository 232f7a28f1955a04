"""GPU parity of cv_fastkde (density grids, LCV, 1-SE) against the reference's
golden fixtures and the float64 oracle."""
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
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def fl():
    import fastlpr_amd
    from fastlpr_amd import _lib
    _lib.load()
    return fastlpr_amd


def _oracle(x, h, opt):
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return R.cv_fastkde(x, h, dict(opt), mode="f64")


@pytest.mark.parametrize("name", recipes.KDE + recipes.KDE_ACC_HI)
def test_kde_matches_reference_golden(fl, name):
    g = np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False)
    x, _, h, opt = recipes.make(name)
    assert json.loads(str(g["options"])) == opt
    r = fl.cv_fastkde(x, h, dict(opt))
    assert r.lcv["id1se"] == int(g["id1se"])
    assert r.lcv["idmax"] == int(g["idmax"])
    assert _rel(r.fhat, g["fhat"]) < 5e-5
    assert _rel(r.lcv["lcv_m"], g["lcv_m"]) < 1e-5
    assert np.allclose(np.asarray(r.grid), g["grid"], rtol=1e-12, atol=0)
    np.testing.assert_array_equal(r.h, g["h1se"])


@pytest.mark.parametrize("name", recipes.KDE + recipes.KDE_ACC_HI)
def test_kde_fp64_equals_fp64_oracle(fl, name):
    x, _, h, opt = recipes.make(name)
    o = _oracle(x, h, opt)
    r = fl.cv_fastkde(x, h, dict(opt, precision="fp64"))
    assert r.lcv["id1se"] == o.lcv["id1se"]
    assert r.lcv["idmax"] == o.lcv["idmax"]
    assert _rel(r.fhat, o.fhat) < 1e-8
    assert _rel(r.metadata["fhat_all"], o.fhat_all) < 1e-8
    assert _rel(r.lcv["lcv_m"], o.lcv["lcv_m"]) < 1e-9
    assert _rel(r.lcv["lcv_sd"], o.lcv["lcv_sd"]) < 1e-7


@pytest.mark.parametrize("name", recipes.KDE + recipes.KDE_ACC_HI)
def test_kde_fp32_within_1e5_of_fp64_oracle(fl, name):
    x, _, h, opt = recipes.make(name)
    o = _oracle(x, h, opt)
    r = fl.cv_fastkde(x, h, dict(opt))
    assert r.lcv["id1se"] == o.lcv["id1se"]
    assert _rel(r.fhat, o.fhat) < 1e-5
    assert _rel(r.lcv["lcv_m"], o.lcv["lcv_m"]) < 1e-5


@pytest.mark.parametrize("acc", [8, 11])
def test_kde_3d_oversampled_accuracy_equals_oracle(fl, acc):
    """3-D KDE at accuracy > 6 (the oversampled spread of flpr_spread_wide and
    the deconvolving resample) against the float64 oracle: the float64 form to
    1e-8, the float32 form to 1e-5, with the same LCV selection."""
    rng = np.random.default_rng(20 + acc)
    x = np.clip(0.5 + 0.15 * rng.standard_normal((2500, 3)), 0, 1)
    # no bandwidth too small for the grid (the reference's cv_fastkde indexes
    # its densities with the unfiltered list, api.py:997, 1074-1097)
    h = np.column_stack([np.logspace(np.log10(0.1), np.log10(0.4), 5)] * 3)
    opt = {"N": [12, 10, 12], "accuracy": acc}
    o = _oracle(x, h, opt)
    r64 = fl.cv_fastkde(x, h, dict(opt, precision="fp64"))
    assert r64.lcv["id1se"] == o.lcv["id1se"]
    assert _rel(r64.fhat, o.fhat) < 1e-8
    assert _rel(r64.lcv["lcv_m"], o.lcv["lcv_m"]) < 1e-8
    r32 = fl.cv_fastkde(x, h, dict(opt))
    assert r32.lcv["id1se"] == o.lcv["id1se"]
    assert _rel(r32.fhat, o.fhat) < 1e-5
    assert _rel(r32.lcv["lcv_m"], o.lcv["lcv_m"]) < 1e-5


def test_kde_large_2d_properties(fl):
    """1e6 points: the density integrates to one on the grid, is non-negative,
    the interpolant reproduces the grid at the nodes, and device inputs give
    device outputs that equal the host-input results."""
    import torch
    rs = np.random.default_rng(3)
    n = 1_000_000
    x = np.column_stack([rs.normal(0.5, 0.1, n), rs.gamma(3.0, 0.1, n)])
    hl = fl.get_hlist([6, 6], [[0.02, 0.3], [0.02, 0.3]])
    opt = {"N": [128, 96]}
    r = fl.cv_fastkde(x, hl, dict(opt))
    dx = np.prod([a[1] - a[0] for a in r.grid])
    assert r.fhat.shape == (128, 96)
    assert np.all(r.fhat >= 0)
    assert abs(np.sum(r.fhat) * dx - 1.0) < 1e-9
    fa = r.metadata["fhat_all"]
    assert fa.shape == (128, 96, 36)
    assert np.all(np.abs(np.sum(fa, axis=(0, 1)) * dx - 1.0) < 1e-9)
    g0, g1 = np.meshgrid(r.grid[0], r.grid[1], indexing="ij")
    pts = np.column_stack([g0.ravel(), g1.ravel()])[::97]
    np.testing.assert_allclose(r.fpp(pts), r.fhat.ravel()[::97], rtol=1e-12, atol=1e-12)
    assert r.lcv["matrix"].shape == (6, 6)
    rd = fl.cv_fastkde(torch.as_tensor(x, device="cuda"), hl, dict(opt))
    assert isinstance(rd.fhat, torch.Tensor) and rd.fhat.is_cuda
    assert rd.lcv["id1se"] == r.lcv["id1se"]
    np.testing.assert_array_equal(rd.fhat.cpu().numpy(), r.fhat)
