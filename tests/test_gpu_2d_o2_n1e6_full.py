"""The headline workload at its full size against the float64 oracle.

2-D local quadratic, n = 1e6 points x ~ U([0,1]^2) (float32 inputs),
y = sin(2 pi x1) cos(2 pi x2) + (0.05 + 0.2 x1) N(0,1), grid 256 x 256,
30 log-spaced bandwidths in [0.005, 0.5] (isotropic), 10 Hutchinson probes:
the mean pass, then the variance pass on (y - yhat)^2.  Both precision forms
are checked against oracle.restate.cv_fastlpr(mode="f64") (the same discrete
estimator in float64, FFT route):

* yhat, normwise max |a - b| / max |b|: <= 1e-8 (fp64 form), <= 1e-5 (fp32,
  and the precision form "mixed": float32 spread with compensated summation,
  float32 solves with one float64 refinement step);
* the selected grid at 4096 nodes drawn with default_rng(1): same bounds;
* the GCV curve and the Hutchinson DOF over all 30 bandwidths: same bounds;
* the selected index (1-SE) equal.

The variance pass of both sides takes the same input, (y - yhat_oracle)^2, so
that the comparison isolates the variance fit.  cv_fastlpr_pair is checked
against the same pair as two cv_fastlpr calls.  Each oracle pass takes about
a minute and 25 GB of host memory; each is computed once per module.
"""
import numpy as np
import pytest
import torch

import fastlpr_amd
from oracle import restate as R

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

N_PTS = 1_000_000
GRID = [256, 256]
TOL = {"fp64": 1e-8, "fp32": 1e-5, "mixed": 1e-5}


def _inputs():
    rng = np.random.default_rng(0)
    x = rng.uniform(0, 1, (N_PTS, 2)).astype(np.float32)
    s = 0.05 + 0.2 * x[:, 0]
    y = (np.sin(2 * np.pi * x[:, 0]) * np.cos(2 * np.pi * x[:, 1])
         + s * rng.standard_normal(N_PTS)).astype(np.float32)
    h = R.get_hlist(30, [0.005, 0.5])[:, 0]
    return x.astype(np.float64), y.astype(np.float64), np.column_stack([h, h])


X, Y, HL = _inputs()
NODES = np.random.default_rng(1).choice(GRID[0] * GRID[1], 4096, replace=False)
OPT = {"order": 2, "N": GRID}


def _rel(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def oracle_mean():
    return R.cv_fastlpr(X, Y, HL, OPT, mode="f64")


@pytest.fixture(scope="module")
def var_input(oracle_mean):
    return (Y - oracle_mean.yhat) ** 2


@pytest.fixture(scope="module")
def oracle_var(var_input):
    return R.cv_fastlpr(X, var_input, HL, dict(OPT, y_type_out="variance"), mode="f64")


@pytest.fixture(scope="module")
def dev_x():
    return torch.as_tensor(X, device="cuda")


def _check(r, o, prec, what):
    tol = TOL[prec]
    yh = r.yhat.cpu().numpy() if isinstance(r.yhat, torch.Tensor) else r.yhat
    e_y = _rel(yh, o.yhat)
    e_g = _rel(r.fpp_yhat.values.reshape(-1)[NODES], np.asarray(o.grid_values).reshape(-1)[NODES])
    e_gcv = _rel(r.gcv_yhat["gcv_m"], o.gcv["gcv_m"])
    e_dof = _rel(r.gcv_yhat["dof"], o.gcv["dof"])
    print(f"{what} {prec}: yhat {e_y:.2e} grid {e_g:.2e} gcv {e_gcv:.2e} dof {e_dof:.2e} "
          f"id1se {r.gcv_yhat['id1se']} / {o.gcv['id1se']}")
    assert r.gcv_yhat["id1se"] == o.gcv["id1se"]
    assert e_y <= tol, e_y
    assert e_g <= tol, e_g
    assert e_gcv <= tol, e_gcv
    assert e_dof <= tol, e_dof


@pytest.mark.parametrize("prec", ["fp64", "fp32", "mixed"])
def test_mean_pass_full_size(oracle_mean, dev_x, prec):
    yd = torch.as_tensor(Y, device="cuda")
    r = fastlpr_amd.cv_fastlpr(dev_x, yd, HL, dict(OPT, precision=prec))
    _check(r, oracle_mean, prec, "mean pass")


@pytest.mark.parametrize("prec", ["fp64", "fp32", "mixed"])
def test_variance_pass_full_size(oracle_var, var_input, dev_x, prec):
    yd = torch.as_tensor(var_input, device="cuda")
    r = fastlpr_amd.cv_fastlpr(dev_x, yd, HL, dict(OPT, precision=prec, y_type_out="variance"))
    _check(r, oracle_var, prec, "variance pass")
    # the variance normalisation factor 1 / d
    assert abs(r.d - oracle_var.d) <= TOL[prec] * abs(oracle_var.d)


def test_pair_equals_two_calls_full_size(dev_x):
    yd = torch.as_tensor(Y, device="cuda")
    opt = dict(OPT, precision="fp32")
    mu, sigma = fastlpr_amd.cv_fastlpr_pair(dev_x, yd, HL, opt)
    c1 = fastlpr_amd.cv_fastlpr(dev_x, yd, HL, opt)
    c2 = fastlpr_amd.cv_fastlpr(dev_x, (yd - c1.yhat) ** 2, HL, dict(opt, y_type_out="variance"))
    assert sigma.metadata["dof_shared"]
    for a, b in ((mu, c1), (sigma, c2)):
        assert a.gcv_yhat["id1se"] == b.gcv_yhat["id1se"]
        assert _rel(a.yhat.cpu().numpy(), b.yhat.cpu().numpy()) <= 1e-12
        assert _rel(a.gcv_yhat["gcv_m"], b.gcv_yhat["gcv_m"]) <= 1e-12
