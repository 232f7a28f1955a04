"""The 512 batched uneven problems, problem by problem: problem i draws with
default_rng(1000 + i) its size N = round(10^U(3, 6)), then its order
integers(0, 3); i < 256 is 1-D on 1024 nodes, otherwise 2-D on 128 x 128; 20
bandwidths log-spaced in [0.02, 0.5]; every tenth problem has complex
y = (a + ib) f(x) with a, b ~ N(0, 1).  The response f (not given by the
configuration) is fixed here as sin(2 pi x1) (+ cos(2 pi x2) in 2-D) plus 0.1
noise.  A sample of the problems small enough for the float64 oracle, covering
1-D / 2-D, orders 0-2, complex y and the cubic GCV/DOF interpolation of 2-D
problems under 5000 points."""
import warnings

import numpy as np
import pytest

from oracle import restate as R

pytestmark = pytest.mark.gpu


def problem(i):
    rng = np.random.default_rng(1000 + i)
    n = int(round(10 ** rng.uniform(3, 6)))
    order = int(rng.integers(0, 3))
    d = 1 if i < 256 else 2
    x = rng.uniform(0, 1, (n, d))
    f = np.sin(2 * np.pi * x[:, 0]) + (np.cos(2 * np.pi * x[:, 1]) if d == 2 else 0.0)
    f = f + 0.1 * rng.standard_normal(n)
    if i % 10 == 0:
        a, b = rng.standard_normal(2)
        f = (a + 1j * b) * f
    hh = np.logspace(np.log10(0.02), np.log10(0.5), 20)
    h = np.column_stack([hh] * d)
    opt = {"order": order, "N": 1024 if d == 1 else [128, 128]}
    return x, f, h, opt


# (problem, d, n, order, complex): 5 (1, 1765, 2), 10 (1, 1927, 2, c), 19 (1, 4475, 0),
# 130 (1, 4784, 1, c), 140 (1, 2498, 0, c), 256 (2, 8995, 1), 261 (2, 8319, 2),
# 275 (2, 1787, 2, cubic), 284 (2, 4459, 0, cubic), 310 (2, 3306, 0, c, cubic),
# 320 (2, 6348, 2, c)
SAMPLE = [5, 10, 19, 130, 140, 256, 261, 275, 284, 310, 320]


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("i", SAMPLE)
def test_uneven_problem_matches_oracle(i):
    import fastlpr_amd as fl
    x, y, h, opt = problem(i)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        o = R.cv_fastlpr(x, y, h, dict(opt), mode="f64")
        r64 = fl.cv_fastlpr(x, y, h, dict(opt, precision="fp64"))
        r32 = fl.cv_fastlpr(x, y, h, dict(opt))
    cubic = x.shape[1] == 2 and x.shape[0] < 5000
    assert r64.gcv_yhat["id1se"] == o.gcv["id1se"]
    assert _rel(r64.yhat, o.yhat) < 1e-8
    assert _rel(r64.gcv_yhat["gcv_m"], o.gcv["gcv_m"]) < (1e-7 if cubic else 1e-8)
    assert r32.gcv_yhat["id1se"] == o.gcv["id1se"]
    assert _rel(r32.yhat, o.yhat) <= 1e-5
    assert _rel(r32.gcv_yhat["gcv_m"], o.gcv["gcv_m"]) < 1e-5
