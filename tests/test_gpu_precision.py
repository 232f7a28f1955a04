"""The precision form (precision="mixed": float32 spread with compensated
summation, float32 convolution, float32 per-node solves with one float64
refinement step and a float64 fallback above a condition estimate of 1e6)
against the float64 oracle: yhat and the GCV curve within 1e-5 (normwise),
the same selected bandwidth.  The full-size config-2 case is in
test_gpu_2d_o2_n1e6_full."""
import warnings

import numpy as np
import pytest

from oracle import restate as R

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _cases():
    rng = np.random.default_rng(11)
    n = 20000
    x1 = rng.uniform(0, 1, (n, 1))
    y1 = np.sin(2 * np.pi * x1[:, 0]) + 0.2 * rng.standard_normal(n)
    h1 = np.logspace(np.log10(0.01), 0, 12)[:, None]
    x2 = rng.uniform(0, 1, (n, 2))
    y2 = np.sin(2 * np.pi * x2[:, 0]) * np.cos(2 * np.pi * x2[:, 1]) + 0.1 * rng.standard_normal(n)
    hh = np.logspace(np.log10(0.01), np.log10(0.4), 10)
    h2 = np.column_stack([hh, hh])
    out = {}
    for o in (0, 1, 2):
        out[f"1d_o{o}"] = (x1, y1, h1, {"order": o, "N": 256})
        out[f"2d_o{o}"] = (x2, y2, h2, {"order": o, "N": [64, 64]})
    out["2d_o2_var"] = (x2, (y2 - np.sin(2 * np.pi * x2[:, 0]) * np.cos(2 * np.pi * x2[:, 1])) ** 2,
                        h2, {"order": 2, "N": [64, 64], "y_type_out": "variance"})
    return out


CASES = _cases()


@pytest.mark.parametrize("name", sorted(CASES))
def test_mixed_precision_matches_oracle(name):
    import fastlpr_amd as fl
    x, y, h, opt = CASES[name]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        o = R.cv_fastlpr(x, y, h, dict(opt), mode="f64")
        r = fl.cv_fastlpr(x, y, h, dict(opt, precision="mixed"))
    e_y, e_g = _rel(r.yhat, o.yhat), _rel(r.gcv_yhat["gcv_m"], o.gcv["gcv_m"])
    print(f"{name}: mixed yhat {e_y:.2e} gcv {e_g:.2e}")
    assert r.gcv_yhat["id1se"] == o.gcv["id1se"]
    assert e_y <= 1e-5 and e_g <= 1e-5
