"""Edge cases of cv_fastlpr / cv_fastkde on the GPU against the float64 oracle:
the smallest grids, a handful of points, a dimension with no spread, one
bandwidth, bandwidths too small for the grid (dropped with a warning, or all
of them replaced by Silverman's rule), several real or complex response
columns, an xrange wider than the data, a constant response, and a long 1-D
grid."""
import warnings
import zlib

import numpy as np
import pytest

from oracle import restate as R

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / den) if den > 0 else float(np.max(np.abs(a)))


def _case(name):
    # a stable per-case seed (str hashes are salted per process)
    rs = np.random.default_rng(zlib.crc32(name.encode()))
    if name == "grid2_1d":
        x = rs.uniform(0, 1, (400, 1)); y = np.sin(3 * x[:, 0]) + 0.1 * rs.standard_normal(400)
        return x, y, np.array([[0.2], [0.5], [1.0]]), {"order": 1, "N": 2}
    if name == "grid2x3_2d":
        x = rs.uniform(0, 1, (600, 2)); y = x[:, 0] - x[:, 1] + 0.1 * rs.standard_normal(600)
        return x, y, np.array([[0.4, 0.4], [0.8, 0.8]]), {"order": 1, "N": [2, 3],
                                                            "dof_interp_method": "linear",
                                                            "gcv_interp_method": "linear"}
    if name == "few_points":
        x = np.array([[0.1], [0.45], [0.9]]); y = np.array([1.0, 2.0, 0.5])
        return x, y, np.array([[0.3], [0.6], [1.2]]), {"order": 1, "N": 16}
    if name == "flat_dim":
        x = np.column_stack([rs.uniform(0, 1, 800), np.full(800, 0.25)])
        y = np.cos(2 * x[:, 0]) + 0.1 * rs.standard_normal(800)
        return x, y, np.array([[0.2, 0.2], [0.5, 0.5]]), {"order": 1, "N": [20, 6]}
    if name == "one_bandwidth":
        x = rs.uniform(0, 1, (500, 1)); y = x[:, 0] ** 2 + 0.05 * rs.standard_normal(500)
        return x, y, np.array([[0.15]]), {"order": 2, "N": 64}
    if name == "some_too_small":
        x = rs.uniform(0, 1, (500, 1)); y = np.sin(4 * x[:, 0]) + 0.1 * rs.standard_normal(500)
        return x, y, np.array([[1e-4], [2e-4], [0.1], [0.3]]), {"order": 1, "N": 40}
    if name == "all_too_small":
        x = rs.uniform(0, 1, (500, 1)); y = np.sin(4 * x[:, 0]) + 0.1 * rs.standard_normal(500)
        return x, y, np.array([[1e-4], [2e-4]]), {"order": 1, "N": 40}
    if name == "multi_sum":
        x = rs.uniform(0, 1, (900, 1))
        y = np.column_stack([np.sin(3 * x[:, 0]), np.cos(2 * x[:, 0]), x[:, 0]]) \
            + 0.1 * rs.standard_normal((900, 3))
        return x, y, np.logspace(-1.5, 0, 6)[:, None], {"order": 1, "N": 50,
                                                        "gcv_selection": "sum"}
    if name == "complex_multi":
        x = rs.uniform(0, 1, (900, 1))
        f = np.sin(3 * x[:, 0]) + 0.1 * rs.standard_normal(900)
        y = np.column_stack([(1 + 2j) * f, (0.5 - 1j) * f ** 2])
        return x, y, np.logspace(-1.5, 0, 6)[:, None], {"order": 1, "N": 50,
                                                        "gcv_selection": "average"}
    if name == "xrange_1d":
        x = rs.uniform(0.2, 0.7, (700, 1)); y = np.sin(5 * x[:, 0]) + 0.1 * rs.standard_normal(700)
        return x, y, np.logspace(-1.5, -0.3, 5)[:, None], {"order": 1, "N": 64,
                                                           "xrange": [[0.0, 1.0]]}
    if name == "xrange_2d":
        x = rs.uniform(0.2, 0.7, (2000, 2)); y = x[:, 0] * x[:, 1] + 0.05 * rs.standard_normal(2000)
        return x, y, np.column_stack([np.logspace(-1.2, -0.3, 4)] * 2), \
            {"order": 1, "N": [30, 28], "xrange": [[0.0, 1.0], [0.1, 0.9]],
             "dof_interp_method": "linear", "gcv_interp_method": "linear"}
    if name == "constant_y":
        x = rs.uniform(0, 1, (600, 1)); y = np.full(600, 2.5)
        return x, y, np.logspace(-1.5, 0, 5)[:, None], {"order": 2, "N": 32}
    if name == "long_grid":
        x = rs.uniform(0, 1, (20000, 1)); y = np.sin(9 * x[:, 0]) + 0.1 * rs.standard_normal(20000)
        return x, y, np.logspace(-2.5, -0.5, 8)[:, None], {"order": 1, "N": 4096}
    raise KeyError(name)


NAMES = ["grid2_1d", "grid2x3_2d", "few_points", "one_bandwidth", "some_too_small",
         "all_too_small", "multi_sum", "complex_multi", "xrange_1d", "xrange_2d", "constant_y",
         "long_grid"]


@pytest.mark.parametrize("name", NAMES)
def test_edge_case_matches_oracle(name):
    import fastlpr_amd as fl
    x, y, h, opt = _case(name)
    with warnings.catch_warnings(record=True) as wo:
        warnings.simplefilter("always")
        o = R.cv_fastlpr(x, y, h, dict(opt), mode="f64")
    with warnings.catch_warnings(record=True) as wg:
        warnings.simplefilter("always")
        r = fl.cv_fastlpr(x, y, h, dict(opt, precision="fp64"))
    ow = sorted({str(w.message) for w in wo if "bandwidth" in str(w.message).lower()})
    gw = sorted({str(w.message) for w in wg if "bandwidth" in str(w.message).lower()})
    assert gw == ow
    np.testing.assert_allclose(np.atleast_2d(r.gcv_yhat["bandwidths"]),
                               np.atleast_2d(o.gcv["bandwidths"]), rtol=1e-12)
    assert r.gcv_yhat["id1se"] == o.gcv["id1se"]
    assert _rel(r.yhat, o.yhat) < 1e-8
    assert _rel(r.fpp_yhat.values, o.grid_values) < 1e-8
    g_eng, g_or = np.asarray(r.gcv_yhat["gcv_m"]), np.asarray(o.gcv["gcv_m"])
    fin = np.isfinite(g_or)
    assert np.array_equal(np.isfinite(g_eng), fin)
    if fin.any():
        assert _rel(g_eng[fin], g_or[fin]) < 1e-8


@pytest.mark.parametrize("name,opt", [("flat_dim", {}), ("grid2x3_2d", None)])
def test_edge_cases_the_reference_rejects(name, opt):
    """A dimension without spread (the raw grid axis is not ascending) and the
    default cubic interpolation on an axis of fewer than four nodes: the
    oracle, following the reference, raises ValueError, and so does the GPU
    path, before any moment is computed."""
    import fastlpr_amd as fl
    x, y, h, o = _case(name)
    o = dict(o)
    if opt is None:
        o.pop("dof_interp_method"), o.pop("gcv_interp_method")
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        with pytest.raises(ValueError):
            R.cv_fastlpr(x, y, h, dict(o), mode="f64")
        with pytest.raises(ValueError):
            fl.cv_fastlpr(x, y, h, dict(o, precision="fp64"))


def test_kde_few_points_and_flat_dimension():
    import fastlpr_amd as fl
    x = np.array([[0.2], [0.25], [0.7]])
    h = np.array([[0.3], [0.6], [0.9]])
    o = R.cv_fastkde(x, h, {}, mode="f64")
    r = fl.cv_fastkde(x, h, {"precision": "fp64"})
    assert r.lcv["id1se"] == o.lcv["id1se"]
    assert _rel(r.fhat, o.fhat) < 1e-8
    xf = np.column_stack([np.linspace(0, 1, 300), np.full(300, 0.4)])
    with pytest.raises(ValueError):
        R.cv_fastkde(xf, np.full((2, 2), 0.3), {}, mode="f64")
    with pytest.raises(ValueError):
        fl.cv_fastkde(xf, np.full((2, 2), 0.3), {"precision": "fp64"})
