"""Batched problems (cv_fastlpr_batch / BatchPlan) against one cv_fastlpr
call per problem.

The workload is the 512 uneven problems (problem i from
default_rng(1000 + i): n = round(10^U(3, 6)), order integers(0, 3), 1-D on
1024 nodes for i < 256 else 2-D on 128 x 128, 20 bandwidths in [0.02, 0.5],
complex y for every tenth problem; the recipe of
test_gpu_uneven_problems).  Every problem's batched result must equal its
single call: the float64 form to 1e-12 (normwise yhat, the GCV curve, the
DOF) with the same selected index; the float32 form to 1e-9 (its float32
moments round a spread grid whose float64 sums may differ in the last bit
between runs).  A sample is also
checked against the float64 oracle.  The graph test captures the batched
fit and replays it several times, with eager work between the replays,
against the eager result.
"""
import warnings

import numpy as np
import pytest
import torch

from oracle import restate as R
from test_gpu_uneven_problems import problem

pytestmark = pytest.mark.gpu

P_ALL = 512


def _rel(a, b):
    a = a.cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    b = b.cpu().numpy() if isinstance(b, torch.Tensor) else np.asarray(b)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / den) if den > 0 else float(np.max(np.abs(a)))


@pytest.fixture(scope="module")
def uneven512():
    xs, ys, hs, opts = [], [], [], []
    for i in range(P_ALL):
        x, y, h, o = problem(i)
        xs.append(torch.as_tensor(x, device="cuda"))
        ys.append(torch.as_tensor(y, device="cuda"))
        hs.append(h)
        opts.append(o)
    return xs, ys, hs, opts


def _compare(rb, rs, tol):
    assert rb.gcv_yhat["id1se"] == rs.gcv_yhat["id1se"]
    e = max(_rel(rb.yhat, rs.yhat), _rel(rb.gcv_yhat["gcv_m"], rs.gcv_yhat["gcv_m"]),
            _rel(rb.gcv_yhat["dof"], rs.gcv_yhat["dof"]))
    assert e <= tol, e
    return e


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-9)])
def test_batch_equals_single_calls(uneven512, prec, tol):
    import fastlpr_amd as fl
    xs, ys, hs, opts = uneven512
    o2 = [dict(o, precision=prec) for o in opts]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        rb = fl.cv_fastlpr_batch(xs, ys, hs, o2)
        worst = 0.0
        for i in range(P_ALL):
            rs = fl.cv_fastlpr(xs[i], ys[i], hs[i], o2[i])
            worst = max(worst, _compare(rb[i], rs, tol))
    print(f"{prec}: worst batched-vs-single difference {worst:.2e}")


def test_batch_sample_matches_oracle(uneven512):
    import fastlpr_amd as fl
    xs, ys, hs, opts = uneven512
    sample = [5, 10, 19, 130, 140, 256, 261, 320]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        rb = fl.cv_fastlpr_batch([xs[i] for i in sample], [ys[i] for i in sample],
                                 [hs[i] for i in sample], [opts[i] for i in sample])
        for i, r in zip(sample, rb):
            x, y, h, o = problem(i)
            ora = R.cv_fastlpr(x, y, h, dict(o), mode="f64")
            assert r.gcv_yhat["id1se"] == ora.gcv["id1se"]
            assert _rel(r.yhat, ora.yhat) <= 1e-5
            assert _rel(r.gcv_yhat["gcv_m"], ora.gcv["gcv_m"]) <= 1e-5


def test_batch_mixed_kinds_and_variance():
    """Real and complex responses, several columns and variance fits in one
    batch call; each against its single call."""
    import fastlpr_amd as fl
    rs = np.random.default_rng(3)
    xs, ys, hs, opts = [], [], [], []
    for i in range(12):
        d = 1 + i % 2
        n = int(rs.integers(2000, 20000))
        x = rs.uniform(0, 1, (n, d))
        f = np.sin(2 * np.pi * x[:, 0]) + 0.2 * rs.standard_normal(n)
        if i % 4 == 1:
            f = (1 + 2j) * f
        elif i % 4 == 2:
            f = np.column_stack([f, np.cos(3 * x[:, 0])])
        o = {"order": i % 3, "N": 64 if d == 1 else [32, 40], "precision": "fp64"}
        if i % 4 == 3:
            f = f ** 2
            o["y_type_out"] = "variance"
        hh = np.logspace(-1.5, -0.4, 8)
        xs.append(x)
        ys.append(f)
        hs.append(np.column_stack([hh] * d))
        opts.append(o)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        rb = fl.cv_fastlpr_batch(xs, ys, hs, opts)
        for i in range(12):
            r1 = fl.cv_fastlpr(xs[i], ys[i], hs[i], opts[i])
            _compare(rb[i], r1, 1e-12)
            assert isinstance(rb[i].yhat, np.ndarray)
            if opts[i].get("y_type_out") == "variance":
                assert abs(rb[i].d - r1.d) <= 1e-12 * abs(r1.d)


def test_batch_graph_replay(uneven512):
    """The captured batch replays to the eager result several times, with the
    program's own eager work (another batch fit) between the replays, and a
    new response set fed through the same graph."""
    import fastlpr_amd as fl
    xs, ys, hs, opts = uneven512
    sel = list(range(0, P_ALL, 16))
    bx = [xs[i] for i in sel]
    by = [ys[i] for i in sel]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        plan = fl.BatchPlan(bx, [hs[i] for i in sel], [opts[i] for i in sel])
        eager = plan.fit(by)
        plan.capture(by)
        by2 = [2.0 * y + 0.5 for y in by]
        eager2 = fl.BatchPlan(bx, [hs[i] for i in sel], [opts[i] for i in sel]).fit(by2)
        for rep in range(3):
            got = plan.fit(by)
            for a, b in zip(got, eager):
                _compare(a, b, 1e-9)
            fl.cv_fastlpr(xs[1], ys[1], hs[1], opts[1])      # eager work between replays
            got2 = plan.fit(by2)
            for a, b in zip(got2, eager2):
                _compare(a, b, 1e-9)


def test_batch_autotuned_launch_shapes(uneven512):
    """The autotuner measures tile / points-per-item candidates per class and
    keeps the fastest; the tuned plan computes the same fits (the knobs only
    reorder float64 additions)."""
    import fastlpr_amd as fl
    from fastlpr_amd import tune
    xs, ys, hs, opts = uneven512
    sel = list(range(3, P_ALL, 24))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        plan = fl.BatchPlan([xs[i] for i in sel], [hs[i] for i in sel], [opts[i] for i in sel],
                            tune=True)
        got = plan.fit([ys[i] for i in sel])
        classes = [c for v in plan._kindsets.values() for c in v]
        assert classes and all(c.tabs.get("tune_ms") or tune.lookup(c) for c in classes)
        for c in classes:
            assert c.tabs["launch_shape"] in tune.CANDIDATES[c.d] + [tune.DEFAULT[c.d]]
        for i, r in zip(sel, got):
            _compare(r, fl.cv_fastlpr(xs[i], ys[i], hs[i], opts[i]), 1e-11)
