"""The learned bandwidth bracket on the GPU.

* The MLP kernels (bf16 MFMA forward and backward, loss, Adam) against the
  CPU oracle (oracle/mlp.py), whose gradients are pinned by finite
  differences in test_oracle_mlp.
* The feature kernel against the oracle's loop on a real spread grid.
* The trained model: on synthetic data sets it has not seen, the bracket's
  GCV minimum is within 2 % of the full sweep's on most of them; a bracketed
  fit's scores equal the full sweep's scores at the same bandwidths (the
  regularisation is the whole list's); whenever the bracketed result does not
  fall back, it matches the full sweep's choice within the 2 % rule; and the
  bracketed fit is faster than the full sweep.
"""
import time
import warnings

import numpy as np
import pytest
import torch

from oracle import mlp as M

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_mlp_forward_backward_match_oracle():
    from fastlpr_amd import bracket as B
    rs = np.random.default_rng(0)
    p = (rs.standard_normal(M.N_PARAMS) * 0.2).astype(np.float32)
    X = rs.standard_normal((300, M.IN)).astype(np.float32)
    dO = rs.standard_normal((300, M.OUT)).astype(np.float32)
    pd = torch.as_tensor(p, device="cuda")
    Xd = torch.as_tensor(X, device="cuda")
    acts = B.forward(pd, Xd)
    g = B.backward(pd, Xd, acts, torch.as_tensor(dO, device="cuda")).cpu().numpy()
    ora = M.forward(p, X)
    for name, t in zip(("Z1", "H1", "Z2", "H2", "O"), acts):
        e = _rel(t.cpu().numpy(), ora[name])
        assert e < 2e-3, (name, e)
    # the oracle's backward from the kernels' own activations (so that bf16
    # rounding of the same float32 values is compared)
    acts_np = {k: t.cpu().numpy().astype(np.float64) for k, t in
               zip(("Z1", "H1", "Z2", "H2", "O"), acts)}
    go = M.backward(p, X, acts_np, dO.astype(np.float64))
    for name, sl in (("W1", slice(0, 8192)), ("b1", slice(8192, 8320)),
                     ("W2", slice(8320, 24704)), ("W3", slice(24832, 25088))):
        e = _rel(g[sl], go[sl])
        assert e < 5e-3, (name, e)
    assert _rel(g, go) < 5e-3


def test_mlp_loss_and_adam_kernels():
    from fastlpr_amd import _lib
    rs = np.random.default_rng(1)
    O = rs.standard_normal((50, 2)).astype(np.float32)
    T = rs.standard_normal((50, 2)).astype(np.float32)
    Od, Td = torch.as_tensor(O, device="cuda"), torch.as_tensor(T, device="cuda")
    dO = torch.empty_like(Od)
    loss = torch.zeros(1, device="cuda")
    _lib.call("flpr_mlp_loss", 50, Od.data_ptr(), Td.data_ptr(), 1.0, 0.5, dO.data_ptr(),
              loss.data_ptr(), torch.cuda.current_stream().cuda_stream)
    lo, go = M.loss(O.astype(np.float64), T.astype(np.float64), 1.0, 0.5)
    assert abs(float(loss.item()) - lo) < 1e-5 * abs(lo)
    np.testing.assert_allclose(dO.cpu().numpy(), go, rtol=1e-5, atol=1e-7)
    n = 1000
    p = rs.standard_normal(n).astype(np.float32)
    g = rs.standard_normal(n).astype(np.float32)
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    pd, gd, md, vd = (torch.as_tensor(a.copy(), device="cuda") for a in (p, g, m, v))
    for step in (1, 2, 3):
        _lib.call("flpr_mlp_adam", n, pd.data_ptr(), gd.data_ptr(), md.data_ptr(), vd.data_ptr(),
                  1e-2, 0.9, 0.999, 1e-8, step, 0.0, torch.cuda.current_stream().cuda_stream)
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        p = p - 1e-2 * (m / (1 - 0.9 ** step)) / (np.sqrt(v / (1 - 0.999 ** step)) + 1e-8)
    np.testing.assert_allclose(pd.cpu().numpy(), p, rtol=1e-5, atol=1e-6)


def test_features_match_oracle():
    from fastlpr_amd import bracket as B
    rs = np.random.default_rng(2)
    n = 30000
    x = rs.uniform(0, 1, (n, 2))
    y = np.sin(5 * x[:, 0]) * np.cos(3 * x[:, 1]) + 0.2 * rs.standard_normal(n)
    xd, yd = torch.as_tensor(x, device="cuda"), torch.as_tensor(y, device="cuda")
    sg = B.spread_grids([xd], [yd], (48, 40))
    feat = B.grid_features(sg).cpu().numpy()[0]
    G0, G1 = sg["G"]
    grid = sg["grid"].cpu().numpy().reshape(2, G0, G1)
    # the spread channels themselves: channel 0 sums to the total footprint
    # weight, channel 1 weights y by the same footprints
    assert grid[0].min() >= 0 and np.isclose(grid[1].sum() / grid[0].sum(), y.mean(), rtol=1e-2)
    assert float(sg["ysq"].item()) == pytest.approx(float(np.sum(y * y)), rel=1e-12)
    want = M.features(grid[0], grid[1], float(np.sum(y * y)), float(sg["floor"].item()), n,
                      float(sg["dx"].item()))
    np.testing.assert_allclose(feat, want, rtol=2e-5, atol=2e-5)


H30 = np.logspace(np.log10(0.005), np.log10(0.5), 30)
HL30 = np.column_stack([H30, H30])
OPT = {"order": 2, "N": [256, 256]}
UNSEEN = list(range(30000, 30024))      # training used data sets 0 .. 19999


def test_bracket_matches_full_sweep_on_unseen_data():
    """On 24 data sets the model never saw: (1) without fallback, the bracket's
    GCV scores equal the full sweep's at the same bandwidths (the bracket's
    lambda is the whole list's); (2) the final choice, with the fallback,
    is within 2 % of the full sweep's GCV minimum every time; (3) the
    bracket alone holds a bandwidth within 2 % of the minimum on most of
    them (the hit rate)."""
    import fastlpr_amd as fl
    from fastlpr_amd import bracket as B
    hits = fallbacks = 0
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        for i in UNSEEN:
            x, y = B.synthetic(i, torch.device("cuda"))
            full = fl.cv_fastlpr(x, y, HL30, OPT)
            br = fl.cv_fastlpr(x, y, HL30, dict(OPT, bandwidth_bracket="learned"))
            info = br.metadata["bracket"]
            g = np.asarray(full.gcv_yhat["gcv_m"])
            idx = info["indices"]
            gmin = np.min(g)
            hits += int(np.min(g[idx]) <= gmin * (1 + B.HIT_TOL))
            if info["fallback"]:
                fallbacks += 1
                assert br.gcv_yhat["id1se"] == full.gcv_yhat["id1se"]
                assert _rel(br.yhat.cpu().numpy(), full.yhat.cpu().numpy()) < 1e-12
            else:
                gb = np.asarray(br.gcv_yhat["gcv_m"])
                assert len(gb) == len(idx)
                assert _rel(gb, g[idx]) < 1e-10
                chosen = g[idx][br.gcv_yhat["id1se"]]
                assert chosen <= gmin * (1 + B.HIT_TOL), (i, chosen / gmin)
    print(f"bracket hit rate {hits}/{len(UNSEEN)}, fallbacks {fallbacks}/{len(UNSEEN)}")
    assert hits >= 0.75 * len(UNSEEN)
    assert fallbacks <= 0.5 * len(UNSEEN)


def test_bracket_is_faster_than_full_sweep():
    """Fits per second of the bracketed fit against the full sweep on the
    same data (n = 1e6, 256 x 256, order 2, 30 bandwidths)."""
    import fastlpr_amd as fl
    from fastlpr_amd import bracket as B
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        data = []
        for i in UNSEEN[:12]:
            x, y = B.synthetic(i, torch.device("cuda"), n_range=(1e6, 1e6 + 1))
            r = fl.cv_fastlpr(x, y, HL30, dict(OPT, bandwidth_bracket="learned"))
            if not r.metadata["bracket"]["fallback"]:
                data.append((x, y))
        assert data
        times = {}
        for mode in ("full", "learned"):
            fl.cv_fastlpr(data[0][0], data[0][1], HL30, dict(OPT, bandwidth_bracket=mode))
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(3):
                for x, y in data:
                    fl.cv_fastlpr(x, y, HL30, dict(OPT, bandwidth_bracket=mode))
            torch.cuda.synchronize()
            times[mode] = (time.perf_counter() - t) / (3 * len(data))
    print(f"full sweep {1 / times['full']:.0f} fits/s, learned bracket "
          f"{1 / times['learned']:.0f} fits/s")
    assert times["learned"] < times["full"]
