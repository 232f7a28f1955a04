"""Seeded input recipes shared by the golden-fixture writer and the tests.

Every case returns (x, y, hlist, options).  Inputs are float64 numpy arrays
built from ``numpy.random.default_rng(seed)``.
"""
from __future__ import annotations

import numpy as np


def _logspace_h(k, lo, hi):
    return np.logspace(np.log10(lo), np.log10(hi), k)


def _iso(h, d):
    return np.column_stack([h] * d)


def make(name: str):
    if name.startswith("lpr1d_o"):
        # lpr1d_o{order}_m{M}
        order = int(name[len("lpr1d_o")])
        M = int(name.split("_m")[1])
        rng = np.random.default_rng(11)
        n = 1500
        x = rng.uniform(0, 1, (n, 1))
        y = np.sin(2 * np.pi * x[:, 0]) + 0.2 * rng.standard_normal(n)
        return x, y, _logspace_h(12, 0.03, 1.0)[:, None], {"order": order, "N": M}
    if name == "lpr1d_cplx_o1":
        rng = np.random.default_rng(12)
        n = 1200
        x = rng.uniform(0, 1, (n, 1))
        f = np.sin(2 * np.pi * x[:, 0]) + 0.2 * rng.standard_normal(n)
        y = (0.7 - 1.3j) * f
        return x, y, _logspace_h(10, 0.03, 1.0)[:, None], {"order": 1, "N": 64}
    if name == "lpr1d_var_o1":
        rng = np.random.default_rng(13)
        n = 1500
        x = rng.uniform(0, 1, (n, 1))
        s = 0.05 + 0.3 * x[:, 0]
        r2 = (s * rng.standard_normal(n)) ** 2
        return x, r2, _logspace_h(10, 0.05, 1.0)[:, None], {"order": 1, "N": 64, "y_type_out": "variance"}
    if name.startswith("acc"):
        # acc{a}_{d}d[_nodof]: accuracy a (0 = nearest-cell binning, > 6 =
        # oversampled grid with deconvolution), order 1
        parts = name.split("_")
        a, dd = int(parts[0][3:]), int(parts[1][0])
        rng = np.random.default_rng(50 + a + 10 * dd)
        n = 3000 if dd == 1 else 6000
        x = rng.uniform(0, 1, (n, dd))
        y = np.sin(2 * np.pi * x[:, 0]) + 0.2 * rng.standard_normal(n)
        N = 100 if dd == 1 else [30, 33]
        opt = {"order": 1, "N": N, "accuracy": a}
        if "nodof" in parts:
            opt["calc_dof"] = False
        return x, y, _iso(_logspace_h(6, 0.04, 0.6), dd), opt
    if name == "lpr2d_o1_cubic":
        # fewer than 5000 points in 2-D: the reference interpolates GCV / DOF cubically
        rng = np.random.default_rng(22)
        n = 1500
        x = rng.uniform(0, 1, (n, 2))
        y = np.sin(2 * np.pi * x[:, 0]) * np.cos(2 * np.pi * x[:, 1]) + 0.1 * rng.standard_normal(n)
        return x, y, _iso(_logspace_h(8, 0.03, 0.5), 2), {"order": 1, "N": [24, 24]}
    if name == "lpr3d_o1_cubic":
        rng = np.random.default_rng(32)
        n = 1500
        x = rng.uniform(0, 1, (n, 3))
        y = np.sum((x - 0.5) ** 2, axis=1) + 0.1 * rng.standard_normal(n)
        return x, y, _iso(_logspace_h(5, 0.1, 0.6), 3), {"order": 1, "N": [10, 10, 10]}
    if name.startswith("lpr2d_o"):
        # lpr2d_o{order}_m{M0}x{M1}
        order = int(name[len("lpr2d_o")])
        dims = name.split("_m")[1]
        M = [int(v) for v in dims.split("x")]
        rng = np.random.default_rng(21)
        n = 6000
        x = rng.uniform(0, 1, (n, 2))
        y = np.sin(2 * np.pi * x[:, 0]) * np.cos(2 * np.pi * x[:, 1]) + 0.1 * rng.standard_normal(n)
        return x, y, _iso(_logspace_h(8, 0.02, 0.4), 2), {"order": order, "N": M}
    if name == "lpr3d_o1":
        rng = np.random.default_rng(31)
        n = 5000
        x = rng.uniform(0, 1, (n, 3))
        y = np.sum((x - 0.5) ** 2, axis=1) + 0.1 * rng.standard_normal(n)
        return x, y, _iso(_logspace_h(5, 0.1, 0.6), 3), {"order": 1, "N": [12, 12, 12]}
    if name == "large1d_o1_n1e5":
        # 1D local linear, N=1e5, x~U(0,1), y=sin(2 pi x)+0.1 N(0,1), M=1024, 20 h in [0.01, 1], seed 0
        rng = np.random.default_rng(0)
        n = 100000
        x = rng.uniform(0, 1, (n, 1))
        y = np.sin(2 * np.pi * x[:, 0]) + 0.1 * rng.standard_normal(n)
        return x, y, _logspace_h(20, 0.01, 1.0)[:, None], {"order": 1, "N": 1024}
    if name.startswith("ivl"):
        # ivl{d}d_o{order}: mean fit, variance fit of its squared residuals,
        # then intervals and predictions (x_new from predict_points)
        d, order = int(name[3]), int(name.split("_o")[1])
        rng = np.random.default_rng(70 + d + 10 * order)
        n = 2000 if d == 1 else 6000
        x = rng.uniform(0, 1, (n, d))
        sd = 0.05 + 0.3 * x[:, 0]
        y = np.sin(2 * np.pi * x[:, 0]) + (np.cos(2 * np.pi * x[:, -1]) if d > 1 else 0) \
            + sd * rng.standard_normal(n)
        N = 100 if d == 1 else [30, 28]
        return x, y, _iso(_logspace_h(8, 0.03, 0.6), d), {"order": order, "N": N}
    if name.startswith("kde"):
        # kde{d}d[_acc{a}]
        d = int(name[3])
        rng = np.random.default_rng(40 + d)
        n = {1: 800, 2: 1500, 3: 1500}[d]
        x = rng.standard_normal((n, d))
        opt = {}
        if "_acc" in name:
            opt["accuracy"] = int(name.split("_acc")[1])
        return x, None, _iso(_logspace_h(10, 0.1, 1.0), d), opt
    raise KeyError(name)


LPR_SMALL = [
    "lpr1d_o0_m100", "lpr1d_o1_m100", "lpr1d_o2_m100",
    "lpr1d_o0_m128", "lpr1d_o1_m128", "lpr1d_o2_m128",
    "lpr1d_cplx_o1", "lpr1d_var_o1",
    "lpr2d_o0_m32x32", "lpr2d_o1_m32x32", "lpr2d_o1_m31x33",
    "lpr3d_o1",
]
LPR_CUBIC = ["lpr2d_o1_cubic", "lpr3d_o1_cubic"]
LPR_ACC = ["acc0_1d", "acc2_1d", "acc0_2d", "acc3_2d"]
LPR_ACC_HI = ["acc7_1d", "acc12_1d", "acc7_2d", "acc9_2d", "acc7_2d_nodof"]
LPR_LARGE = ["large1d_o1_n1e5"]
IVL = ["ivl1d_o1", "ivl1d_o2", "ivl2d_o1"]


def predict_points(name: str):
    """New predictor locations inside the data range for the ivl recipes."""
    x = make(name)[0]
    rng = np.random.default_rng(99)
    lo, hi = x.min(0), x.max(0)
    return lo + (hi - lo) * rng.uniform(0, 1, (500, x.shape[1]))


KDE = ["kde1d", "kde2d", "kde3d"]
KDE_ACC_HI = ["kde1d_acc7", "kde2d_acc8"]
