"""Host-side API behaviour that needs no GPU: validation errors, bandwidth
lists and the host 1-SE rule (used by cv_fastkde), against the reference's
documented semantics (api.py:32-164, 835-1000; bandwidth.py:22-137).  The
device selection (Hutchinson, GCV, 1-SE) is checked in test_gpu_select."""
import numpy as np
import pytest

import fastlpr_amd
from fastlpr_amd import api
from oracle import restate as R


def test_get_hlist_matches_oracle_and_closed_form():
    h = fastlpr_amd.get_hlist(5, [0.1, 1.0])
    assert h.shape == (5, 1)
    np.testing.assert_allclose(h[:, 0], np.logspace(-1, 0, 5))
    h2 = fastlpr_amd.get_hlist([3, 2], [[0.1, 1.0], [0.2, 2.0]])
    # first dimension varies fastest (MATLAB ndgrid order)
    np.testing.assert_allclose(h2[:3, 1], 0.2)
    np.testing.assert_allclose(h2[3:, 1], 2.0)
    np.testing.assert_array_equal(h2, R.get_hlist([3, 2], [[0.1, 1.0], [0.2, 2.0]]))
    assert fastlpr_amd.get_hlist(1, [0.5, 0.5]).tolist() == [[0.5]]


def test_prepare_bandwidths_shapes_and_errors():
    rng = np.array([[0.1, 1.0], [0.1, 1.0]])
    assert api._prepare_bandwidths(0.3, 2, {}, rng).shape == (1, 2)
    assert api._prepare_bandwidths([0.3, 0.4], 2, {}, rng).shape == (1, 2)
    assert api._prepare_bandwidths(None, 2, {}, rng).shape == (400, 2)
    with pytest.raises(ValueError):
        api._prepare_bandwidths([0.3], 2, {}, rng)
    with pytest.raises(ValueError):
        api._prepare_bandwidths([[0.1, -0.2]], 2, {}, rng)
    with pytest.raises(ValueError):
        api._prepare_bandwidths([[0.1, np.nan]], 2, {}, rng)


def test_one_se_rule_matches_oracle_and_rules():
    rs = np.random.default_rng(0)
    for _ in range(50):
        k = 12
        s = rs.uniform(0.5, 1.5, k)
        e = rs.uniform(0, 0.2, k)
        s[rs.integers(0, k)] = np.nan
        bw = np.sort(rs.uniform(0.01, 1, (k, 1)), axis=0)
        for ps in (True, False):
            for dstd in (0.0, 1.0, 3.0):
                assert api._one_se_rule(s, e, bw, ps, dstd) == R.one_se_rule(s, e, bw, ps, dstd)
    # dstd = 0 picks the minimum; all-NaN raises
    assert api._one_se_rule(np.array([3.0, 1.0, 2.0]), np.zeros(3), np.ones((3, 1)), True, 0.0) == 1
    with pytest.raises(ValueError):
        api._one_se_rule(np.array([np.nan, np.inf]), np.zeros(2), np.ones((2, 1)), True)


@pytest.mark.parametrize("bad", ["nan_x", "inf_y", "order", "shape", "dims"])
def test_cv_fastlpr_validation_errors(bad):
    x = np.random.default_rng(0).uniform(size=(50, 1))
    y = np.sin(x[:, 0])
    opts = {"order": 1}
    if bad == "nan_x":
        x = x.copy()
        x[3, 0] = np.nan
    elif bad == "inf_y":
        y = y.copy()
        y[2] = np.inf
    elif bad == "order":
        opts = {"order": 3}
    elif bad == "shape":
        y = y[:40]
    elif bad == "dims":
        x = np.random.default_rng(0).uniform(size=(50, 11))
    with pytest.raises(ValueError):
        fastlpr_amd.cv_fastlpr(x, y, [[0.2]] if bad != "dims" else None, opts)


@pytest.mark.parametrize("acc,dims,exc", [(-1, 1, ValueError), (13, 1, NotImplementedError),
                                          (13, 3, NotImplementedError)])
def test_unsupported_accuracy_raises_before_device_work(acc, dims, exc):
    x = np.random.default_rng(0).uniform(size=(60, dims))
    with pytest.raises(exc):
        fastlpr_amd.cv_fastlpr(x, x[:, 0], [[0.3] * dims], {"order": 1, "accuracy": acc})
    with pytest.raises(exc):
        fastlpr_amd.cv_fastkde(x, [[0.3] * dims], {"accuracy": acc})


def test_fine_geometry_oversampling_constants():
    """kernel.py:97-156: R = 2 up to accuracy 11, 3 above; Mr = R L; the fine
    grid holds the data support plus the footprint on every axis."""
    import math
    from fastlpr_amd import engine as E
    st = E.Stats(np.array([0.0, 0.0]), np.array([1.0, 1.0]), np.array([0.5, 0.5]),
                 np.array([0.29, 0.29]))
    geo = E.build_geometry(st, (40, 33), 0, margin=E.fine_margin((40, 33)))
    assert all(geo.geom.E[k] >= (40, 33)[k] for k in range(2))
    for acc, R in ((7, 2), (11, 2), (12, 3)):
        f = E.fine_geometry(geo, acc)
        for k in range(2):
            L = geo.L[k]
            assert f.L[k] == R * L and f.msp == acc
            tau = math.pi * acc / (L * L * R * (R - 0.5))
            assert abs(f.inv4tau[k] * 4 * tau - 1) < 1e-15
            assert f.G[k] == 2 * (R * ((40, 33)[k] // 2 + 1) + acc) + 1


def test_cv_fastkde_validation_errors():
    import fastlpr_amd as fl
    with pytest.raises(ValueError, match="at least 2 observations"):
        fl.cv_fastkde(np.array([[0.3]]))
    with pytest.raises(ValueError, match="NaN"):
        fl.cv_fastkde(np.array([[0.3], [np.nan], [0.1]]))
    with pytest.raises(ValueError, match="Inf"):
        fl.cv_fastkde(np.array([[0.3], [np.inf], [0.1]]))


def test_cubic_tables_match_scipy_regular_grid_interpolator():
    """The not-a-knot collocation inverse and the de Boor basis the GPU kernel
    mirrors reproduce scipy's cubic interpolation: make_interp_spline(k=3) and
    RegularGridInterpolator(method='cubic') with a direct solve to rounding;
    RegularGridInterpolator's default iterative solve (what the reference
    runs) only to its 1e-5 residual."""
    from scipy.interpolate import RegularGridInterpolator, make_interp_spline
    from scipy.sparse.linalg import spsolve
    from fastlpr_amd import engine as E
    rs = np.random.default_rng(5)
    for N in (4, 5, 7, 16, 40):
        ax = np.linspace(-0.3, 1.7, N)
        t = E._nak_knots(ax)
        basis, first = E._cubic_basis(t, ax, N)
        A = np.zeros((N, N))
        for r in range(4):
            A[np.arange(N), first + r] += basis[:, r]
        y = rs.standard_normal(N)
        c = np.linalg.solve(A, y)
        pts = np.r_[ax, rs.uniform(ax[0], ax[-1], 50)]
        b, f0 = E._cubic_basis(t, pts, N)
        mine = np.sum(b * c[f0[:, None] + np.arange(4)[None, :]], axis=1)
        np.testing.assert_allclose(mine, make_interp_spline(ax, y, k=3)(pts), rtol=1e-12,
                                   atol=1e-12)
        exact = RegularGridInterpolator((ax,), y, method="cubic", solver=spsolve)(pts[:, None])
        np.testing.assert_allclose(mine, exact, rtol=1e-12, atol=1e-12)
        loose = RegularGridInterpolator((ax,), y, method="cubic")(pts[:, None])
        assert np.max(np.abs(mine - loose)) < 1e-4 * np.max(np.abs(y))


def test_interval_nu_constants():
    """nu(d, order) from the Gaussian kernel moments equals the closed forms
    and the reference's table (confidence.py:33-43)."""
    import math
    from fastlpr_amd import api as A
    ref = {(1, 0): 0.282094791774, (1, 1): 0.282094791774, (1, 2): 0.476034961118,
           (2, 0): 0.079577471546, (2, 1): 0.079577471546, (2, 2): 0.198943678865,
           (3, 0): 0.022448390266, (3, 1): 0.022448390266, (3, 2): 0.077166341538}
    for (d, o), v in ref.items():
        assert abs(A._nu(d, o) - v) < 1e-11, (d, o)
    assert abs(A._nu(1, 0) - 1 / (2 * math.sqrt(math.pi))) < 1e-15
    assert abs(A._nu(1, 2) - 27 / (32 * math.sqrt(math.pi))) < 1e-15
    assert abs(A._nu(3, 1) - 1 / (8 * math.pi ** 1.5)) < 1e-15
    with pytest.raises(ValueError):
        A._nu(4, 1)


def test_value_channel_limit_rejected_before_device_work():
    """The solve reduces at most 32 value channels (y columns + probes); a
    larger set raises NotImplementedError before the GPU is touched (this
    runs without one)."""
    rs = np.random.default_rng(3)
    x = rs.uniform(0, 1, (200, 1))
    y = rs.standard_normal(200)
    h = np.array([[0.1], [0.3]])
    for K in (32, 40):
        with pytest.raises(NotImplementedError, match="value channels"):
            fastlpr_amd.cv_fastlpr(x, y, h, {"order": 1, "num_dof_sample": K})
    # complex y counts twice: 2 + 31 > 32
    with pytest.raises(NotImplementedError, match="value channels"):
        fastlpr_amd.cv_fastlpr(x, (1 + 1j) * y, h, {"order": 1, "num_dof_sample": 31})
    # a random_matrix gives the probe count by its columns
    with pytest.raises(NotImplementedError, match="value channels"):
        fastlpr_amd.cv_fastlpr(x, y, h, {"order": 1, "random_matrix": np.ones((200, 33))})
    api._check_channels(1, 31)  # 32 channels: accepted


def test_interpolant_rejects_wrong_point_columns():
    """Query points with the wrong number of columns raise ValueError before
    any kernel launch (the interpolation kernels read d coordinates per point)."""
    from types import SimpleNamespace
    import torch
    from fastlpr_amd import engine
    geo = SimpleNamespace(N=(8, 8))
    with pytest.raises(ValueError, match=r"shape \(n, 2\)"):
        engine.check_points(geo, torch.zeros((5, 1), dtype=torch.float64))
    with pytest.raises(ValueError, match=r"shape \(n, 2\)"):
        engine.check_points(geo, torch.zeros((5,), dtype=torch.float64))
    engine.check_points(geo, torch.zeros((5, 2), dtype=torch.float64))
    interp = api.GridInterpolant(geo, [np.arange(8.0)] * 2, torch.zeros((1, 64)), (8, 8),
                                 False, 1, True, True)
    with pytest.raises(ValueError, match=r"shape \(n, 2\)"):
        interp(np.zeros((3, 1)))
    with pytest.raises(ValueError, match=r"shape \(n, 2\)"):
        interp(np.zeros((3, 3)))
