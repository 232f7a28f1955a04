"""flpr_select (Hutchinson DOF, GCV, 1-SE on the device) against the oracle's
restatement of regression.py:1276-1354, regression.py:378-421 and api.py:73-164.

The kernel receives the sums the fit leaves on the device (probe dots p'phat,
y dots y'yhat, yhat'yhat, y'y); the oracle receives the vectors themselves.
The DOF path sees identical inputs and must match bit for bit; the residual
sums differ only by the rounding of y'y - 2 y'yhat + yhat'yhat against
sum (y - yhat)^2.
"""
import ctypes

import numpy as np
import pytest

from oracle import restate as R

pytestmark = pytest.mark.gpu


def _case(seed, n, K, dh, ny_c, cplx, nan_b=None):
    rs = np.random.default_rng(seed)
    p = rs.standard_normal((n, K))
    lev = np.linspace(0.02, 0.6, dh)
    phat = lev[None, None, :] * p[:, :, None] + 0.05 * rs.standard_normal((n, K, dh))
    if nan_b is not None:
        phat[:, 0, nan_b] = np.nan  # one probe's trace becomes NaN -> n
    if cplx:
        y = rs.standard_normal((n, ny_c)) + 1j * rs.standard_normal((n, ny_c))
        yh = [y + s * (rs.standard_normal((n, ny_c)) + 1j * rs.standard_normal((n, ny_c)))
              for s in np.linspace(0.05, 0.9, dh)]
    else:
        y = rs.standard_normal((n, ny_c))
        yh = [y + s * rs.standard_normal((n, ny_c)) for s in np.linspace(0.05, 0.9, dh)]
    h = np.exp(rs.uniform(np.log(0.01), np.log(1.0), (dh, 2)))
    h[3] = h[5][::-1]  # a tie in sum(h)
    return p, phat, y, yh, h


def _real_cols(a):
    """(n, c) complex -> (n, 2c) interleaved (re, im); real unchanged."""
    if np.iscomplexobj(a):
        return np.ascontiguousarray(a).view(np.float64).reshape(a.shape[0], -1)
    return a


@pytest.mark.parametrize("K", [1, 3, 10])
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("agg", ["first", "average", "sum"])
@pytest.mark.parametrize("dstd", [0.0, 1.0])
def test_select_matches_oracle(K, cplx, agg, dstd):
    _check_select(K, cplx, agg, dstd, 9, None)


@pytest.mark.parametrize("dh", [9, 64, 70])
@pytest.mark.parametrize("nan_score", [None, 0, 4])
def test_select_pick_sizes_and_nan_scores(dh, nan_score):
    """The one-wave selection (dh <= 64) and the serial one (dh > 64), with a
    NaN score first or inside (np.argmin's first NaN, the 1-SE rule over the
    finite scores)."""
    _check_select(3, False, "average", 1.0, dh, nan_score)


def _check_select(K, cplx, agg, dstd, dh, nan_score):
    import torch
    from fastlpr_amd import _lib, engine
    _lib.load()
    n, ny_c = 400, 1 if cplx else 3
    p, phat, y, yh, h = _case(17 + K, n, K, dh, ny_c, cplx, nan_b=2)
    if nan_score is not None:
        yh[nan_score] = yh[nan_score] * np.nan
    dof, dse, pen_std, pen_mean, pph, pp = R.hutchinson(p, phat, n, cplx)
    scores, errs = [], []
    for b in range(dh):
        yy = y[:, 0] if agg == "first" and ny_c == 1 else y
        hh = yh[b][:, 0] if agg == "first" and ny_c == 1 else yh[b]
        s, e, _ = R.gcv_score(yy, hh, dof[b], pen_std[b],
                              None if pen_mean is None else pen_mean[b], agg)
        scores.append(s)
        errs.append(e)
    scores, errs = np.array(scores), np.array(errs)
    sel = R.one_se_rule(scores, errs, h, True, dstd)
    yr = _real_cols(y)
    ny = yr.shape[1]
    dots = np.zeros((dh, ny + K))
    quad = np.zeros((dh, ny))
    for b in range(dh):
        hr = _real_cols(yh[b])
        dots[b, :ny] = np.sum(yr * hr, axis=0)
        quad[b] = np.sum(hr * hr, axis=0)
        dots[b, ny:] = pph[:, b]
    dev = torch.device("cuda", 0)
    red = torch.tensor(np.concatenate([dots.ravel(), quad.ravel()]), device=dev)
    ysq = torch.tensor(np.sum(yr * yr, axis=0), device=dev)
    ppd = torch.tensor(pp, device=dev)
    hs = torch.tensor(h.sum(axis=1), device=dev)
    out = torch.zeros(4 * dh + 3, dtype=torch.float64, device=dev)
    seld = torch.full((1,), -7, dtype=torch.int32, device=dev)
    _lib.call("flpr_select", red.data_ptr(), dh, ny + K, ny, K, ysq.data_ptr(), ppd.data_ptr(),
              hs.data_ptr(), n, int(cplx), engine.AGGREGATIONS[agg], dstd, out.data_ptr(),
              seld.data_ptr(), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    o = out.cpu().numpy()
    np.testing.assert_array_equal(o[2 * dh:3 * dh], dof)
    np.testing.assert_array_equal(o[3 * dh:4 * dh], dse)
    np.testing.assert_allclose(o[:dh], scores, rtol=1e-10)
    np.testing.assert_allclose(o[dh:2 * dh], errs, rtol=1e-10, atol=0)
    assert int(o[4 * dh]) == sel == int(seld.item())
    assert int(o[4 * dh + 1]) == int(np.argmin(scores))
    assert int(o[4 * dh + 2]) == int(np.sum(np.isfinite(scores)))


def test_select_rejects_bad_arguments():
    from fastlpr_amd import _lib
    lib = _lib.load()
    # nval must equal ny + nprobe; no device memory is touched on rejection
    rc = lib.flpr_select(None, 4, 5, 1, 3, None, None, None, 100, 0, 0, 0.0, None, None, None)
    assert rc == -1


def test_conv_entry_points_reject_bad_arguments():
    """flpr_conv_strips and flpr_conv validate before touching device memory."""
    from fastlpr_amd import _lib
    lib = _lib.load()
    fake = 1 << 20   # never dereferenced: the checks fail first
    # the tap rows must come in whole groups of npow
    assert lib.flpr_conv_strips(fake, fake, 5, 3, 256, fake, None) == -1
    assert lib.flpr_conv_strips(fake, fake, 6, 3, 0, fake, None) == -1
    # nmax must be positive (the float32 form addresses its strip rows by it)
    assert lib.flpr_conv(0, 0, fake, fake, fake, 0, fake, fake, 1, 1, 267, 256, 1, 0, 1,
                         None, 0, None) == -1
    # max_powers: 1..3 for the float32 form, 1 for the float64 form
    for fp64, mp in ((0, 0), (0, 4), (1, 2), (1, 0)):
        assert lib.flpr_conv(0, fp64, fake, fake, fake, 256, fake, fake, 1, 1, 267, 256, 1, 0,
                             mp, None, 0, None) == -1
