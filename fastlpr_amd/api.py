"""Drop-in ``cv_fastlpr`` / ``cv_fastkde`` on the MI355X engine.

Signatures, option keys, defaults, validation errors and warnings follow the
reference API (api.py:32-164, 167-832, 835-1184).  All numeric work on the
data runs in the HIP kernels of libfastlpr_amd.so (see ``engine``); the host
does validation, option handling, probe generation (NumPy MT19937, so the
Hutchinson probes are the reference's), and the O(dh) GCV / LCV selection.

Inputs may be NumPy arrays (results come back as NumPy) or torch tensors on a
GPU (results stay on the device as torch tensors, nothing crosses PCIe).

Extra options (no counterpart in the reference):
  ``precision``: "fp32" (default; float32 MFMA convolution with float64
  spread and float64 solves), "fp64" (float64 end to end) or "mixed" (the
  precision form: float32 spread with compensated summation, float32
  convolution, float32 solves with one float64 refinement step and a
  float64 fallback above a condition estimate of 1e6).
  ``bandwidth_bracket``: "full" (default, every bandwidth) or "learned"
  (2-D real single-column mean fits: a trained MLP proposes a bracket of
  five consecutive bandwidths around the GCV minimum, only those are solved,
  with the regularisation of the whole list; when the chosen one sits on an
  inner edge of the bracket the full sweep runs instead; see ``bracket``).
"""
from __future__ import annotations

import math
import warnings
from typing import Dict, Optional, Sequence, Tuple

import numpy as np
import torch

from . import engine
from .bandwidth import get_hlist
from .structures import KDEOutput, RegressionOutput

__all__ = ["cv_fastlpr", "cv_fastkde", "get_hlist", "GridInterpolant"]


# ---------------------------------------------------------------------------
# option and input helpers
# ---------------------------------------------------------------------------

def set_defaults(options: Optional[dict], **defaults) -> dict:
    """Shallow copy, then setdefault (utils/__init__.py:18-60)."""
    out = dict(options) if options is not None else {}
    for k, v in defaults.items():
        out.setdefault(k, v)
    return out


def _device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("fastlpr_amd needs a ROCm GPU (torch.cuda.is_available() is False); "
                           "there is no CPU fallback")
    return torch.device("cuda", torch.cuda.current_device())


def _is_dev(a) -> bool:
    return isinstance(a, torch.Tensor) and a.is_cuda


def _as_column_matrix(x):
    """1-D -> column, complex columns split into [real, imag] (utils/helpers.py:31-71)."""
    if isinstance(x, torch.Tensor):
        t = x
        if t.ndim == 1:
            t = t[:, None]
        if t.ndim != 2:
            raise ValueError("Input must be a 1D or 2D array.")
        if t.is_complex():
            cplx = torch.any(t.imag != 0, dim=0)
            t = torch.cat([t.real, t.imag[:, cplx]], dim=1) if bool(cplx.any()) else t.real
        return t.to(torch.float64)
    arr = np.asarray(x)
    if arr.ndim == 1:
        arr = arr[:, None]
    if arr.ndim != 2:
        raise ValueError("Input must be a 1D or 2D array.")
    if np.iscomplexobj(arr):
        cplx = np.any(arr.imag != 0, axis=0)
        arr = np.column_stack([arr.real, arr.imag[:, cplx]]) if np.any(cplx) else arr.real
    return arr.astype(float)


def _prepare_bandwidths(h, dims: int, options: dict, default_range) -> np.ndarray:
    """Shape and validate bandwidths (api.py:32-70).  default_range is the
    (dims, 2) range array or a callable producing it (called only when h is None)."""
    if h is None:
        dh = options.get("dh", None)
        if dh is None:
            dh = 20 if dims == 1 else [20] * dims
        bw = get_hlist(dh, default_range() if callable(default_range) else default_range)
    else:
        a = np.asarray(h.cpu() if isinstance(h, torch.Tensor) else h, dtype=float)
        if a.ndim == 0:
            bw = np.full((1, dims), a.item())
        elif a.ndim == 1:
            if a.size != dims:
                raise ValueError("Scalar bandwidth must match the number of dimensions.")
            bw = a.reshape(1, -1)
        else:
            if a.shape[1] != dims:
                raise ValueError(
                    "Bandwidth array must have the same dimensionality as the input data.")
            bw = a.reshape(-1, dims)
    if bw.shape[0] == 0:
        raise ValueError("Bandwidth array is empty: give at least one bandwidth.")
    if not np.all(np.isfinite(bw)):
        raise ValueError("Bandwidth values must be finite (no NaN or Inf).")
    if np.any(bw <= 0):
        raise ValueError(
            "Bandwidth values must be strictly positive. "
            f"Found non-positive values: {bw[bw <= 0]}")
    return bw


def _one_se_rule(scores, errs, bandwidths, prefer_smaller: bool, dstd: float = 1.0) -> int:
    """1-SE selection over finite scores; ties on sum(h) broken by the smaller
    score (api.py:73-164)."""
    scores = np.asarray(scores, dtype=float)
    errs = np.asarray(errs, dtype=float)
    bandwidths = np.asarray(bandwidths, dtype=float)
    valid = np.isfinite(scores)
    if not np.any(valid):
        raise ValueError(
            "All GCV scores are NaN or Inf. Regression failed for all bandwidths. "
            "Try: (1) Use larger bandwidth h, (2) Increase grid points, "
            "(3) Check for data quality issues")
    vidx = np.where(valid)[0]
    vs, ve = scores[valid], errs[valid]
    if prefer_smaller:
        best = int(np.argmin(vs))
        cand = vs <= vs[best] + dstd * ve[best]
        cidx = vidx[cand]
        if cidx.size == 0:
            return int(vidx[best])
        sums = bandwidths[cidx].sum(axis=1)
        ties = np.where(sums == sums.max())[0]
        return int(cidx[ties[int(np.argmin(vs[cand][ties]))]])
    best = int(np.argmax(vs))
    cand = vs >= vs[best] - dstd * ve[best]
    cidx = vidx[cand]
    if cidx.size == 0:
        return int(vidx[best])
    return int(cidx[np.argmax(bandwidths[cidx].sum(axis=1))])


def _check_finite(a, what: str, fn: str):
    if isinstance(a, torch.Tensor):
        t = torch.view_as_real(a) if a.is_complex() else a
        nan = bool(torch.isnan(t).any())
        inf = bool(torch.isinf(t).any())
    else:
        nan = bool(np.any(np.isnan(a)))
        inf = bool(np.any(np.isinf(a)))
    if nan:
        raise ValueError(f"{what} contains NaN values. Please remove or impute NaN values "
                         f"before calling {fn}.")
    if inf:
        raise ValueError(f"{what} contains Inf values. Please remove or handle Inf values "
                         f"before calling {fn}.")


def _check_accuracy(acc: int, dims: int) -> None:
    """NUFFT accuracies the engine runs: 0 (nearest-cell binning) to 6 on the
    L-periodic grid, 7..12 on the oversampled grid."""
    if acc < 0:
        raise ValueError("options['accuracy'] must be a non-negative integer.")
    if acc > 12:
        raise NotImplementedError("accuracy > 12 is not supported on the GPU")


def _check_axes(stats, xrange=None):
    """A dimension without spread gives a raw grid axis linspace(v, v, N) that
    is not strictly ascending; the reference's RegularGridInterpolator on it
    (api.py:686-696, regression.py:672-713) raises this ValueError."""
    if xrange is not None:
        return
    flat = np.flatnonzero(~(np.asarray(stats.xmax) > np.asarray(stats.xmin)))
    if flat.size:
        raise ValueError(f"The points in dimension {int(flat[0])} must be strictly ascending "
                         "or descending")


def _check_cubic(N):
    """scipy's cubic RegularGridInterpolator (the reference's default for 2-D /
    3-D under 5000 points, api.py:588-594) needs four nodes per axis."""
    for k, nk in enumerate(N):
        if nk < 4:
            raise ValueError(f"There are {nk} points in dimension {k}, but method cubic "
                             "requires at least  4 points per dimension.")


MAX_VALUE_CHANNELS = 32


def _check_channels(ny: int, K: int) -> None:
    """The per-node solve reduces at most 32 value channels (response columns,
    complex ones counted twice, plus the Hutchinson probes) in one launch; a
    larger set is rejected before any device work."""
    if ny + K > MAX_VALUE_CHANNELS:
        raise NotImplementedError(
            f"the GPU engine solves at most {MAX_VALUE_CHANNELS} value channels per fit: "
            f"{ny} real response column(s) + num_dof_sample={K} probes = {ny + K}")


def _grid_shape(opts: dict, n: int, dims: int) -> Tuple[int, ...]:
    """Grid nodes per dimension: options['N'] or ceil(n^(1/d)) (model.py:78-83, core.py:76-82)."""
    N = opts.get("N", None)
    if N is None:
        dn = int(np.ceil(n ** (1.0 / dims)))
        N = [dn] * dims
    elif np.isscalar(N):
        N = [int(N)] * dims
    else:
        N = [int(v) for v in np.asarray(N).ravel()]
    if len(N) != dims:
        raise ValueError("options['N'] must give one grid size per dimension.")
    if min(N) < 2:
        raise ValueError("The GPU engine needs at least 2 grid nodes per dimension.")
    return tuple(N)


# ---------------------------------------------------------------------------
# Hutchinson probes: NumPy MT19937 stream, cached on the device
# ---------------------------------------------------------------------------

_PROBES: Dict[tuple, tuple] = {}


def _probes(n: int, K: int, seed, random_matrix, dev):
    """Probe matrix p (n, K) = Generator(MT19937(seed)).standard_normal((n, K))
    (regression.py:1184-1194) and p'p per column, on the host and on the device.
    The generator fills the matrix row by row from one stream, so the matrix
    of n rows is the first n rows of the matrix of any larger row count with
    the same K and seed: one matrix per (K, seed), at the largest n asked so
    far, is kept on the device and every smaller n takes a prefix of it.  A
    caller-supplied random_matrix is used as is.  Returns (p on the device,
    p'p on the host, p'p on the device)."""
    if random_matrix is not None:
        p = np.asarray(random_matrix.cpu() if isinstance(random_matrix, torch.Tensor)
                       else random_matrix, dtype=float)
        if p.ndim != 2 or p.shape[0] != n:
            raise ValueError(f"random_matrix has wrong shape: expected ({n}, *), got {p.shape}")
        pp = np.sum(p * p, axis=0)
        return (torch.as_tensor(p, device=dev).contiguous(), pp,
                torch.as_tensor(pp, device=dev))
    key = (n, K, int(seed), str(dev))
    hit = _PROBES.get(key)
    if hit is None:
        base_h, base = _probe_matrix(n, K, seed, dev)
        pp = np.sum(base_h[:n] * base_h[:n], axis=0)
        hit = (base[:n], pp, torch.as_tensor(pp, device=dev))
        if len(_PROBES) > 4096:
            _PROBES.clear()
        _PROBES[key] = hit
    return hit


_PROBE_BASE: Dict[tuple, tuple] = {}


def _probe_matrix(n: int, K: int, seed, dev):
    """The probe matrix of (K, seed) with at least n rows, on the host and on
    the device (regenerated larger when n exceeds the rows held)."""
    key = (K, int(seed), str(dev))
    hit = _PROBE_BASE.get(key)
    if hit is None or hit[0].shape[0] < n:
        p = np.random.Generator(np.random.MT19937(seed)).standard_normal((n, K))
        hit = (p, torch.as_tensor(p, device=dev).contiguous())
        if len(_PROBE_BASE) > 4:
            _PROBE_BASE.clear()
        _PROBE_BASE[key] = hit
    return hit


def clear_caches():
    """Drop the cached probe matrices, conv plans and host geometry / bandwidth tables."""
    _PROBES.clear()
    _PROBE_BASE.clear()
    engine._PLANS.clear()
    engine._GEOMS.clear()
    engine._BTABS.clear()
    engine._SCHED.clear()


# ---------------------------------------------------------------------------
# interpolant returned as fpp_yhat / fpp
# ---------------------------------------------------------------------------

class GridInterpolant:
    """Multilinear interpolant of grid values (RegularGridInterpolator 'linear'
    semantics, linear extrapolation outside the grid) evaluated by the
    flpr_interp_eval kernel.  Grid values stay on the device."""

    def __init__(self, geo, axes, values: torch.Tensor, grid_shape, cplx: bool, ncol: int,
                 as_numpy: bool, squeeze: bool):
        self._geo = geo
        self.grid_axes = tuple(axes)
        self._vals = values  # (ncol_real, nodes) float64 on device
        self._shape = tuple(grid_shape)
        self._cplx = cplx
        self._ncol = ncol
        self._as_numpy = as_numpy
        self._squeeze = squeeze

    @property
    def values(self) -> np.ndarray:
        v = self._vals.cpu().numpy()
        if self._cplx:
            v = v[0::2] + 1j * v[1::2]
        v = np.moveaxis(v.reshape((v.shape[0],) + self._shape), 0, -1)
        return v[..., 0] if self._squeeze else v

    def device_values(self) -> torch.Tensor:
        """The grid values as a device tensor, laid out like ``values``."""
        v = self._vals
        if self._cplx:
            v = torch.complex(v[0::2], v[1::2])
        v = torch.movedim(v.reshape((v.shape[0],) + self._shape), 0, -1)
        return v[..., 0] if self._squeeze else v

    def __call__(self, points):
        dev = self._vals.device
        if isinstance(points, torch.Tensor):
            pts = points.to(device=dev, dtype=torch.float64)
        else:
            pts = torch.as_tensor(np.atleast_2d(np.asarray(points, dtype=float)), device=dev)
        if pts.ndim == 1:
            pts = pts[:, None] if len(self._shape) == 1 else pts[None, :]
        engine.check_points(self._geo, pts)
        pts = pts.contiguous()
        out, _ = engine.interp_eval(self._geo, pts, self._vals, self._vals.shape[0])
        if self._cplx:
            out = torch.complex(out[:, 0::2], out[:, 1::2])
        if self._squeeze:
            out = out[:, 0]
        return out.cpu().numpy() if self._as_numpy else out


# ---------------------------------------------------------------------------
# cv_fastlpr
# ---------------------------------------------------------------------------

def _point_sums(fit, xd, yd, probes, ny: int, K: int, gcv_method: str, dof_method: str):
    """Probe dots and residual sums from the grid solutions interpolated at the
    points (the reference's direct route, regression.py:1256-1280 and
    api.py:661-723), used when either interpolation is cubic: the spline
    coefficients and the sums at the points in two kernels
    (engine.point_sums).  Returns the fit with its m_y narrowed to the y
    columns and its red buffer filled in the layout flpr_select reads (y
    dots zero, residual sums in place of the quadratic forms), and the
    matching all-zero sum of squares."""
    dh, nval, nodes = fit.m_y.shape
    red = torch.zeros((dh * nval + dh * ny,), dtype=torch.float64, device=xd.device)
    engine.point_sums(fit.geo, xd, yd, ny, probes, K, fit.m_y, dh, gcv_method == "cubic",
                      dof_method == "cubic", red)
    fit.m_y = fit.m_y[:, :ny, :]
    fit.red = red
    return fit, torch.zeros((ny,), dtype=torch.float64, device=xd.device)


def _y_channels(y_arr, dev):
    """(n, dy) real/complex -> (n, ny) float64 device, complex as interleaved (re, im)."""
    if isinstance(y_arr, torch.Tensor):
        t = y_arr.to(dev)
        if t.is_complex():
            t = torch.view_as_real(t.to(torch.complex128)).reshape(t.shape[0], -1)
        return t.to(torch.float64).contiguous()
    if np.iscomplexobj(y_arr):
        yc = np.asarray(y_arr, dtype=np.complex128)
        return torch.as_tensor(yc.view(np.float64).reshape(yc.shape[0], -1), device=dev).contiguous()
    return torch.as_tensor(np.asarray(y_arr, dtype=np.float64), device=dev).contiguous()


def _is_complex(a) -> bool:
    return bool(a.is_complex()) if isinstance(a, torch.Tensor) else bool(np.iscomplexobj(a))


def _queue_inputs(x, y, x_arr, y_arr, n: int, variance: bool, prior, stats_pending=None):
    """x and the response channels as float64 on the device, and the x
    statistics pass queued behind them (not read back).  Returns (dev, xd,
    yd, yd_in, y_raw_orig, pending): yd is log(y + 1/n) for a variance fit
    (api.py:390-396), yd_in the response before it, y_raw_orig its first
    column (variance fits).  pending is ("y", sums) when prior holds the x
    side of a fit of this same x (then only y's sums of squares and
    non-finite count are queued, engine.y_stats), else ("x", the queued x
    statistics pass, or stats_pending when the caller queued it already for
    these same device x and y), and None when y has too many columns for the
    pass."""
    dev = x.device if _is_dev(x) else (y.device if _is_dev(y) else _device())
    if not torch.cuda.is_available():
        _device()
    xd = (x_arr.to(dev) if isinstance(x_arr, torch.Tensor)
          else torch.as_tensor(x_arr, device=dev)).to(torch.float64).contiguous()
    yd = _y_channels(y_arr, dev)
    y_raw_orig = yd[:, 0].contiguous() if variance else None
    yd_in = yd
    if variance:
        yd = torch.log(yd + 1.0 / n)  # api.py:390-396
    xs = prior.get("xshare") if prior is not None else None
    if yd.shape[1] > 32:
        pending = None
    elif xs and xs.get("x") is xd and prior["key"][:3] == (xd.data_ptr(), xd._version, xd.shape):
        pending = ("y", engine.y_stats(yd))
    elif stats_pending is not None and not variance:
        pending = ("x", stats_pending)
    else:
        pending = ("x", engine.x_stats_launch(xd, yd))
    return dev, xd, yd, yd_in, y_raw_orig, pending


def cv_fastlpr(x, y, h: Optional[Sequence[Sequence[float]]] = None,
               options: Optional[dict] = None) -> RegressionOutput:
    """Local polynomial regression (order 0/1/2) with GCV bandwidth selection
    (api.py:167-832), computed on the GPU.

    x (N, d), y (N,) / (N, dy) / (1, N), real or complex; h scalar, (d,) or (k, d);
    options as the reference: order, regularization, num_dof_sample, dof_seed,
    random_matrix, y_type_out ('mean' | 'variance'), calc_dof, gcv_selection,
    dstd, accuracy, N, xrange, flag_power2, dof_interp_method,
    gcv_interp_method, interp_method; plus ``precision`` ('fp32' | 'fp64') and
    ``bandwidth_bracket`` ('full' | 'learned').
    """
    finish, _ = _fit(x, y, h, options)
    return finish()


def cv_fastlpr_pair(x, y, h: Optional[Sequence[Sequence[float]]] = None,
                    options: Optional[dict] = None,
                    variance_options: Optional[dict] = None):
    """The heteroscedastic two-pass fit (examples/example_hetero.py:70-135):

        mu    = cv_fastlpr(x, y, h, options)
        sigma = cv_fastlpr(x, (y - mu.yhat)**2, h,
                           {**options, **variance_options, 'y_type_out': 'variance'})

    returned as (mu, sigma), equal to those two calls.  The degrees of freedom
    of a fit depend only on x, the bandwidths, the order, the grid and the
    probes, not on y (regression.py:1095-1354), so when the variance options
    leave all of those unchanged the second pass reuses the first pass's
    probe sums instead of spreading, convolving and solving the 10 probe
    channels again.  Real single-column y only.
    """
    as_numpy = not (_is_dev(x) or _is_dev(y))
    x_arr = _as_column_matrix(x)
    dev = x.device if _is_dev(x) else (y.device if _is_dev(y) else _device())
    xd = (x_arr.to(dev) if isinstance(x_arr, torch.Tensor)
          else torch.as_tensor(x_arr, device=dev)).to(torch.float64).contiguous()
    yt = y if isinstance(y, torch.Tensor) else torch.as_tensor(np.asarray(y))
    if yt.is_complex() or (yt.dim() == 2 and yt.shape[0] > 1 and yt.shape[1] > 1):
        raise NotImplementedError("cv_fastlpr_pair takes a real single-column response")
    yd = yt.to(dev).to(torch.float64).reshape(-1).contiguous()
    # the statistics pass goes first (x and y are already on the device), so
    # that the option handling of the first fit overlaps it; both passes are
    # queued before either result is read back, so the device runs them back
    # to back
    pending = engine.x_stats_launch(xd, yd.view(-1, 1))
    finish_mu, share = _fit(xd, yd, h, options, as_numpy=False, keep_x=True,
                            stats_pending=pending)
    vopts = dict(options or {})
    vopts.update(variance_options or {})
    vopts["y_type_out"] = "variance"
    r2 = (yd - share["yhat"][:, 0]) ** 2
    finish_sigma, share_v = _fit(xd, r2, h, vopts, prior=share, as_numpy=False)
    sel = (share["sel_out"], share_v["sel_out"])
    share.clear()
    if sel[0] is not None and sel[1] is not None:
        # both passes' selection results in one read (one host synchronisation)
        k = sel[0].numel()
        host = torch.empty((k + sel[1].numel(),), dtype=torch.float64, pin_memory=True)
        host[:k].copy_(sel[0], non_blocking=True)
        host[k:].copy_(sel[1], non_blocking=True)
        torch.cuda.current_stream(xd.device).synchronize()
        hv = host.numpy()
        mu, sigma = finish_mu(hv[:k]), finish_sigma(hv[k:])
    else:
        mu = finish_mu()
        sigma = finish_sigma()
    if as_numpy:
        mu, sigma = _to_numpy(mu), _to_numpy(sigma)
    return mu, sigma


def _to_numpy(r: RegressionOutput) -> RegressionOutput:
    def cv(t):
        return t.cpu().numpy() if isinstance(t, torch.Tensor) else t
    md = dict(r.metadata)
    for k in ("fitted_at_samples", "residuals"):
        md[k] = cv(md[k])
    interp = r.fpp_yhat
    if isinstance(interp, GridInterpolant):
        interp._as_numpy = True
    return RegressionOutput(yhat=cv(r.yhat), grid=r.grid, h=r.h, gcv_yhat=r.gcv_yhat,
                            fpp_yhat=interp, metadata=md, dof=r.dof, dof_stderr=r.dof_stderr,
                            d=r.d, s0=cv(r.s0))


def _fit(x, y, h, options, prior=None, as_numpy=None, keep_x=False, stats_pending=None):
    """cv_fastlpr's body, split in two: the call queues all device work and
    returns (finish, share); finish() reads the results back (one host
    synchronisation) and builds the RegressionOutput.  share identifies the
    x side of the fit (the key), holds the probe sums and the fitted values
    (share['yhat'], n x ny real columns) on the device and, with keep_x, the
    rest of the x-side state (engine.run_grid share_out), so that a later fit
    of the same x with a matching key (prior) only spreads, convolves and
    solves its own response columns (engine.run_grid_y) and needs no x
    statistics.  stats_pending: the statistics pass of exactly these device
    x and y, already queued (engine.x_stats_launch), used instead of a new
    one."""
    if as_numpy is None:
        as_numpy = not (_is_dev(x) or _is_dev(y))
    x_arr = _as_column_matrix(x)
    if isinstance(y, torch.Tensor):
        y_arr = y
    else:
        y_arr = np.asarray(y)
    if y_arr.ndim == 1:
        y_arr = y_arr[:, None]
    elif y_arr.ndim == 2:
        if y_arr.shape[0] == 1 and y_arr.shape[1] > 1:
            y_arr = y_arr.T
    else:
        raise ValueError(f"Response y must be 1D or 2D array. Got {y_arr.ndim}D array.")
    dy = y_arr.shape[1]
    if y_arr.shape[0] != x_arr.shape[0]:
        raise ValueError(
            f"Response y must have same number of samples as x. "
            f"x has {x_arr.shape[0]} samples, y has {y_arr.shape[0]} samples.")
    if x_arr.shape[0] == 0:
        raise ValueError("Input x must have at least one sample.")
    if y_arr.shape[0] == 0:
        raise ValueError("Response y must have at least one sample.")
    if not isinstance(x_arr, torch.Tensor):
        _check_finite(x_arr, "Input x", "cv_fastlpr")
    if not isinstance(y_arr, torch.Tensor):
        _check_finite(y_arr, "Response y", "cv_fastlpr")

    opts = set_defaults(options, order=0, regularization=1e-8, num_dof_sample=10,
                        kernel_type="gaussian", y_type_out="mean", calc_dof=True,
                        gcv_selection="first")
    n, dims = x_arr.shape
    if opts.get("accuracy") is None:
        opts["accuracy"] = max(6 - int(np.ceil(np.log10(n))), 4)
    order = int(opts["order"])
    if order not in (0, 1, 2):
        raise ValueError("Polynomial order must be 0, 1, or 2.")
    if dims > 10:
        raise ValueError(f"dx={dims} exceeds maximum supported dimension (10). "
                         "High-dimensional regression suffers from the curse of dimensionality.")
    elif dims > 6:
        warnings.warn(f"dx={dims} is high-dimensional and may cause memory issues. "
                      "Consider dimensionality reduction techniques.", UserWarning)
    if dims > 3:
        raise NotImplementedError("the GPU engine supports 1-3 dimensions")
    if str(opts.get("kernel_type", "gaussian")).lower() not in ("gauss", "gaussian", "normal"):
        raise NotImplementedError("the GPU engine implements the Gaussian kernel")
    variance = str(opts.get("y_type_out", "mean")) == "variance"
    _check_accuracy(int(opts["accuracy"]), dims)
    # explicit bandwidths are validated before any device work; the default
    # range needs the x statistics
    bw_given = _prepare_bandwidths(h, dims, opts, None) if h is not None else None
    prec = str(opts.get("precision", "fp32")).lower()
    if prec not in ("fp32", "fp64", "mixed"):
        raise ValueError("options['precision'] must be 'fp32', 'fp64' or 'mixed'")
    fp64 = prec == "fp64"
    mixed = prec == "mixed"
    N = _grid_shape(opts, n, dims)
    y_is_complex = _is_complex(y_arr)
    # value channels of the solve, checked before any device work
    if bool(opts.get("calc_dof", True)):
        K_req = int(opts.get("num_dof_sample", 10))
        rm = opts.get("random_matrix", None)
        if rm is not None:
            rshape = np.shape(rm.cpu() if isinstance(rm, torch.Tensor) else rm)
            if len(rshape) == 2:
                K_req = int(rshape[1])
        _check_channels(dy * (2 if y_is_complex else 1), max(K_req, 0))
    # every check above runs before any device work
    dev, xd, yd, yd_in, y_raw_orig, pending = _queue_inputs(x, y, x_arr, y_arr, n, variance,
                                                            prior, stats_pending)
    ny = yd.shape[1]
    if ny > 32:
        raise NotImplementedError("the GPU engine takes at most 32 real response columns")
    bad_y_dev = None
    if pending[0] == "y":
        # the same x as the fit that left its x side in prior: its statistics,
        # and y's sums of squares and non-finite check on the device (read
        # back with the results)
        s1 = prior["xshare"]["st"]
        ys = pending[1]
        stats = engine.Stats(s1.xmin, s1.xmax, s1.mean, s1.std, 0, 0, None, ys[2:])
        bad_y_dev = ys[1:2]
    else:
        # one device pass: x statistics, non-finite checks of device inputs, sum y^2
        stats = engine.x_stats_read(pending[1])
        if stats.bad_x:
            _check_finite(x_arr, "Input x", "cv_fastlpr")
        if stats.bad_y and isinstance(y_arr, torch.Tensor):
            _check_finite(y_arr, "Response y", "cv_fastlpr")
    # default range 0.1..1 x std(ddof=1) from the device statistics (helpers.py:74-92)
    bandwidths = bw_given if bw_given is not None else _prepare_bandwidths(
        h, dims, opts, lambda: np.column_stack((0.1 * stats.std, 1.0 * stats.std)))
    reg = float(opts["regularization"])
    calc_dof = bool(opts.get("calc_dof", True))
    xrange = opts.get("xrange", None)
    _check_axes(stats, xrange)
    flag_power2 = bool(opts.get("flag_power2", True))
    lin_default = "linear" if (dims == 1 or n >= 5000) else "cubic"

    share = {"key": None, "probe_dots": None}
    reuse = False
    binfo = None
    if not calc_dof:
        method = opts.get("interp_method", "linear")
        if method != "linear":
            raise NotImplementedError("only linear interpolation is implemented on the GPU")
        fit = engine.run_grid(xd, yd, ny, None, 0, bandwidths, order, N, int(opts["accuracy"]),
                              fp64, reg, xrange, flag_power2, stats=stats, with_ones_eps=False,
                              y_pairs=y_is_complex, mixed=mixed)
        sel = 0
        sel_dev = sel_out = None
    else:
        dof_method = str(opts.get("dof_interp_method", lin_default))
        gcv_method = str(opts.get("gcv_interp_method", lin_default))
        for key, meth in (("dof_interp_method", dof_method), ("gcv_interp_method", gcv_method)):
            if meth not in ("linear", "cubic"):
                raise NotImplementedError(f"{key}={meth!r}: the GPU engine implements "
                                          "'linear' and 'cubic'")
        cubic = "cubic" in (dof_method, gcv_method)
        if cubic:
            _check_cubic(N)
        if cubic and xrange is not None:
            raise NotImplementedError("cubic GCV/DOF interpolation with options['xrange'] "
                                      "is not implemented on the GPU")
        K = int(opts.get("num_dof_sample", 10))
        if K < 1:
            raise ValueError(
                "All GCV scores are NaN or Inf. Regression failed for all bandwidths. "
                "num_dof_sample must be at least 1 for the GCV degrees of freedom.")
        rm = opts.get("random_matrix", None)
        probes, _pp, pp_dev = _probes(n, K, opts.get("dof_seed", 42), rm, dev)
        K = probes.shape[1]
        # S and the probes are spread at the sample-size accuracy whatever the
        # option says (api.py:569-576 calls compute_design_matrix without one,
        # regression.py:86-87, 756-758); only y takes the option's (api.py:632-642)
        acc_n = max(6 - int(np.ceil(np.log10(n))), 4)
        key = (xd.data_ptr(), xd._version, xd.shape, np.asarray(bandwidths).tobytes(), order,
               tuple(N), acc_n, int(opts["accuracy"]), prec, reg, repr(xrange), flag_power2, K,
               probes.data_ptr(), y_is_complex, ny)
        reuse = (prior is not None and not cubic and prior["key"] == key)
        xshare = {} if (keep_x and not cubic) else None
        agg = str(opts.get("gcv_selection", "first"))
        bracket_mode = str(opts.get("bandwidth_bracket", "full")).lower()
        if bracket_mode not in ("full", "learned"):
            raise ValueError("options['bandwidth_bracket'] must be 'full' or 'learned'")
        use_bracket = (bracket_mode == "learned" and dims == 2 and not cubic and not reuse
                       and prior is None and not keep_x and not y_is_complex and ny == 1
                       and not variance and xrange is None)

        def sweep(run_h, select_h=None):
            """run_grid over the bandwidths run_h, then the device selection."""
            if reuse and prior.get("xshare"):
                f = engine.run_grid_y(prior["xshare"], xd, yd, ny, y_is_complex)
            else:
                f = engine.run_grid(xd, yd, ny, None if reuse else probes, 0 if reuse else K,
                                    run_h, order, N, acc_n,
                                    fp64, reg, xrange, flag_power2, want_rss=True, stats=stats,
                                    readback=False, keep_all=cubic,
                                    accuracy_y=int(opts["accuracy"]), y_pairs=y_is_complex,
                                    share_out=xshare, mixed=mixed, select_h=select_h)
            if xshare:
                share["xshare"] = xshare
            dh_ = f.h.shape[0]
            if reuse:
                # y dots, the first pass's probe dots, then the quadratic forms
                red = torch.empty((dh_ * (ny + K) + dh_ * ny,), dtype=torch.float64, device=dev)
                dots = red[:dh_ * (ny + K)].view(dh_, ny + K)
                dots[:, :ny] = f.red[:dh_ * ny].view(dh_, ny)
                dots[:, ny:] = prior["probe_dots"]
                red[dh_ * (ny + K):] = f.red[dh_ * ny:]
                f.red = red
            elif not cubic:
                share["key"] = key
                share["probe_dots"] = f.red[:dh_ * (ny + K)].view(dh_, ny + K)[:, ny:].clone()
            ysq = stats.ysq_dev
            if cubic:
                f, ysq = _point_sums(f, xd, yd, probes, ny, K, gcv_method, dof_method)
            # Hutchinson DOF, GCV and the 1-SE choice stay on the device; the
            # interpolation at the chosen bandwidth is queued behind them and
            # the host reads everything back once, at the end
            so, sd = engine.select(f, ysq, pp_dev, n, ny, K, y_is_complex, agg,
                                   float(opts.get("dstd", 0.0)))
            return f, so, sd

        binfo = None
        if use_bracket:
            from . import bracket
            binfo = {"n": n}
            fit, sel_out, sel_dev = sweep(bandwidths, bracket.selector(bracket.model(),
                                                                       stats.ysq_dev, binfo))
            if "indices" not in binfo:
                binfo = None
        else:
            fit, sel_out, sel_dev = sweep(bandwidths)
    if fit.warn:
        warnings.warn(fit.warn, stacklevel=2)

    geo = fit.geo
    dh = fit.h.shape[0]
    if variance and ny != 1:
        raise NotImplementedError("y_type_out='variance' takes a real single-column response")
    # yhat, residuals (and the variance back-transform) at the selected
    # bandwidth, on the device; the factor 1/d lands in the readback buffer
    fac_slot = sel_out[4 * dh + 3:4 * dh + 4] if (variance and sel_out is not None) else None
    out, resid_r, grid_var, fac = engine.fit_output(
        geo, xd, fit.m_y, ny, sel_dev, y_raw_orig if variance else yd, variance, fac_slot)
    if y_is_complex:
        yhat = torch.view_as_complex(out.view(n, dy, 2))
        resid = torch.view_as_complex(resid_r.view(n, dy, 2))
    else:
        yhat, resid = out, resid_r
    if dy == 1:
        yhat = yhat[:, 0]
    resid = resid.reshape(-1)
    share["yhat"] = out
    if sel_out is not None and bad_y_dev is not None:
        sel_out[4 * dh + 4:4 * dh + 5].copy_(bad_y_dev.reshape(1))

    def conv(t):
        return t.cpu().numpy() if as_numpy else t

    def finish(hv=None):
        r = _finish(fit, sel_out, fac if variance else None, bad_y_dev, yd_in, dh, opts,
                    order, geo, yhat, resid, grid_var, N, y_is_complex, dy, as_numpy,
                    reuse, conv, hv)
        if binfo is not None:
            # the minimum on an inner edge of the bracket may lie outside it:
            # then the full sweep decides
            s_b = int(r.gcv_yhat["id1se"])
            idx = binfo["indices"]
            if (s_b == 0 and idx[0] > 0) or (s_b == len(idx) - 1 and idx[-1] < binfo["n_full"] - 1):
                r = cv_fastlpr(x, y, h, dict(opts, bandwidth_bracket="full"))
                binfo["fallback"] = True
            r.metadata["bracket"] = {k: binfo[k] for k in ("indices", "center", "width",
                                                           "fallback", "hlist")}
        return r

    share["sel_out"] = sel_out
    return finish, share


def _finish(fit, sel_out, fac, bad_y_dev, y_in, dh, opts, order, geo, yhat, resid, grid_var, N,
            y_is_complex, dy, as_numpy, reuse, conv, hv=None):
    """The one host read of a fit (selection results, the variance scale, the
    deferred non-finite count of y) and the RegressionOutput; the grid values
    and S_00 of the selected bandwidth are views of the fit's arrays.  hv: the
    selection results already read back (sel_out's values on the host)."""
    d_norm = None
    if sel_out is not None:
        if hv is None:
            hv = sel_out.cpu().numpy()
        if bad_y_dev is not None and hv[4 * dh + 4] > 0:
            _check_finite(y_in, "Response y", "cv_fastlpr")
        if fac is not None:
            d_norm = 1.0 / float(hv[4 * dh + 3])
        sel = int(hv[4 * dh])
        if int(hv[4 * dh + 2]) == 0:
            raise ValueError(
                "All GCV scores are NaN or Inf. Regression failed for all bandwidths. "
                "Try: (1) Use larger bandwidth h, (2) Increase grid points, "
                "(3) Check for data quality issues")
        scores, errs = hv[:dh].copy(), hv[dh:2 * dh].copy()
        dof_all, dof_se = hv[2 * dh:3 * dh].copy(), hv[3 * dh:4 * dh].copy()
        sel, idmin = int(hv[4 * dh]), int(hv[4 * dh + 1])
        gcv_result = {"bandwidths": fit.h, "gcv_m": scores, "gcv_sd": errs, "idmin": idmin,
                      "id1se": sel, "h1se": fit.h[sel], "hmin": fit.h[idmin],
                      "dof": dof_all, "dof_stderr": dof_se}
        dof_sel, dof_err_sel = float(dof_all[sel]), float(dof_se[sel])
    else:
        sel = 0   # without DOF the first bandwidth is the result (api.py:436-470)
        if bad_y_dev is not None and float(bad_y_dev.item()) > 0:
            _check_finite(y_in, "Response y", "cv_fastlpr")
        if fac is not None:
            d_norm = 1.0 / float(fac.item())
        gcv_result = {"bandwidths": fit.h, "gcv_m": np.array([np.nan]),
                      "gcv_sd": np.array([np.nan]), "idmin": 0, "id1se": 0,
                      "h1se": fit.h[0], "hmin": fit.h[0], "dof": np.array([np.nan]),
                      "dof_stderr": np.array([np.nan])}
        dof_sel = dof_err_sel = np.nan

    grid_vals = grid_var.view(1, -1) if grid_var is not None else fit.m_y[sel]
    s0 = fit.s0_all[sel].reshape(N)
    interp = GridInterpolant(geo, geo.axes_raw, grid_vals, N, y_is_complex, dy, as_numpy,
                             dy == 1)
    metadata = {"options": opts, "order": order, "bandwidths_tested": fit.h,
                "fitted_at_samples": conv(yhat), "residuals": conv(resid),
                "mean": geo.mean, "std": geo.std, "dof": dof_sel, "dof_stderr": dof_err_sel,
                "dof_shared": reuse}
    return RegressionOutput(yhat=conv(yhat), grid=tuple(geo.axes_raw), h=fit.h[sel],
                            gcv_yhat=gcv_result, fpp_yhat=interp, metadata=metadata,
                            dof=dof_sel, dof_stderr=dof_err_sel, d=d_norm, s0=conv(s0))


# ---------------------------------------------------------------------------
# cv_fastkde
# ---------------------------------------------------------------------------

K0_GAUSS = 1.0 / np.sqrt(2.0 * np.pi)


def cv_fastkde(x, h: Optional[Sequence[Sequence[float]]] = None,
               options: Optional[dict] = None) -> KDEOutput:
    """Kernel density estimation with likelihood cross-validation (api.py:835-1184),
    computed on the GPU.

    The density grid of every bandwidth is the order-0 moment S_00 (the same
    spread and convolution as cv_fastlpr, with only the ones channel),
    normalised to integrate to one on the raw grid and clipped at zero.  LCV
    uses the leave-one-out identity f_-i(x_i) = n/(n-1) (f(x_i) - K(0)/(n prod h))
    and the 1-SE rule preferring the larger sum(h) with dstd = 1 (the
    reference passes no dstd to its selection).
    """
    as_numpy = not _is_dev(x)
    x_arr = _as_column_matrix(x)
    n, dims = x_arr.shape
    if n < 2:
        raise ValueError(f"Input x must have at least 2 observations. Got {n} observations.")
    if not isinstance(x_arr, torch.Tensor):
        _check_finite(x_arr, "Input x", "cv_fastkde")
    opts = set_defaults(options, order=0, calc_dof=False, y_type_out="mean")
    opts = set_defaults(opts, kernel_type="gaussian", verbose=False, dh=20)
    if opts.get("accuracy") is None:
        opts["accuracy"] = max(6 - int(np.ceil(np.log10(n))), 4)
    if opts.get("N") is None:
        an = int(np.ceil(n ** (1.0 / dims)))
        opts["N"] = [an] * dims if dims > 1 else an
    if dims > 3:
        raise NotImplementedError("the GPU engine supports 1-3 dimensions")
    if str(opts.get("kernel_type", "gaussian")).lower() not in ("gauss", "gaussian", "normal"):
        raise NotImplementedError("the GPU engine implements the Gaussian kernel")
    if opts.get("output_interp_method", "linear") != "linear":
        raise NotImplementedError("only linear interpolation is implemented on the GPU")
    _check_accuracy(int(opts["accuracy"]), dims)
    bw_given = _prepare_bandwidths(h, dims, opts, None) if h is not None else None
    prec = str(opts.get("precision", "fp32")).lower()
    if prec not in ("fp32", "fp64", "mixed"):
        raise ValueError("options['precision'] must be 'fp32', 'fp64' or 'mixed'")
    fp64 = prec == "fp64"
    mixed = prec == "mixed"
    N = _grid_shape(opts, n, dims)
    dev = x.device if _is_dev(x) else _device()
    xd = (x_arr.to(dev) if isinstance(x_arr, torch.Tensor)
          else torch.as_tensor(x_arr, device=dev)).to(torch.float64).contiguous()
    stats = engine.x_stats(xd, None)
    if stats.bad_x:
        _check_finite(x_arr, "Input x", "cv_fastkde")
    _check_axes(stats)
    bandwidths = bw_given if bw_given is not None else _prepare_bandwidths(
        h, dims, opts, lambda: np.column_stack((0.1 * stats.std, 1.0 * stats.std)))

    fit = engine.run_grid(xd, None, 0, None, 0, bandwidths, 0, N, int(opts["accuracy"]), fp64,
                          0.0, None, bool(opts.get("flag_power2", True)), stats=stats,
                          moments_only=True, kde=True)
    if fit.warn:
        warnings.warn(fit.warn, stacklevel=2)
    geo = fit.geo
    hl = fit.h
    dh = hl.shape[0]
    # normalise so that sum(f) dx = 1 on the raw grid, clip at zero (api.py:1062-1096)
    dx_grid = float(np.prod([ax[1] - ax[0] for ax in geo.axes_raw]))
    s = fit.s0_all.all()
    # row sums in chunks of the row: a plain sum over dim 1 reduces each long
    # row on few blocks
    chunks = math.gcd(s.shape[1], 4096)
    sdx = (s.reshape(dh, chunks, -1).sum(dim=2).sum(dim=1, keepdim=True) if chunks >= 64
           else s.sum(dim=1, keepdim=True)) * dx_grid
    # s / sdx where sdx > 0, else s; then clipped at zero, in place
    dens = torch.div(s, torch.where(sdx > 0, sdx, torch.ones_like(sdx))).clamp_min_(0.0)
    k0h = K0_GAUSS / np.prod(hl, axis=1)
    lv = engine.lcv(geo, xd, dens, k0h, fit.order_idx).cpu().numpy()
    scores, errs = lv[:dh].copy(), lv[dh:].copy()
    sel = _one_se_rule(scores, errs, hl, prefer_smaller=False)
    idmax = int(np.argmax(scores))
    unique_axes = [np.unique(hl[:, k]) for k in range(dims)]
    shape_lcv = tuple(len(a) for a in unique_axes)
    matrix = scores.reshape(shape_lcv) if int(np.prod(shape_lcv)) == len(scores) else None
    lcv_result = {"hlist": hl, "lcv_m": scores, "lcv_sd": errs, "idmax": idmax, "id1se": sel,
                  "hmax": hl[idmax], "h1se": hl[sel], "unique_axes": unique_axes,
                  "matrix": matrix}
    fsel = dens[sel]
    fhat_all = torch.movedim(dens.reshape((dh,) + tuple(N)), 0, -1)
    interp = GridInterpolant(geo, geo.axes_raw, fsel.reshape(1, -1).contiguous(), N, False, 1,
                             as_numpy, True)

    def conv(t):
        return t.cpu().numpy() if as_numpy else t

    metadata = {"options": opts, "bandwidths_tested": hl, "fhat_all": conv(fhat_all)}
    return KDEOutput(fhat=conv(fsel.reshape(N)), grid=tuple(geo.axes_raw), h=hl[sel],
                     lcv=lcv_result, fpp=interp, metadata=metadata)


# ---------------------------------------------------------------------------
# prediction and pointwise intervals
# ---------------------------------------------------------------------------

def fastlpr_predict(regression: RegressionOutput, x_new):
    """Evaluate a fitted regression at new points through its interpolant
    (predict.py:15-122): linear interpolation of the selected grid values with
    linear extrapolation, on the GPU; complex fits stay complex."""
    if regression.fpp_yhat is None:
        raise ValueError("Regression output does not contain an interpolant (fpp).")
    if isinstance(x_new, torch.Tensor):
        pts = x_new if x_new.ndim == 2 else x_new[:, None]
        return regression.fpp_yhat(pts)
    pts = np.asarray(x_new, dtype=float)
    if pts.ndim == 1:
        pts = pts[:, None]
    return np.asarray(regression.fpp_yhat(pts))


def _gauss_moment(k: int, squared: bool) -> float:
    """int phi(u) u^k du (= (k-1)!! for even k) or int phi(u)^2 u^k du
    (= Gamma((k+1)/2) / (2 pi) for even k); 0 for odd k."""
    if k % 2:
        return 0.0
    if squared:
        return math.gamma((k + 1) / 2.0) / (2.0 * math.pi)
    out = 1.0
    for j in range(k - 1, 0, -2):
        out *= j
    return out


def _nu(d: int, order: int) -> float:
    """Asymptotic variance constant e1' M_K^-1 M_K2 M_K^-1 e1 of the local
    polynomial estimator at an interior point for the product Gaussian kernel
    (confidence.py:24-46), from the kernel moments of each dimension."""
    if d not in (1, 2, 3) or order not in (0, 1, 2):
        raise ValueError(f"nu constant not available for d={d}, order={order}. "
                         "Supported: d in {1,2,3}, order in {0,1,2}.")
    pw = engine.poly_powers(d, order)
    nt = len(pw)
    MK = np.ones((nt, nt))
    MK2 = np.ones((nt, nt))
    for i in range(nt):
        for j in range(nt):
            for k in range(d):
                e = pw[i][k] + pw[j][k]
                MK[i, j] *= _gauss_moment(e, False)
                MK2[i, j] *= _gauss_moment(e, True)
    Minv = np.linalg.inv(MK)
    return float((Minv @ MK2 @ Minv)[0, 0])


def fastlpr_interval(mu: RegressionOutput, sigma, alpha: float = 0.05,
                     type: str = "confidence"):
    """Pointwise confidence (mean) or prediction (new observation) intervals on
    the grid (confidence.py:60-216): se^2 = sigma nu / (prod(h) s0), half width
    z se (confidence) or z sqrt(sigma + se^2) (prediction), z the standard
    normal 1 - alpha/2 quantile.  sigma is the variance fit
    (y_type_out='variance') on the same grid, or a number / array.  Returns
    (grid..., 2) [upper, lower], on the device when mu's grids are."""
    if not isinstance(mu, RegressionOutput):
        raise TypeError("mu must be a RegressionOutput object from cv_fastlpr")
    if mu.yhat is None:
        raise ValueError("mu must have .yhat field (fitted values)")
    if not isinstance(sigma, (RegressionOutput, float, int, np.ndarray, torch.Tensor)):
        raise TypeError("sigma must be RegressionOutput, float, or ndarray")
    if isinstance(sigma, RegressionOutput) and sigma.yhat is None:
        raise ValueError("sigma must have .yhat field (fitted values)")
    if not (0 < alpha < 1):
        raise ValueError(f"alpha must be in (0, 1), got {alpha}")
    kind = type.lower()
    if kind not in ("confidence", "prediction"):
        raise ValueError(f"type must be 'confidence' or 'prediction', got '{type}'")
    if mu.s0 is None:
        raise ValueError("mu.s0 is None. The RegressionOutput must contain s_0 "
                         "(zero-th kernel moment).")
    from statistics import NormalDist
    z = NormalDist().inv_cdf(1 - alpha / 2)
    dev = isinstance(mu.s0, torch.Tensor)

    def grid_of(r):
        if r.fpp_yhat is not None and hasattr(r.fpp_yhat, "device_values"):
            v = r.fpp_yhat.device_values()
            return v if dev else v.cpu().numpy()
        return r.yhat

    mu_values = grid_of(mu)
    if isinstance(sigma, RegressionOutput):
        sigma_values = grid_of(sigma)
    elif dev:
        sigma_values = torch.as_tensor(sigma, dtype=torch.float64, device=mu.s0.device)
    else:
        sigma_values = np.asarray(sigma, dtype=float)
    xp = torch if dev else np
    if dev:
        sigma_pos = torch.clamp_min(sigma_values, 0.0)
        s0 = torch.clamp_min(mu.s0, 1e-10)
    else:
        sigma_pos = np.maximum(sigma_values, 0)
        s0 = np.maximum(mu.s0, 1e-10)
    if tuple(sigma_pos.shape) != tuple(s0.shape):
        raise ValueError(f"Shape mismatch: sigma has shape {tuple(sigma_pos.shape)} but s0 has "
                         f"shape {tuple(s0.shape)}. Ensure mu and sigma use the same grid.")
    if tuple(mu_values.shape) != tuple(s0.shape):
        raise ValueError(f"Shape mismatch: mu has shape {tuple(mu_values.shape)} but s0 has "
                         f"shape {tuple(s0.shape)}. Ensure mu grid values and s0 are aligned.")
    d = len(mu.grid)
    ell = int(mu.metadata["order"]) if mu.metadata else 0
    prod_h = float(np.prod(np.atleast_1d(np.asarray(mu.h, dtype=float))))
    se_sq = sigma_pos * _nu(d, ell) / (prod_h * s0)
    if kind == "confidence":
        half = z * xp.sqrt(se_sq)
    else:
        half = z * xp.sqrt(sigma_pos + se_sq)
    return xp.stack([mu_values + half, mu_values - half], -1)
