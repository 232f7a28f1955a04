"""NumPy restatement of the reference fastLPR hot path (oracle; test infrastructure only).

Citations are ``path:line`` relative to the reference checkout
(``fastLPR_py/src/fastlpr/...`` abbreviated to the module name).

The dataflow restated here is the reference's, step by step:

    standardize + grid            model.py:58-93, core.py:64-103
    kernel-moment spectra         convolution.py:318-652, 747-794
    knot mapping                  convolution.py:138-199, nufftn_type1.py:140-176
    Gaussian spread + FFT         nufftn_type1.py:378-772 (spreading_fused.py:67-94)
    spectrum product + IFFT       convolution.py:797-953
    S, T, solve                   regression.py:31-159, 162-277, 1357-1553
    Hutchinson DOF                regression.py:1095-1354
    GCV, 1-SE                     regression.py:280-471, api.py:73-164, 664-740
    final interpolation           api.py:742-786
    KDE + LCV                     api.py:835-1184

``mode="ref"`` keeps the reference's float32 arithmetic at accuracy<5
(type_utils.py:145-168, convolution.py:904-906); ``mode="f64"`` keeps float64
throughout.  Accumulation of the spread is done in float64 (``np.bincount``)
and then rounded to the grid dtype, which is what the reference's Cython
backend does (nufftn_type1.py:618-628).
"""
from __future__ import annotations

import math
import warnings
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import scipy.fft as sfft
from scipy.interpolate import RegularGridInterpolator

SQRT_2PI = math.sqrt(2.0 * math.pi)


# ----------------------------------------------------------------------------
# options, bandwidths, selection
# ----------------------------------------------------------------------------

def set_defaults(options: Optional[dict], **defaults) -> dict:
    """Shallow copy + setdefault (utils/__init__.py:18-60)."""
    out = dict(options) if options is not None else {}
    for k, v in defaults.items():
        out.setdefault(k, v)
    return out


def as_column_matrix(x) -> np.ndarray:
    """1D -> column; complex columns split into [real, imag] (utils/helpers.py:31-71)."""
    arr = np.asarray(x)
    if arr.ndim == 1:
        arr = arr[:, None]
    if arr.ndim != 2:
        raise ValueError("Input must be a 1D or 2D array.")
    if np.iscomplexobj(arr):
        cplx = np.any(arr.imag != 0, axis=0)
        arr = np.column_stack([arr.real, arr.imag[:, cplx]]) if np.any(cplx) else arr.real
    return arr.astype(float)


def default_hrange(x: np.ndarray) -> np.ndarray:
    """[0.1, 1.0] x std(ddof=1) per dimension (utils/helpers.py:74-92)."""
    s = np.std(x, axis=0, ddof=1)
    s[s == 0] = 1.0
    return np.column_stack((0.1 * s, 1.0 * s))


def get_hlist(n, rng, spacing=None) -> np.ndarray:
    """Log-spaced bandwidth grid, Cartesian product in Fortran order (bandwidth.py:110-137)."""
    r = np.asarray(rng, dtype=float)
    if r.ndim == 1:
        r = r.reshape(1, 2)
    d = r.shape[0]
    counts = [int(n)] * d if np.isscalar(n) else [int(v) for v in n]
    if len(counts) != d:
        raise ValueError("n must match the number of rows in range.")
    if spacing is None:
        spacing = lambda lo, hi, num: np.logspace(np.log10(lo), np.log10(hi), num)
    axes = [spacing(lo, hi, k) if k > 1 else np.array([lo], dtype=float)
            for (lo, hi), k in zip(r, counts)]
    mesh = np.meshgrid(*axes, indexing="ij")
    return np.column_stack([m.ravel("F") for m in mesh])


def prepare_bandwidths(h, dims: int, options: dict, default_range) -> np.ndarray:
    """Shape/validate bandwidths (api.py:32-70)."""
    if h is None:
        dh = options.get("dh", None)
        if dh is None:
            dh = 20 if dims == 1 else [20] * dims
        bw = get_hlist(dh, default_range)
    else:
        a = np.asarray(h, dtype=float)
        if a.ndim == 0:
            bw = np.full((1, dims), a.item())
        elif a.ndim == 1:
            if a.size != dims:
                raise ValueError("Scalar bandwidth must match the number of dimensions.")
            bw = a.reshape(1, -1)
        else:
            if a.shape[1] != dims:
                raise ValueError("Bandwidth array must have the same dimensionality as the input data.")
            bw = a.reshape(-1, dims)
    if not np.all(np.isfinite(bw)):
        raise ValueError("Bandwidth values must be finite (no NaN or Inf).")
    if np.any(bw <= 0):
        raise ValueError("Bandwidth values must be strictly positive.")
    return bw


def one_se_rule(scores, errs, bandwidths, prefer_smaller: bool, dstd: float = 1.0) -> int:
    """1-SE selection ignoring non-finite scores (api.py:73-164)."""
    scores = np.asarray(scores, float)
    errs = np.asarray(errs, float)
    bandwidths = np.asarray(bandwidths, float)
    valid = np.isfinite(scores)
    if not np.any(valid):
        raise ValueError("All GCV scores are NaN or Inf.")
    vidx = np.where(valid)[0]
    vs, ve = scores[valid], errs[valid]
    if prefer_smaller:
        b = int(np.argmin(vs))
        thr = vs[b] + dstd * ve[b]
        cmask = vs <= thr
        cidx = vidx[cmask]
        if len(cidx) == 0:
            return int(vidx[b])
        sums = np.sum(bandwidths[cidx], axis=1)
        cs = vs[cmask]
        tie = np.where(sums == sums.max())[0]
        return int(cidx[tie[int(np.argmin(cs[tie]))]])
    b = int(np.argmax(vs))
    thr = vs[b] - dstd * ve[b]
    cmask = vs >= thr
    cidx = vidx[cmask]
    if len(cidx) == 0:
        return int(vidx[b])
    sums = np.sum(bandwidths[cidx], axis=1)
    return int(cidx[np.argmax(sums)])


# ----------------------------------------------------------------------------
# geometry
# ----------------------------------------------------------------------------

def accuracy_for(n: int) -> int:
    """max(6 - ceil(log10 n), 4) (api.py:345-348)."""
    return max(6 - int(np.ceil(np.log10(n))), 4)


def grid_params(N: int, acc: int):
    """Spread width, oversampled size, Gaussian width, ratio (kernel.py:97-156)."""
    R = 1 if acc <= 6 else (2 if acc <= 11 else 3)
    Msp = int(acc)
    tau = (np.pi * Msp) / (N * N * R * (R - 0.5))
    return Msp, R * N, float(tau), R


def standardize(x: np.ndarray):
    """z-score (ddof=1) then centre at the midpoint of min/max (model.py:58-72)."""
    mean = x.mean(axis=0)
    std = x.std(axis=0, ddof=1)
    std[std == 0] = 1.0
    xs = (x - mean) / std
    centre = (xs.min(axis=0) + xs.max(axis=0)) / 2
    return xs - centre, mean, std, centre


def raw_grid(x: np.ndarray, N: Sequence[int], xrange=None) -> List[np.ndarray]:
    """linspace(min, max, N) per dimension, no padding (core.py:64-103)."""
    if xrange is None:
        lo, hi = x.min(axis=0), x.max(axis=0)
    else:
        r = np.asarray(xrange, float)
        lo, hi = r[:, 0], r[:, 1]
    return [np.linspace(lo[d], hi[d], int(N[d]), dtype=float) for d in range(x.shape[1])]


def padded_size(N: np.ndarray, flag_power2: bool = True):
    """L, qin, qout (convolution.py:390-409)."""
    N = np.asarray(N, dtype=int)
    L = 2 ** np.ceil(np.log2(2 * N - 1)).astype(int) if flag_power2 else N + 2
    qin0 = (L - N) // 2
    qout0 = np.ceil((L - N) / 2).astype(int)
    return L, qin0, qout0


def poly_powers(dims: int, order: int) -> List[Tuple[int, ...]]:
    """Monomial exponents in the reference's term order (design_matrix_optimized.py:30-82)."""
    if order == 0:
        return [(0,) * dims]
    out = []
    if dims == 1:
        out = [(p,) for p in range(order + 1)]
    elif dims == 2:
        for t in range(order + 1):
            for px in range(t, -1, -1):
                out.append((px, t - px))
    elif dims == 3:
        for t in range(order + 1):
            for px in range(t, -1, -1):
                for py in range(t - px, -1, -1):
                    out.append((px, py, t - px - py))
    else:
        raise ValueError("dims > 3 not supported by the restatement")
    return out


def pair_list(nt: int) -> List[Tuple[int, int]]:
    """S entries (i, j>=i) row-major = column-major lower (convolution.py:488-497,
    design_matrix_optimized.py:289-297)."""
    return [(i, j) for i in range(nt) for j in range(i, nt)]


def diag_index(k: int, nt: int) -> int:
    """Position of S_kk in the pair list (regression.py:243-262)."""
    return k * nt - k * (k - 1) // 2


# ----------------------------------------------------------------------------
# kernel-moment spectra
# ----------------------------------------------------------------------------

def kernel_grids(xs: np.ndarray, h: np.ndarray, N: np.ndarray, order: int):
    """Gaussian kernel x monomials sampled on the centred grid (convolution.py:412-512,
    kernel.py:330-357, design_matrix_optimized.py:216-298).

    Returns (kd_list, ihbad) where kd_list[s] has shape (N..., dh); for order 0 the
    list has one entry.  The Gaussian uses the 1D constant 1/sqrt(2 pi) for every d
    and divides by prod(h).
    """
    dims = xs.shape[1]
    lo, hi = xs.min(axis=0), xs.max(axis=0)
    axes = [np.linspace(lo[d], hi[d], int(N[d])) for d in range(dims)]
    mesh = np.meshgrid(*axes, indexing="ij")
    dh = h.shape[0]
    u = [mesh[d][..., None] / h[None, :, d].reshape((1,) * dims + (dh,)) for d in range(dims)]
    sq = sum(ud * ud for ud in u)
    K = np.exp(-0.5 * sq) / SQRT_2PI / np.prod(h, axis=1).reshape((1,) * dims + (dh,))
    powers = poly_powers(dims, order)
    if order == 0:
        kd = [K]
    else:
        kd = []
        for i, j in pair_list(len(powers)):
            term = K.copy()
            for d in range(dims):
                e = powers[i][d] + powers[j][d]
                if e:
                    term = term * u[d] ** e
            kd.append(term)
    thr = np.finfo(np.float64).eps * 10
    ihbad = np.abs(np.sum(K, axis=tuple(range(dims)))) < thr
    return kd, ihbad


def filter_bandwidths(xs, h, N, order):
    """Drop too-small bandwidths; Silverman fallback when all are bad (convolution.py:542-605)."""
    kd, ihbad = kernel_grids(xs, h, N, order)
    if np.any(ihbad) and not np.all(ihbad):
        warnings.warn(f"{int(np.sum(ihbad))} bandwidth(s) are too small for this grid and will be removed",
                      stacklevel=2)
        h = h[~ihbad]
        kd = [k[..., ~ihbad] for k in kd]
    elif np.all(ihbad):
        n, dims = xs.shape
        hs = (4.0 / (dims + 2) / n) ** (1.0 / (dims + 4))
        warnings.warn(f"All bandwidths too small. Using Silverman's rule: h = {hs:.4f}", stacklevel=2)
        h = np.array([[hs] * dims])
        kd, _ = kernel_grids(xs, h, N, order)
    return kd, h, ihbad


def kernel_spectrum(kd: np.ndarray, L: np.ndarray, N: np.ndarray) -> np.ndarray:
    """Zero-pad with (L-N)//2+1 leading zeros, FFT every spatial axis, no shift
    (convolution.py:747-794)."""
    dims = len(N)
    pads = []
    for d in range(dims):
        before = (int(L[d]) - int(N[d])) // 2 + 1
        pads.append((before, int(L[d]) - int(N[d]) - before))
    pads.append((0, 0))
    out = np.pad(kd, pads).astype(np.complex128)
    for d in range(dims):
        out = sfft.fft(out, axis=d)
    return out


# ----------------------------------------------------------------------------
# NUFFT type 1 (Gaussian gridding)
# ----------------------------------------------------------------------------

def knot_positions(xs: np.ndarray, N: np.ndarray, L: np.ndarray) -> np.ndarray:
    """Map scaled x to xmod in [0, 2 pi) exactly as nufft_transform + scale_knots do
    (convolution.py:169-199, nufftn_type1.py:140-176)."""
    n, dims = xs.shape
    xmin, xmax = xs.min(axis=0), xs.max(axis=0)
    scale = xmax - xmin
    scale[scale == 0] = 1.0
    xn = np.empty_like(xs)
    qin0 = (L - N) // 2
    qin1 = qin0 + N - 1
    for d in range(dims):
        Ld = int(L[d])
        idx = np.linspace(-0.5, 0.5 - 1 / Ld, Ld)
        if N[d] != Ld:
            ks = idx[qin1[d]] - idx[qin0[d]]
            xn[:, d] = (xs[:, d] - xmin[d]) / scale[d] * ks + idx[qin0[d]]
        else:
            ks = idx[-1] - idx[0]
            xn[:, d] = (xs[:, d] - xmin[d]) / scale[d] * ks - 0.5
    xn += 0.5
    np.mod(xn, 1.0, out=xn)
    # nufft_type1_wrapper: x - 0.5, df = 1/L -> N_nufft = L, Fs = N_nufft
    k = xn - 0.5
    Nn = np.round(1.0 / (1.0 / L.astype(float))).astype(int)
    k = k * (Nn / Nn)
    kmin = (k.min(axis=0) + k.max(axis=0)) * 0.5 - 0.5
    k = k + (-0.5 - kmin)
    k = k * (2 * np.pi)
    return np.mod(k, 2 * np.pi)


def spread(xmod: np.ndarray, values: np.ndarray, L: np.ndarray, acc: int, mode: str) -> np.ndarray:
    """Gaussian gridding onto the Mr^d torus (nufftn_type1.py:520-636, spreading_fused.py:67-94).

    values: (n, C) real or complex.  Returns Ftau of shape (Mr..., C) in the
    grid dtype (complex64 for accuracy<5 in mode "ref", complex128 otherwise).
    """
    n, dims = xmod.shape
    C = values.shape[1]
    Msp, Mr, tau, R = zip(*[grid_params(int(L[d]), acc) for d in range(dims)])
    Mr = np.array(Mr)
    hx = 2 * np.pi / Mr
    m = np.round(xmod / hx).astype(np.int64)
    strides = np.ones(dims, dtype=np.int64)
    for d in range(dims - 2, -1, -1):
        strides[d] = strides[d + 1] * Mr[d + 1]
    acc_re = np.zeros((int(np.prod(Mr)), C))
    acc_im = np.zeros((int(np.prod(Mr)), C))
    is_c = np.iscomplexobj(values)
    # per-dimension tap weights exp(-diff^2 / (4 tau)), diff = xmod - hx (m + j)
    taps = [np.arange(-Msp[d], Msp[d] + 1) for d in range(dims)]
    wd, idxd = [], []
    for d in range(dims):
        mj = m[:, d:d + 1] + taps[d][None, :]
        diff = xmod[:, d:d + 1] - hx[d] * mj
        if acc == 0:
            # accuracy 0: every point binned to its nearest cell (nufftn_type1.py:442-516)
            wd.append(np.zeros_like(diff))
        else:
            wd.append(diff * diff / (4 * tau[d]))
        idxd.append(np.mod(mj, Mr[d]))
    # build the full tensor-product neighbourhood
    lin = np.zeros((n, 1), dtype=np.int64)
    arg = np.zeros((n, 1))
    for d in range(dims):
        lin = (lin[:, :, None] + (idxd[d] * strides[d])[:, None, :]).reshape(n, -1)
        arg = (arg[:, :, None] + wd[d][:, None, :]).reshape(n, -1)
    w = np.exp(-arg)
    flat = lin.ravel()
    size = int(np.prod(Mr))
    for c in range(C):
        v = values[:, c]
        acc_re[:, c] = np.bincount(flat, weights=(w * v.real[:, None]).ravel(), minlength=size)
        if is_c:
            acc_im[:, c] = np.bincount(flat, weights=(w * v.imag[:, None]).ravel(), minlength=size)
    grid = acc_re + 1j * acc_im
    # the binning path (acc 0) keeps its complex128 grid (nufftn_type1.py:474-488)
    cdt = np.complex64 if (mode == "ref" and 0 < acc < 5) else np.complex128
    return grid.astype(cdt).reshape(tuple(Mr) + (C,)), (Msp, Mr, tau, R)


def nufft_type1(xmod, values, L, acc, mode):
    """Spread, FFT over axes d-1..0, fftshift, / (n R^d), extract N=L, optional
    deconvolution (nufftn_type1.py:693-770); then ifftshift (convolution.py:235-236)."""
    n, dims = xmod.shape
    F, (Msp, Mr, tau, R) = spread(xmod, values, L, acc, mode)
    for ax in range(dims - 1, -1, -1):
        F = sfft.fft(F, axis=ax)
    F = sfft.fftshift(F, axes=tuple(range(dims)))
    F = F / (n * (R[0] ** dims))
    sl = []
    for d in range(dims):
        q0 = int(np.ceil((Mr[d] - L[d]) / 2))
        sl.append(slice(q0, q0 + int(L[d])))
    F = F[tuple(sl)]
    if acc > 6:
        tau_arr = np.array(tau)
        f = [np.linspace(-int(L[d]) // 2, int(L[d]) - int(L[d]) // 2 - 1, int(L[d])) for d in range(dims)]
        mesh = np.meshgrid(*f, indexing="ij")
        Kn = np.sqrt(np.prod(tau_arr) / np.pi ** dims) * np.exp(-sum(tau_arr[d] * mesh[d] ** 2 for d in range(dims)))
        Kn = np.maximum(Kn, np.finfo(Kn.dtype).tiny)
        F = F * (1.0 / Kn)[..., None]
    for ax in range(dims):
        F = sfft.ifftshift(F, axes=ax)
    return F


def convolve(y_ft: np.ndarray, kdf: np.ndarray, N: np.ndarray, qout0: np.ndarray, real: bool):
    """Spectrum product (L..., dh, C), IFFT per axis, real part, extract qout
    (convolution.py:904-951)."""
    dims = len(N)
    k = kdf.astype(y_ft.dtype) if y_ft.dtype.itemsize < kdf.dtype.itemsize else kdf
    m = k[..., :, None] * y_ft[..., None, :]
    for ax in range(dims):
        m = sfft.ifft(m, axis=ax)
    if real:
        m = np.real(m)
    sl = tuple(slice(int(qout0[d]), int(qout0[d]) + int(N[d])) for d in range(dims))
    return m[sl]


# ----------------------------------------------------------------------------
# solve, DOF, GCV
# ----------------------------------------------------------------------------

def solve(S: List[np.ndarray], T: List[np.ndarray], order: int, reg: float, nt: int, mode: str):
    """m = T / max(S, max(S) reg) for order 0; m = e1' (S + lam I)^-1 T with
    lam = 1e-6 max_{k,nodes,h} |S_kk| for order >= 1 (regression.py:162-277).

    S[s]: (grid..., dh); T[i]: (grid..., dh, C).  The reference evaluates Cramer
    formulas (lwp_formulas.py:133-380); a generic LU solve is the same
    intercept e1' S^-1 T.
    """
    if order == 0:
        s = S[0]
        thr = np.max(s) * reg
        return T[0] / np.maximum(s, thr)[..., None]
    max_diag = 0.0
    for k in range(nt):
        max_diag = max(max_diag, np.max(np.abs(S[diag_index(k, nt)])))
    lam = 1e-6 * max_diag
    if mode == "ref":
        lam = np.float32(lam)
    dt = S[0].dtype
    A = np.empty(S[0].shape + (nt, nt), dtype=dt)
    for s_idx, (i, j) in enumerate(pair_list(nt)):
        v = S[s_idx] + (lam if i == j else 0)
        A[..., i, j] = v
        A[..., j, i] = v
    B = np.stack(T, axis=-2)  # (grid..., dh, nt, C)
    if np.iscomplexobj(B):
        A = A.astype(np.complex64 if dt == np.float32 else np.complex128)
    X = np.linalg.solve(A, B)
    return X[..., 0, :]


def interp_linear(axes, values, pts):
    """RegularGridInterpolator 'linear', fill_value=None (api.py:686-696)."""
    return RegularGridInterpolator(tuple(axes), values, method="linear",
                                   bounds_error=False, fill_value=None)(pts)


def _spline_kwargs(method, mode):
    """RegularGridInterpolator 'cubic' solves for the spline coefficients
    iteratively (scipy make_ndbspl with gcrotmk, relative residual 1e-5), as
    the reference does; mode "f64" solves the same system directly, which is
    the exact not-a-knot tensor spline."""
    if method == "cubic" and mode == "f64":
        from scipy.sparse.linalg import spsolve
        return {"solver": spsolve}
    return {}


def interp_dof(axes, values, pts, method, mode="ref"):
    """_interpolate_with_linear_extrapolation for in-bounds points (regression.py:929-1092)."""
    return RegularGridInterpolator(tuple(axes), values, method=method, bounds_error=False,
                                   fill_value=None, **_spline_kwargs(method, mode))(pts)


def hutchinson(p, phat, n, y_is_complex):
    """Trace estimate and penalties (regression.py:1262-1354). phat: (n, K, dh)."""
    K = p.shape[1]
    pp = np.sum(p * p, axis=0)
    pph = np.sum(p[:, :, None] * phat, axis=0)
    tr = n * (pph / pp[:, None])
    tr = np.where(np.isnan(tr), n, tr)
    tr = np.clip(tr, 0.0, n)
    dof = np.mean(tr, axis=0)
    dh = tr.shape[1]
    if K > 1:
        stderr = np.std(tr, axis=0, ddof=1) / np.sqrt(K)
        pd = 1 - tr / n
        if y_is_complex:
            pd = 2 * pd - 1
        pc = np.clip(pd, 0.0, 1.0)
        pen = 1.0 / np.maximum(pc, 0.01) ** 2
        pen_std = np.std(pen, axis=0, ddof=1)
        pen_mean = np.mean(pen, axis=0)
        if y_is_complex:
            dof = n * (1 - np.mean(pc, axis=0))
    else:
        stderr = np.zeros(dh)
        pen_std = np.zeros(dh)
        pen_mean = None
    return dof, stderr, pen_std, pen_mean, pph, pp


def gcv_score(y, yhat, dof, pen_std, pen_mean, aggregation="first"):
    """GCV = mse * penalty, stderr = mse * penalty_std (regression.py:280-471)."""
    y = np.asarray(y)
    yhat = np.asarray(yhat)
    if y.ndim == 1:
        y = y[:, None]
    if yhat.ndim == 1:
        yhat = yhat[:, None]
    n = y.shape[0]
    rss = np.sum(np.abs(y - yhat) ** 2, axis=0)
    mse = rss / n
    pd = np.clip(1.0 - dof / n, 0.0, 1.0)
    pen = float(pen_mean) if pen_mean is not None else 1.0 / max(pd, 0.01) ** 2
    g = mse * pen
    if aggregation == "first":
        score, m = float(g[0]), float(mse[0])
    elif aggregation == "average":
        score, m = float(np.mean(g)), float(np.mean(mse))
    elif aggregation == "sum":
        score, m = float(np.sum(g)), float(np.sum(mse))
    else:
        raise ValueError(f"Unknown aggregation method: {aggregation}.")
    se = m * pen_std if pen_std > 0 else 0.0
    return score, se, rss


# ----------------------------------------------------------------------------
# the fit
# ----------------------------------------------------------------------------

@dataclass
class Fit:
    yhat: np.ndarray
    grid: tuple
    h: np.ndarray
    gcv: dict
    grid_values: np.ndarray
    m_all: np.ndarray
    s0: np.ndarray
    d: Optional[float]
    dof: float
    dof_stderr: float
    trace: dict = field(default_factory=dict)


class Problem:
    """Everything that depends only on x, h and the options (the S side)."""

    def __init__(self, x_raw, bandwidths, opts, mode):
        self.mode = mode
        self.x_raw = x_raw
        n, dims = x_raw.shape
        self.n, self.dims = n, dims
        self.order = int(opts["order"])
        self.acc = int(opts["accuracy"])
        xs, mean, std, centre = standardize(x_raw)
        self.xs, self.mean, self.std = xs, mean, std
        Nopt = opts.get("N", None)
        if Nopt is None:
            dn = int(np.ceil(n ** (1.0 / dims)))
            Nopt = dn if dims == 1 else [dn] * dims
        N = [int(Nopt)] * dims if np.isscalar(Nopt) else [int(v) for v in Nopt]
        self.N = np.array(N)
        self.axes_raw = raw_grid(x_raw, N, opts.get("xrange", None))
        self.L, self.qin0, self.qout0 = padded_size(self.N, bool(opts.get("flag_power2", True)))
        kd, h, self.ihbad = filter_bandwidths(xs, bandwidths, self.N, self.order)
        self.h = h
        self.dh = h.shape[0]
        self.powers = poly_powers(dims, self.order)
        self.nt = len(self.powers)
        self.kdf = [kernel_spectrum(k, self.L, self.N) for k in kd]
        self.xmod = knot_positions(xs, self.N, self.L)

    def data_spectrum(self, values, acc=None):
        vals = np.asarray(values)
        if vals.ndim == 1:
            vals = vals[:, None]
        vals = vals.astype(np.complex128)
        return nufft_type1(self.xmod, vals, self.L, self.acc if acc is None else acc, self.mode)

    def design(self, add_eps=True, acc=None):
        """S entries (regression.py:31-159) at accuracy acc (default: the
        option's)."""
        ones_ft = self.data_spectrum(np.ones(self.n), acc)
        S = [convolve(ones_ft, k, self.N, self.qout0, True)[..., 0] for k in self.kdf]
        if self.order == 0 and add_eps:
            S[0] = S[0] + np.finfo(float).eps
        return S

    def weighted(self, values, real, acc=None):
        """T entries for the first nt kernels (regression.py:1497-1533)."""
        y_ft = self.data_spectrum(values, acc)
        return [convolve(y_ft, self.kdf[i], self.N, self.qout0, real) for i in range(self.nt)]


def cv_fastlpr(x, y, h=None, options=None, mode="ref", return_trace=False) -> Fit:
    """Restatement of api.cv_fastlpr (api.py:167-832)."""
    x_arr = as_column_matrix(x)
    y_arr = np.asarray(y)
    if y_arr.ndim == 1:
        y_arr = y_arr[:, None]
    elif y_arr.ndim == 2 and y_arr.shape[0] == 1 and y_arr.shape[1] > 1:
        y_arr = y_arr.T
    n, dims = x_arr.shape
    opts = set_defaults(options, order=0, regularization=1e-8, num_dof_sample=10,
                        kernel_type="gaussian", y_type_out="mean", calc_dof=True,
                        gcv_selection="first")
    if opts.get("accuracy") is None:
        opts["accuracy"] = accuracy_for(n)
    order = int(opts["order"])
    if order not in (0, 1, 2):
        raise ValueError("Polynomial order must be 0, 1, or 2.")
    bandwidths = prepare_bandwidths(h, dims, opts, default_hrange(x_arr))
    variance = str(opts.get("y_type_out", "mean")) == "variance"
    y_orig = y_arr.ravel().copy() if variance else None
    P = Problem(x_arr, bandwidths, opts, mode)
    if variance:
        y_arr = np.log(y_arr + 1.0 / len(y_arr))
    reg = float(opts["regularization"])
    real_y = np.isrealobj(y_arr)
    trace = {}

    if not bool(opts.get("calc_dof", True)):
        S = P.design(add_eps=False)
        T = P.weighted(y_arr, real_y)
        m_all = solve(S, T, order, reg, P.nt, mode)        # (grid..., dh, C)
        gv = m_all[..., 0, :]
        gv = gv[..., 0] if gv.shape[-1] == 1 else gv
        yh = interp_linear(P.axes_raw, gv, x_arr)
        d = None
        if variance:
            d = 1.0 / ((1.0 / n) * np.sum(y_orig * np.exp(-yh)))
            yh = np.exp(yh) / d
            gv = np.exp(gv) / d
        gcv = dict(bandwidths=P.h, gcv_m=np.array([np.nan]), gcv_sd=np.array([np.nan]),
                   idmin=0, id1se=0, h1se=P.h[0], hmin=P.h[0],
                   dof=np.array([np.nan]), dof_stderr=np.array([np.nan]))
        return Fit(yh, tuple(P.axes_raw), P.h[0], gcv, gv, m_all, np.real(S[0][..., 0]), d,
                   np.nan, np.nan, trace)

    # With DOF estimation, S and the probe responses use the sample-size
    # accuracy whatever the option says: api.py:569-576 calls
    # compute_design_matrix without an accuracy (regression.py:86-87) and the
    # probe regression fixes it (regression.py:756-758); only the response
    # T (api.py:632-642) takes the option's accuracy.
    acc_n = accuracy_for(n)
    S = P.design(add_eps=True, acc=acc_n)
    # DOF (regression.py:1095-1354)
    K = int(opts.get("num_dof_sample", 10))
    rm = opts.get("random_matrix", None)
    if rm is not None:
        p = np.asarray(rm, dtype=float)
        K = p.shape[1]
    else:
        p = np.random.Generator(np.random.MT19937(opts.get("dof_seed", 42))).standard_normal((n, K))
    if dims == 1 or n >= 5000:
        dflt = "linear"
    else:
        dflt = "cubic"
    dof_method = opts.get("dof_interp_method", dflt)
    Tp = P.weighted(p, True, acc=acc_n)
    mp = solve(S, Tp, order, reg, P.nt, mode)           # (grid..., dh, K)
    phat = interp_dof(P.axes_raw, mp, x_arr, dof_method, mode)  # (n, dh, K)
    phat = np.real(np.transpose(phat, (0, 2, 1)))
    dof, dstd, pen_std, pen_mean, pph, pp = hutchinson(p, phat, n, not real_y)
    trace.update(p_dot_phat=pph, p_dot_p=pp, m_probe=mp)

    T = P.weighted(y_arr, real_y)
    m_all = solve(S, T, order, reg, P.nt, mode)        # (grid..., dh, C)
    gcv_method = opts.get("gcv_interp_method", dflt)
    C = m_all.shape[-1]
    ms = m_all if C > 1 else m_all[..., 0]
    yh_all = RegularGridInterpolator(tuple(P.axes_raw), ms, method=gcv_method,
                                     bounds_error=False, fill_value=None,
                                     **_spline_kwargs(gcv_method, mode))(x_arr)
    agg = str(opts.get("gcv_selection", "first"))
    scores, errs, rss_all = [], [], []
    for ih in range(P.dh):
        yh = yh_all[:, ih] if C == 1 else yh_all[:, ih, :]
        s_, e_, r_ = gcv_score(y_arr, yh, dof[ih], pen_std[ih],
                               None if pen_mean is None else pen_mean[ih], agg)
        scores.append(s_)
        errs.append(e_)
        rss_all.append(r_)
    scores, errs = np.asarray(scores), np.asarray(errs)
    trace.update(rss=np.asarray(rss_all), m_y=m_all)
    dstd_sel = float(opts.get("dstd", 0.0))
    sel = one_se_rule(scores, errs, P.h, True, dstd_sel)
    gv = m_all[..., sel, :]
    gv = gv[..., 0] if C == 1 else gv
    s0 = np.real(S[0][..., sel]) if S[0].ndim > dims else np.real(S[0])
    yh = interp_linear(P.axes_raw, gv, x_arr)
    d = None
    if variance:
        d = 1.0 / ((1.0 / n) * np.sum(y_orig * np.exp(-yh)))
        yh = np.exp(yh) / d
        gv = np.exp(gv) / d
    idmin = int(np.argmin(scores))
    gcv = dict(bandwidths=P.h, gcv_m=scores, gcv_sd=errs, idmin=idmin, id1se=sel,
               h1se=P.h[sel], hmin=P.h[idmin], dof=dof, dof_stderr=dstd)
    return Fit(yh, tuple(P.axes_raw), P.h[sel], gcv, gv, m_all, s0, d,
               float(dof[sel]), float(dstd[sel]), trace)


@dataclass
class KDEFit:
    fhat: np.ndarray
    grid: tuple
    h: np.ndarray
    lcv: dict
    fhat_all: np.ndarray


def cv_fastkde(x, h=None, options=None, mode="ref") -> KDEFit:
    """Restatement of api.cv_fastkde (api.py:835-1184)."""
    x_arr = as_column_matrix(x)
    n, dims = x_arr.shape
    if n < 2:
        raise ValueError(f"Input x must have at least 2 observations. Got {n} observations.")
    opts = set_defaults(options, order=0, calc_dof=False, y_type_out="mean")
    opts = set_defaults(opts, kernel_type="gaussian", verbose=False, dh=20)
    if opts.get("accuracy") is None:
        opts["accuracy"] = accuracy_for(n)
    if opts.get("N") is None:
        an = int(np.ceil(n ** (1.0 / dims)))
        opts["N"] = [an] * dims if dims > 1 else an
    bw = prepare_bandwidths(h, dims, opts, default_hrange(x_arr))
    dh = len(bw)
    xs, mean, std, centre = standardize(x_arr)
    Nv = opts["N"]
    N = np.atleast_1d(np.asarray(Nv if not isinstance(Nv, int) else Nv, dtype=int))
    if N.size == 1:
        N = np.repeat(N, dims)
    # compute_design_matrix with the raw bandwidths in scaled space (api.py:1026-1034)
    L, qin0, qout0 = padded_size(N, bool(opts.get("flag_power2", True)))
    kd, h_f, _ = filter_bandwidths(xs, bw, N, 0)
    kdf = kernel_spectrum(kd[0], L, N)
    xmod = knot_positions(xs, N, L)
    ones_ft = nufft_type1(xmod, np.ones((n, 1), dtype=np.complex128), L, int(opts["accuracy"]), mode)
    s = convolve(ones_ft, kdf, N, qout0, True)[..., 0] + np.finfo(float).eps
    lo, hi = xs.min(axis=0), xs.max(axis=0)
    axes_s = [np.linspace(lo[d], hi[d], N[d]) for d in range(dims)]
    axes_raw = [(axes_s[d] + centre[d]) * std[d] + mean[d] for d in range(dims)]
    dxg = np.prod([a[1] - a[0] for a in axes_raw])
    dens = []
    for ih in range(s.shape[-1]):
        sc = np.real(s[..., ih])
        tot = np.sum(sc) * dxg
        f = sc / tot if tot > 0 else sc
        dens.append(np.maximum(f, 0.0))
    scores, errs = [], []
    K0 = 1.0 / np.sqrt(2.0 * np.pi)
    for ih in range(len(dens)):
        fa = interp_linear(axes_raw, dens[ih], x_arr)
        floo = (n / (n - 1)) * (fa - (K0 / np.prod(bw[ih])) / n)
        lf = np.log(np.maximum(floo, 1e-10))
        scores.append(float(np.sum(lf)))
        errs.append(float(np.std(lf) * np.sqrt(n)))
    scores, errs = np.asarray(scores), np.asarray(errs)
    sel = one_se_rule(scores, errs, bw, False)
    idmax = int(np.argmax(scores))
    lcv = dict(hlist=bw, lcv_m=scores, lcv_sd=errs, idmax=idmax, id1se=sel,
               hmax=bw[idmax], h1se=bw[sel])
    return KDEFit(dens[sel], tuple(axes_raw), bw[sel], lcv, np.stack(dens, axis=-1))


# ----------------------------------------------------------------------------
# prediction and intervals
# ----------------------------------------------------------------------------

NU_TABLE = {  # confidence.py:33-43 (e1' M_K^-1 M_K2 M_K^-1 e1, product Gaussian kernel)
    (1, 0): 0.282094791774, (1, 1): 0.282094791774, (1, 2): 0.476034961118,
    (2, 0): 0.079577471546, (2, 1): 0.079577471546, (2, 2): 0.198943678865,
    (3, 0): 0.022448390266, (3, 1): 0.022448390266, (3, 2): 0.077166341538,
}


def interval(mu_grid, sigma_grid, s0, h, order, alpha=0.05, kind="confidence"):
    """fastlpr_interval (confidence.py:60-216): [upper, lower] on the grid."""
    from scipy.stats import norm
    z = norm.ppf(1 - alpha / 2)
    # NumPy keeps the grids' dtypes (the reference's s0 and mean grid are
    # float32; a Python float times a float32 array stays float32)
    sigma_pos = np.maximum(np.asarray(sigma_grid), 0)
    s0 = np.maximum(np.asarray(s0), 1e-10)
    nu = NU_TABLE[(s0.ndim, int(order))]
    se_sq = sigma_pos * nu / (float(np.prod(np.atleast_1d(h).astype(float))) * s0)
    half = z * np.sqrt(se_sq) if kind == "confidence" else z * np.sqrt(sigma_pos + se_sq)
    mu_grid = np.asarray(mu_grid)
    return np.stack([mu_grid + half, mu_grid - half], axis=-1)
