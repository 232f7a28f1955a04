"""Brute-force float64 definitions that pin the restatement where the
reference's own outputs cannot: 2-D and 3-D order 2, which the reference
cannot run (lwp_formulas.py:436-444 raises FileNotFoundError).  Test
infrastructure only.

The restatement evaluates the reference's moments through FFTs
(nufftn_type1.py:693-736, convolution.py:747-794 and 904-951).  Here the same
numbers come from their definition, with explicit sums and no FFT:

  spread     F(c) = sum_i v_i prod_k exp(-(xmod_ik - hx_k m_k)^2 / (4 tau_k))
             over the (2 Msp + 1)^d cells m = round(xmod_i / hx) + j, taken mod L
             (nufftn_type1.py:565-636)
  moments    M_alpha[node] = (1/n) sum_c F(c) Kpad_alpha[(qout0 + node - c) mod L]
             where Kpad_alpha holds K(u) u^alpha on the centred grid axis
             (u = axis / h, K(u) = exp(-|u|^2 / 2) / (sqrt(2 pi) prod h)) after
             (L - N) // 2 + 1 leading zeros (convolution.py:747-794): the
             circular convolution that the spectrum product and inverse FFT of
             convolution.py:904-951 compute, extracted at qout0
  solve      m = e1' (S + lam I)^-1 T per node by Gauss-Jordan elimination, S
             assembled from M_{p_i + p_j}, lam = 1e-6 max_{k, nodes, h} S_kk
             (regression.py:217, 243-267)

Only the knot positions xmod and the grid axes are taken from the
restatement (both are pinned end to end by the reference's fixtures).
"""
from __future__ import annotations

import itertools
import math
from typing import Dict, List, Sequence, Tuple

import numpy as np


def spread_direct(xmod: np.ndarray, values: np.ndarray, L: Sequence[int], msp: int) -> np.ndarray:
    """Gaussian gridding by explicit loops over points and footprint taps.
    Returns F of shape (L..., C) float64 (R = 1, no oversampling)."""
    n, d = xmod.shape
    C = values.shape[1]
    L = [int(v) for v in L]
    F = np.zeros(tuple(L) + (C,))
    hx = [2 * math.pi / L[k] for k in range(d)]
    tau = [math.pi * msp / (L[k] * L[k] * 0.5) for k in range(d)]
    for i in range(n):
        m0 = [int(np.round(xmod[i, k] / hx[k])) for k in range(d)]
        for off in itertools.product(range(-msp, msp + 1), repeat=d):
            w = 1.0
            idx = []
            for k in range(d):
                m = m0[k] + off[k]
                diff = xmod[i, k] - hx[k] * m
                w *= math.exp(-diff * diff / (4 * tau[k]))
                idx.append(m % L[k])
            F[tuple(idx)] += w * values[i]
    return F


def kernel_moment_pad(axes_scaled: List[np.ndarray], h: np.ndarray, alpha: Tuple[int, ...],
                      L: Sequence[int]) -> np.ndarray:
    """K(u) u^alpha on the centred grid, zero-padded to L with (L - N) // 2 + 1
    leading zeros per axis, from the definition (full d-dimensional kernel)."""
    d = len(axes_scaled)
    mesh = np.meshgrid(*axes_scaled, indexing="ij")
    u = [mesh[k] / h[k] for k in range(d)]
    K = np.exp(-0.5 * sum(uk * uk for uk in u)) / math.sqrt(2 * math.pi) / float(np.prod(h))
    for k in range(d):
        K = K * u[k] ** alpha[k]
    out = np.zeros(tuple(int(v) for v in L))
    sl = []
    for k in range(d):
        N = len(axes_scaled[k])
        b = (int(L[k]) - N) // 2 + 1
        sl.append(slice(b, b + N))
    out[tuple(sl)] = K
    return out


def moments_direct(F: np.ndarray, n: int, axes_scaled: List[np.ndarray], h: np.ndarray,
                   alphas: Sequence[Tuple[int, ...]]) -> Dict[Tuple[int, ...], np.ndarray]:
    """M_alpha[node, c] = (1/n) sum_cells F(cell, c) Kpad_alpha[(qout0 + node - cell) mod L]
    for every output node, by an explicit sum over the cells."""
    d = len(axes_scaled)
    L = F.shape[:d]
    N = [len(a) for a in axes_scaled]
    qout0 = [int(math.ceil((L[k] - N[k]) / 2)) for k in range(d)]
    cells = np.stack(np.meshgrid(*[np.arange(L[k]) for k in range(d)], indexing="ij"), axis=-1)
    Fc = F.reshape(-1, F.shape[-1])
    out = {}
    for a in alphas:
        Kp = kernel_moment_pad(axes_scaled, h, a, L)
        M = np.zeros(tuple(N) + (F.shape[-1],))
        for node in itertools.product(*[range(v) for v in N]):
            idx = tuple((qout0[k] + node[k] - cells[..., k]) % L[k] for k in range(d))
            w = Kp[idx].reshape(-1)
            M[node] = (w @ Fc) / n
        out[tuple(a)] = M
    return out


def gauss_jordan_first(A: np.ndarray, b: np.ndarray) -> np.ndarray:
    """First row of A^-1 b by Gauss-Jordan elimination with partial pivoting."""
    A = np.array(A, dtype=float)
    b = np.array(b, dtype=float).reshape(A.shape[0], -1)
    m = A.shape[0]
    M = np.concatenate([A, b], axis=1)
    for col in range(m):
        piv = col + int(np.argmax(np.abs(M[col:, col])))
        if piv != col:
            M[[col, piv]] = M[[piv, col]]
        M[col] = M[col] / M[col, col]
        for r in range(m):
            if r != col:
                M[r] = M[r] - M[r, col] * M[col]
    return M[0, m:]


def lpr_direct(xmod: np.ndarray, y: np.ndarray, axes_scaled: List[np.ndarray], hlist: np.ndarray,
               powers: Sequence[Tuple[int, ...]], L: Sequence[int], msp: int) -> np.ndarray:
    """Grid values m (N..., dh, C) of the order >= 1 estimator from the definitions
    above: moments by explicit sums, the global lam over all bandwidths, then a
    Gauss-Jordan solve per node."""
    n, d = xmod.shape
    yv = y.reshape(n, -1)
    C = yv.shape[1]
    nt = len(powers)
    F1 = spread_direct(xmod, np.ones((n, 1)), L, msp)
    Fy = spread_direct(xmod, yv, L, msp)
    N = [len(a) for a in axes_scaled]
    S_all, T_all = [], []
    lam_src = 0.0
    for h in hlist:
        salphas = sorted({tuple(pi[k] + pj[k] for k in range(d)) for pi in powers for pj in powers})
        S = moments_direct(F1, n, axes_scaled, h, salphas)
        T = moments_direct(Fy, n, axes_scaled, h, [tuple(p) for p in powers])
        S_all.append(S)
        T_all.append(T)
        for p in powers:
            lam_src = max(lam_src, float(np.max(np.abs(S[tuple(2 * v for v in p)]))))
    lam = 1e-6 * lam_src
    out = np.zeros(tuple(N) + (len(hlist), C))
    for ih in range(len(hlist)):
        S, T = S_all[ih], T_all[ih]
        for node in itertools.product(*[range(v) for v in N]):
            A = np.empty((nt, nt))
            for i in range(nt):
                for j in range(nt):
                    a = tuple(powers[i][k] + powers[j][k] for k in range(d))
                    A[i, j] = S[a][node][0] + (lam if i == j else 0.0)
            B = np.stack([T[tuple(p)][node] for p in powers])
            out[node + (ih,)] = gauss_jordan_first(A, B)
    return out


# ----------------------------------------------------------------------------
# The direct local polynomial fit (no gridding, no FFT)
# ----------------------------------------------------------------------------
#
# The reference's estimator at a grid node z is e1' (S + lam I)^-1 T with
#   S_ij(z) = sum_i K(u_i) u_i^(p_i + p_j),  T_i(z) = sum_i K(u_i) u_i^p_i y_i,
#   u_i = (xs_i - zs) / h,  K(u) = exp(-|u|^2 / 2) / (sqrt(2 pi) prod h),
# xs the standardised points and zs the node on the standardised grid
# (model.py:58-72, core.py:64-103, convolution.py:412-512), computed there
# through Gaussian gridding and FFTs.  The kernel is sampled on the centred
# grid axis (convolution.py:412-512) and zero-padded (convolution.py:747-794),
# so it is cut off where |xs_i - zs| exceeds half the grid's extent along an
# axis; the sums below apply the same cut.  They equal the reference's
# moments where its gridding is exact: accuracy > 6 (the oversampled,
# deconvolved NUFFT, nufftn_type1.py:693-770) with bandwidths of a few grid
# cells (the FFT product convolves the sampled kernel, whose trigonometric
# interpolant is the Gaussian only when the Gaussian is resolved by the grid
# and its cut-off tail is negligible).  At accuracy <= 6 the reference
# spreads without deconvolution (nufftn_type1.py:565-636), which smooths the
# data by the gridding Gaussian (about sqrt(Msp / pi) cells wide) as well.

def standardised(x: np.ndarray) -> np.ndarray:
    """z-score with ddof = 1, then centred on the midpoint of min / max
    (model.py:58-72)."""
    mean = x.mean(axis=0)
    std = x.std(axis=0, ddof=1)
    std = np.where(std == 0, 1.0, std)
    xs = (x - mean) / std
    return xs - (xs.min(axis=0) + xs.max(axis=0)) / 2


def direct_moments(xs: np.ndarray, y: np.ndarray, zs: np.ndarray, h: np.ndarray,
                   powers: Sequence[Tuple[int, ...]], cut=None, chunk: int = 64):
    """S (m, nt, nt) and T (m, nt) at the m standardised nodes zs from the
    point sums, float64; h (d,); cut (d,) or None: points farther than
    cut[k] from the node along axis k get no weight."""
    n, d = xs.shape
    nt = len(powers)
    m = zs.shape[0]
    S = np.zeros((m, nt, nt))
    T = np.zeros((m, nt))
    P = np.array(powers)                                  # (nt, d)
    norm = 1.0 / (math.sqrt(2 * math.pi) * float(np.prod(h)))
    for a in range(0, m, chunk):
        z = zs[a:a + chunk]
        diff = xs[None, :, :] - z[:, None, :]                        # (c, n, d)
        u = diff / h[None, None, :]
        K = np.exp(-0.5 * np.sum(u * u, axis=2)) * norm              # (c, n)
        if cut is not None:
            K = K * np.all(np.abs(diff) <= np.asarray(cut)[None, None, :], axis=2)
        mono = np.prod(u[:, :, None, :] ** P[None, None, :, :], axis=3)   # (c, n, nt)
        Km = K[:, :, None] * mono
        S[a:a + chunk] = np.einsum("cni,cnj->cij", Km, mono)
        T[a:a + chunk] = np.einsum("cni,n->ci", Km, y)
    return S, T


def direct_lpr_nodes(x: np.ndarray, y: np.ndarray, N: Sequence[int], nodes: np.ndarray,
                     h: np.ndarray, powers: Sequence[Tuple[int, ...]], lam: float) -> np.ndarray:
    """The direct local polynomial fit e1' (S + lam I)^-1 T at grid nodes
    (index tuples, (m, d)) of the N-node grid over the data's range,
    bandwidth h (d,), with the kernel cut at half the grid's extent.  lam is
    the ridge (the reference's 1e-6 max |S_kk| over nodes and bandwidths,
    regression.py:217, 243-267, supplied by the caller)."""
    n, d = x.shape
    xs = standardised(x)
    lo, hi = xs.min(axis=0), xs.max(axis=0)
    axes_s = [np.linspace(lo[k], hi[k], int(N[k])) for k in range(d)]
    # the kernel grid's offsets span [-(hi - lo) / 2, (hi - lo) / 2]
    cut = (hi - lo) / 2 * (1 + 1e-12)
    zs = np.stack([axes_s[k][nodes[:, k]] for k in range(d)], axis=1)
    S, T = direct_moments(xs, y, zs, np.asarray(h, float), powers, cut)
    out = np.empty(len(zs))
    for j in range(len(zs)):
        out[j] = gauss_jordan_first(S[j] + lam * np.eye(len(powers)), T[j][:, None])[0]
    return out
