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
