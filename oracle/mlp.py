"""CPU oracle of the learned bandwidth bracket (test infrastructure only).

Restates, in NumPy, what the HIP kernels of flpr_mlp.hip compute, so that
tests can check the kernels against it:

* ``forward`` / ``backward``: the MLP 64 -> 128 -> 128 -> 2 with exact GELU
  (0.5 z (1 + erf(z / sqrt 2))), every GEMM operand rounded to bfloat16
  (round to nearest even) and the products summed in float64 (the kernels
  sum in float32; tests allow for that);
* ``features``: the 64 grid-summary features of a spread grid with the
  channels (sum w, sum w y) and the sum of y^2 over the points.

There is no reference counterpart: the reference always sweeps every
bandwidth (api.py:664-740).  The oracle's gradients are pinned by central
finite differences (tests/test_oracle_mlp.py), its bf16 rounding by the
bit patterns of exactly representable values and of halfway cases, and the
features by a direct loop over a small grid.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

IN, HID, OUT = 64, 128, 2
SHAPES = [("W1", (HID, IN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
          ("W3", (OUT, HID)), ("b3", (OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)
SCALES = [1, 2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32, 48]


def unpack(p: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = p[o:o + k].reshape(shape)
        o += k
    return out


def bf16(a) -> np.ndarray:
    """float32 -> bfloat16 (round to nearest, ties to even) -> float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), a, out)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2 * math.pi)


def mm(a, b, rounding=True):
    """bf16-rounded operands (rounding=True, the kernels' arithmetic) or the
    exact operands, float64 sums."""
    if not rounding:
        return np.asarray(a, np.float64) @ np.asarray(b, np.float64)
    return bf16(a).astype(np.float64) @ bf16(b).astype(np.float64)


def forward(p: np.ndarray, X: np.ndarray, rounding=True):
    w = unpack(p)
    Z1 = mm(X, w["W1"].T, rounding) + w["b1"]
    H1 = gelu(Z1)
    Z2 = mm(H1, w["W2"].T, rounding) + w["b2"]
    H2 = gelu(Z2)
    O = mm(H2, w["W3"].T, rounding) + w["b3"]
    return dict(Z1=Z1, H1=H1, Z2=Z2, H2=H2, O=O)


def backward(p: np.ndarray, X: np.ndarray, acts: dict, dO: np.ndarray,
             rounding=True) -> np.ndarray:
    """Gradient of sum(O * dO) with respect to the packed parameters."""
    w = unpack(p)
    g = {}
    g["W3"] = mm(dO.T, acts["H2"], rounding)
    g["b3"] = dO.sum(axis=0)
    dZ2 = mm(dO, w["W3"], rounding) * gelu_grad(acts["Z2"])
    g["W2"] = mm(dZ2.T, acts["H1"], rounding)
    g["b2"] = dZ2.sum(axis=0)
    dZ1 = mm(dZ2, w["W2"], rounding) * gelu_grad(acts["Z1"])
    g["W1"] = mm(dZ1.T, X, rounding)
    g["b1"] = dZ1.sum(axis=0)
    return np.concatenate([g[name].reshape(-1) for name, _ in SHAPES])


def loss(O, T, w0, w1):
    w = np.array([w0, w1])
    d = O - T
    return float(np.sum(w * d * d) / O.shape[0]), 2.0 * w * d / O.shape[0]


def features(F0: np.ndarray, F1: np.ndarray, ysq: float, floor: float, n: int,
             dx: float) -> np.ndarray:
    """The 64 features of one spread grid (channels sum w and sum w y, G0 x G1
    each) with ysq = sum_i y_i^2 over the n points."""
    G0, G1 = F0.shape
    f = np.zeros(64)
    for di_dir in range(4):
        for si, s in enumerate(SCALES):
            di = 0 if di_dir == 1 else s
            dj = 0 if di_dir == 0 else (-s if di_dir == 3 else s)
            num = den = 0.0
            for i in range(G0):
                i2 = i + di
                if i2 < 0 or i2 >= G0:
                    continue
                for j in range(G1):
                    j2 = j + dj
                    if j2 < 0 or j2 >= G1:
                        continue
                    a, b = F0[i, j], F0[i2, j2]
                    if not (a > floor and b > floor):
                        continue
                    d = F1[i2, j2] / b - F1[i, j] / a
                    wgt = a * b / (a + b)
                    num += wgt * d * d
                    den += wgt
            f[di_dir * len(SCALES) + si] = math.log((num / den if den > 0 else 0.0) + 1e-30)
    s0, s1 = F0.sum(), F1.sum()
    mask = F0 > floor
    sm2 = np.sum(F1[mask] ** 2 / F0[mask])
    cells = G0 * G1
    mean = s1 / s0
    ey2 = ysq / n
    var = max(ey2 - mean * mean, 1e-30)
    lvar = max(ey2 - sm2 / s0, 1e-30)
    m0 = s0 / cells
    cv2 = max(np.sum(F0 * F0) / cells / (m0 * m0) - 1.0, 1e-30)
    f[56] = math.log10(n)
    f[57] = math.log(var)
    f[58] = math.log(lvar)
    f[59] = math.log(max(m0, 1e-30))
    f[60] = math.log(cv2)
    f[61] = np.sum(~mask) / cells
    f[62] = math.log(var / lvar)
    f[63] = math.log10(dx)
    return f
