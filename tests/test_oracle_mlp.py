"""The CPU oracle of the learned bracket (oracle/mlp.py), pinned without a GPU:
bfloat16 rounding against torch's conversion and hand-made halfway cases, the
backward pass against central finite differences of the forward pass (exact
arithmetic; the bf16 path is the same formula on rounded operands), a
planted mistake in the backward pass caught by the same check, and the
feature loop against a vectorised restatement."""
import math

import numpy as np
import pytest
import torch

from oracle import mlp as M


def test_bf16_rounding_matches_torch_and_ties_to_even():
    rs = np.random.default_rng(0)
    a = (rs.standard_normal(10000) * 10.0 ** rs.uniform(-20, 20, 10000)).astype(np.float32)
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(M.bf16(a), ref)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties go to the even mantissa (1)
    assert M.bf16(np.float32(1 + 2 ** -8)) == 1.0
    # 1 + 3 * 2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: the even one is 1 + 2^-6
    assert M.bf16(np.float32(1 + 3 * 2 ** -8)) == np.float32(1 + 2 ** -6)
    assert M.bf16(np.float32(0.5)) == 0.5


def _loss(p, X, T):
    O = M.forward(p, X, rounding=False)["O"]
    return 0.5 * float(np.sum((O - T) ** 2))


@pytest.mark.parametrize("planted", [False, True])
def test_backward_matches_finite_differences(planted):
    rs = np.random.default_rng(1)
    p = rs.standard_normal(M.N_PARAMS) * 0.1
    X = rs.standard_normal((7, M.IN))
    T = rs.standard_normal((7, M.OUT))
    acts = M.forward(p, X, rounding=False)
    dO = acts["O"] - T
    if planted:
        dO = dO.copy()
        dO[:, 1] *= -1.0          # a wrong sign in the output gradient
    g = M.backward(p, X, acts, dO, rounding=False)
    idx = rs.choice(M.N_PARAMS, 60, replace=False)
    num = np.empty(len(idx))
    for q, i in enumerate(idx):
        e = np.zeros(M.N_PARAMS)
        e[i] = 1e-6
        num[q] = (_loss(p + e, X, T) - _loss(p - e, X, T)) / 2e-6
    err = np.max(np.abs(num - g[idx])) / np.max(np.abs(num))
    if planted:
        assert err > 1e-2
    else:
        assert err < 1e-6, err


def test_feature_loop_matches_vectorised_restatement():
    rs = np.random.default_rng(2)
    G0, G1 = 23, 31
    F0 = rs.uniform(0, 2, (G0, G1))
    F0[rs.uniform(size=F0.shape) < 0.1] = 0.0
    F1 = F0 * rs.standard_normal((G0, G1))
    floor = 0.05
    f = M.features(F0, F1, 2.0e4, floor, 12345, 0.013)
    for k, (di, dj) in enumerate([(1, 0), (0, 1), (1, 1), (1, -1)]):
        for si, s in enumerate(M.SCALES):
            a, b = di * s, dj * s
            i0, i1 = 0, max(G0 - a, 0)
            j0, j1 = min(max(0, -b), G1), max(min(G1, G1 - b), 0)
            j1 = max(j1, j0)
            A = F0[i0:i1, j0:j1]
            B = F0[i0 + a:i1 + a, j0 + b:j1 + b]
            m1 = F1[i0:i1, j0:j1]
            m2 = F1[i0 + a:i1 + a, j0 + b:j1 + b]
            ok = (A > floor) & (B > floor)
            with np.errstate(divide="ignore", invalid="ignore"):
                d = np.where(ok, m2 / B - m1 / A, 0.0)
                w = np.where(ok, A * B / (A + B), 0.0)
            want = math.log((np.sum(w * d * d) / np.sum(w) if np.sum(w) > 0 else 0) + 1e-30)
            assert abs(f[k * len(M.SCALES) + si] - want) < 1e-9 * max(1.0, abs(want))
    assert f[56] == pytest.approx(math.log10(12345))
    mask = F0 > floor
    lvar = 2.0e4 / 12345 - np.sum(F1[mask] ** 2 / F0[mask]) / F0.sum()
    assert f[58] == pytest.approx(math.log(lvar))
    assert f[63] == pytest.approx(math.log10(0.013))
