"""Host-side plan of the separable banded convolution, checked on the CPU.

The engine's job tables (slab offsets, Toeplitz base, tap bands, channel
prefixes) are executed here by a NumPy emulation of the flpr_conv semantics
(out[r][i] = sum_s in[r][s] tap[i + base - s] / out[o][i][c] = sum_s tap[...] in[o][s][c])
on a NumPy spread grid built with the same knot mapping, and the resulting
S and T moments are compared with the oracle's FFT convolution in float64.
A plausible indexing mistake (wrong qout/pad offset = one-cell shift, wrong
slab offset, a dropped channel prefix) changes the moments by O(1).

Accuracy > 6 adds the oversampled spread and the deconvolving resample
(flpr_resample's formula g[j] = sum_c F[c] a[(R j - c) mod Mr]) onto the
widened coarse grid (E >= N), emulated here with the engine's fine geometry.
"""
import math

import numpy as np
import pytest
import torch

from fastlpr_amd import engine as E
from oracle import restate as R


def _fine_grid(geo, xmod, vals, acc):
    """Oversampled spread (every footprint index checked against the fine
    grid) and the per-axis deconvolving resample onto the coarse grid."""
    gf = E.fine_geometry(geo, acc)
    g = geo.geom
    n, d = xmod.shape
    Gf = tuple(gf.G[k] for k in range(d))
    fine = np.zeros((vals.shape[1],) + Gf)
    j = np.arange(-acc, acc + 1)
    for i in range(n):
        ws, idx = [], []
        for k in range(d):
            m = int(np.round(xmod[i, k] / gf.hx[k]))
            pm = m - gf.L[k] if m > gf.L[k] // 2 else m
            diff = xmod[i, k] - gf.hx[k] * (m + j)
            ws.append(np.exp(-(diff * diff * gf.inv4tau[k])))
            ix = pm + j + gf.E[k]
            assert ix.min() >= 0 and ix.max() < gf.G[k]
            idx.append(ix)
        if d == 1:
            fine[:, idx[0]] += vals[i][:, None] * ws[0][None]
        else:
            fine[:, idx[0][:, None], idx[1][None, :]] += \
                vals[i][:, None, None] * np.outer(ws[0], ws[1])[None]
    out = fine.astype(complex)
    for k in range(d - 1, -1, -1):
        Mr, Lc = gf.L[k], g.L[k]
        Rr = Mr // Lc
        tau = (math.pi * acc) / (Lc * Lc * Rr * (Rr - 0.5))
        kap = np.arange(-Lc // 2, Lc // 2)
        delta = np.arange(Mr)
        a = (np.sqrt(math.pi / tau) * np.exp(tau * kap * kap.astype(float))[None, :]
             * np.exp(2j * np.pi * np.outer(delta, kap) / Mr)).sum(1) / (Lc * Rr)
        jc = np.arange(g.G[k]) - g.E[k]
        cf = np.arange(gf.G[k]) - gf.E[k]
        A = a[np.mod(Rr * jc[:, None] - cf[None, :], Mr)]
        out = np.moveaxis(np.tensordot(A, np.moveaxis(out, k + 1, 0), axes=(1, 0)), 0, k + 1)
    return out.real


def _emulate(x, y, h, order, N, acc=None, fp64_plan=True):
    n, d = x.shape
    acc = R.accuracy_for(n) if acc is None else acc
    sd = x.std(0, ddof=1)
    st = E.Stats(x.min(0), x.max(0), x.mean(0), np.where(sd == 0, 1, sd))
    if acc > 6:
        geo = E.build_geometry(st, tuple(N), 0, margin=E.fine_margin(tuple(N)))
        assert all(geo.geom.E[k] >= N[k] for k in range(d))
    else:
        geo = E.build_geometry(st, tuple(N), acc)
    g = geo.geom
    h, _, _ = E.filter_bandwidths(geo, h, n)
    dh = h.shape[0]
    C = 2
    plan = E.Plan(d, order, geo.N, geo.L, geo.margin, C, dh, fp64_plan, "cpu")
    xmod = np.zeros((n, d))
    cell = np.zeros((n, d), int)
    for k in range(d):
        t = (x[:, k] - g.mean[k]) / g.std_[k]
        xs = t - g.centre[k]
        xn = (xs - g.xmin_s[k]) / g.scale_s[k] * g.ks[k] + g.idx0[k]
        xn = np.mod(xn + 0.5, 1.0)
        kk = (xn - 0.5) + g.shift[k]
        xmod[:, k] = np.mod(kk * 2 * np.pi, 2 * np.pi)
        cell[:, k] = np.round(xmod[:, k] / g.hx[k]).astype(int)
    grid = np.zeros((C,) + plan.G)
    msp = geo.msp
    vals = np.column_stack([np.ones(n), y])
    j = np.arange(-msp, msp + 1)
    if acc > 6:
        grid = _fine_grid(geo, xmod, vals, acc)
    for i in range(n if acc <= 6 else 0):
        ws, idx = [], []
        for k in range(d):
            m = cell[i, k]
            pm = m - g.L[k] if m > g.L[k] // 2 else m
            diff = xmod[i, k] - g.hx[k] * (m + j)
            ws.append(np.exp(-(diff * diff * g.inv4tau[k])))
            idx.append(pm + j + g.E[k])
        if d == 1:
            grid[:, idx[0]] += vals[i][:, None] * ws[0][None]
        else:
            grid[:, idx[0][:, None], idx[1][None, :]] += vals[i][:, None, None] * np.outer(ws[0], ws[1])[None]
    if acc <= 6:
        grid = _periodize(grid, g, d)
    tap = np.zeros(dh * d * plan.npow * plan.nmax)
    for b in range(dh):
        for k in range(d):
            xg = np.linspace(g.xmin_s[k], g.xmax_s[k], geo.N[k])
            u = xg / h[b, k]
            e = np.exp(-0.5 * u * u)
            if k == 0:
                e = e / math.sqrt(2 * math.pi) / np.prod(h[b]) / n
            for a in range(plan.npow):
                o = ((b * d + k) * plan.npow + a) * plan.nmax
                tap[o:o + geo.N[k]] = e * u ** a
    bands = E.tap_bands(geo, h, True).reshape(-1)
    stage0 = np.zeros(dh * plan.stage_per_h[0])
    _run(plan.pass_tables, [plan.chunk * plan.stage_per_h[k] for k in range(d + 1)],
         grid.reshape(-1), stage0, tap, bands, d, plan.nmax)
    return plan, stage0, h, acc, dict(geo=geo, tap=tap, bands=bands, xmod=xmod, cell=cell)


def _periodize(grid, g, d):
    """flpr_grid_periodize: for G > L on some axis, every cell of the unwrapped
    grid becomes the sum over its residue class modulo L."""
    G = [g.G[k] for k in range(d)]
    L = [g.L[k] for k in range(d)]
    if all(G[k] <= L[k] for k in range(d)):
        return grid
    res = [np.mod(np.arange(G[k]) - g.E[k], L[k]) for k in range(d)]
    out = np.empty_like(grid)
    for c in range(grid.shape[0]):
        per = np.zeros(L)
        if d == 1:
            np.add.at(per, res[0], grid[c])
            out[c] = per[res[0]]
        else:
            np.add.at(per, (res[0][:, None], res[1][None, :]), grid[c])
            out[c] = per[res[0][:, None], res[1][None, :]]
    return out


def _run(tables, inter_sizes, grid, stage0, tap, bands, d, nmax):
    """flpr_conv semantics over a plan's pass tables (a job applies its
    `powers` consecutive tap rows to the same input, row p's output at
    out_off + p * out_stride); every output range is checked against its
    buffer and written at most once per pass."""
    for ci in range(len(tables)):
        inter = {k: np.zeros(inter_sizes[k]) for k in range(1, d)}
        for ps in tables[ci]:
            src = grid if ps.in_stage == d else inter[ps.in_stage]
            dst = stage0 if ps.out_stage == 0 else inter[ps.out_stage]
            written = np.zeros(len(dst), bool)
            assert ps.max_np == max(max(1, int(r[5])) for r in ps.jobs_host)
            for rec in ps.jobs_host:
                ioff, ooff, toff, band, ro, npw, ostride = (int(v) for v in rec[:7])
                lo, hi = bands[2 * band], bands[2 * band + 1]
                span_in = ro * ps.K * (ps.inner if ps.mode == 1 else 1)
                span_out = ro * ps.nout * (ps.inner if ps.mode == 1 else 1)
                assert 0 <= ioff and ioff + span_in <= len(src)
                for p in range(max(npw, 1)):
                    t = np.arange(ps.nout)[:, None] + ps.base - np.arange(ps.K)[None, :]
                    tp = toff + p * nmax
                    A = np.where((t >= lo) & (t <= hi), tap[np.clip(tp + t, 0, len(tap) - 1)],
                                 0.0)
                    op = ooff + p * ostride
                    assert 0 <= op and op + span_out <= len(dst)
                    assert not written[op:op + span_out].any()
                    written[op:op + span_out] = True
                    if ps.mode == 0:
                        In = src[ioff:ioff + ro * ps.K].reshape(ro, ps.K)
                        dst[op:op + ro * ps.nout] = (In @ A.T).reshape(-1)
                    else:
                        In = src[ioff:ioff + ro * ps.K * ps.inner].reshape(ro, ps.K, ps.inner)
                        dst[op:op + ro * ps.nout * ps.inner] = \
                            np.einsum("is,osc->oic", A, In).reshape(-1)


CASES = [(1, o, [40]) for o in (0, 1, 2)] + [(1, 1, [33])] + \
        [(2, o, [12, 10]) for o in (0, 1, 2)] + [(2, 2, [11, 9])] + \
        [(1, 1, [2]), (1, 0, [3]), (1, 2, [4]), (2, 1, [2, 3]), (2, 0, [5, 2])]


CASES_HI = [(1, 1, [40], 7), (1, 2, [33], 12), (2, 1, [12, 10], 7), (2, 2, [11, 9], 9)]


@pytest.mark.parametrize("fp64_plan", [True, False])
@pytest.mark.parametrize("d,order,N,acc", [c + (None,) for c in CASES] + CASES_HI)
def test_banded_plan_reproduces_fft_moments(d, order, N, acc, fp64_plan):
    """Both job tables (the float64 form's one power per job, the float32
    form's multi-power jobs), executed in float64, give the FFT moments."""
    rng = np.random.default_rng(3 + d + order)
    n = 300
    x = rng.uniform(0, 1, (n, d))
    y = np.sin(6 * x[:, 0]) + 0.1 * rng.standard_normal(n)
    h = np.array([[0.05] * d, [0.2] * d, [0.6] * d]) if d == 1 else np.array([[0.1, 0.1], [0.3, 0.5]])
    plan, st0, h, acc, _ = _emulate(x, y, h, order, N, acc, fp64_plan)
    if not fp64_plan and order > 0:
        assert any(ps.max_np > 1 for ps in plan.pass_tables[0])
    P = R.Problem(x, h, {"order": order, "accuracy": acc, "N": N}, "f64")
    S = P.design(add_eps=False)
    T = P.weighted(y, True)
    per_h, nodes, dh = plan.stage_per_h[0], plan.nodes, h.shape[0]
    for p, off in enumerate(plan.s_off):
        em = np.stack([st0[b * per_h + off:b * per_h + off + nodes] for b in range(dh)], -1)
        ref = S[p].reshape(nodes, -1)
        assert np.max(np.abs(em - ref)) <= 1e-12 * np.max(np.abs(ref))
    for i, off in enumerate(plan.t_off):
        em = np.stack([st0[b * per_h + off + nodes:b * per_h + off + 2 * nodes] for b in range(dh)], -1)
        ref = T[i][..., 0].reshape(nodes, -1)
        assert np.max(np.abs(em - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("fp64_plan", [True, False])
@pytest.mark.parametrize("d,order,N", [(1, 1, [40]), (1, 2, [33]), (2, 2, [12, 10]),
                                       (2, 1, [11, 9])])
def test_y_only_passes_write_the_new_channel_moments(d, order, N, fp64_plan):
    """Plan.y_pass_tables (a second fit of the same x): the passes convolve
    only the new channel and write its T moments into channel 1 of the first
    fit's stage-0 slabs, leaving every S moment and the ones channel intact."""
    rng = np.random.default_rng(30 + d + order)
    n = 300
    x = rng.uniform(0, 1, (n, d))
    y = np.sin(6 * x[:, 0]) + 0.1 * rng.standard_normal(n)
    y2 = np.cos(4 * x[:, -1]) + 0.2 * rng.standard_normal(n)
    h = np.array([[0.05] * d, [0.2] * d, [0.6] * d]) if d == 1 else np.array([[0.1, 0.1], [0.3, 0.5]])
    plan, st0, h, acc, aux = _emulate(x, y, h, order, N, None, fp64_plan)
    before = st0.copy()
    geo, xmod, cell = aux["geo"], aux["xmod"], aux["cell"]
    g = geo.geom
    grid2 = np.zeros((1,) + plan.G)
    j = np.arange(-geo.msp, geo.msp + 1)
    for i in range(n):
        ws, idx = [], []
        for k in range(d):
            m = cell[i, k]
            pm = m - g.L[k] if m > g.L[k] // 2 else m
            diff = xmod[i, k] - g.hx[k] * (m + j)
            ws.append(np.exp(-(diff * diff * g.inv4tau[k])))
            idx.append(pm + j + g.E[k])
        w = ws[0] if d == 1 else np.outer(ws[0], ws[1])
        sl = (0, idx[0]) if d == 1 else (0, idx[0][:, None], idx[1][None, :])
        grid2[sl] += y2[i] * w
    yper, tables = plan.y_pass_tables(1)
    _run(tables, [plan.chunk * yper[k] for k in range(d + 1)], grid2.reshape(-1), st0,
         aux["tap"], aux["bands"], d, plan.nmax)
    P = R.Problem(x, h, {"order": order, "accuracy": acc, "N": N}, "f64")
    T2 = P.weighted(y2, True)
    per_h, nodes, dh = plan.stage_per_h[0], plan.nodes, h.shape[0]
    changed = np.zeros(st0.shape, bool)
    for i, off in enumerate(plan.t_off):
        for b in range(dh):
            sl = slice(b * per_h + off + nodes, b * per_h + off + 2 * nodes)
            changed[sl] = True
        em = np.stack([st0[b * per_h + off + nodes:b * per_h + off + 2 * nodes]
                       for b in range(dh)], -1)
        ref = T2[i][..., 0].reshape(nodes, -1)
        assert np.max(np.abs(em - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert np.array_equal(st0[~changed], before[~changed])


def test_plan_stage_layout_counts():
    """2-D order 2 with 12 channels: 15 final moments, 6 of them carry all channels."""
    p = E.Plan(2, 2, (256, 256), (512, 512), 4, 12, 30, False, "cpu")
    s0 = p.stage_slabs[0]
    assert len(s0) == 15
    assert sum(1 for _, cb in s0.values() if cb == 12) == 6
    assert p.stage_per_h[0] == (6 * 12 + 9) * 256 * 256
    assert len(p.s_off) == 21 and len(p.t_off) == 6
    # float32 form: ROW jobs = powers 0-2 of all channels + powers 3-4 of the
    # ones channel; COL jobs per stage-1 slab b = 0..4: 2, 2, 2, 1, 1
    assert [len(t.jobs_host) for t in p.pass_tables[0]] == [2 * 30, 8 * 30]
    assert [t.max_np for t in p.pass_tables[0]] == [3, 3]
    p64 = E.Plan(2, 2, (256, 256), (512, 512), 4, 12, 30, True, "cpu")
    assert [len(t.jobs_host) for t in p64.pass_tables[0]] == [5 * 30, 15 * 30]
    assert [t.max_np for t in p64.pass_tables[0]] == [1, 1]


@pytest.mark.parametrize("fp64", [False, True])
def test_conv_block_schedule_covers_every_tile_once(fp64):
    """conv_schedule (the 2-D order-2 workload: 256 x 256 nodes, 30
    bandwidths, 12 channels): every (job, tile) of both passes appears exactly once, each
    tile decodes inside its job (flpr_conv's tile numbering), idle entries are
    marked -1, and XCD b mod 8 receives a balanced share of the work."""
    from fastlpr_amd import _lib
    from fastlpr_amd import get_hlist
    _lib.load()
    st = E.Stats(np.zeros(2), np.ones(2), np.full(2, 0.5), np.full(2, 0.2887))
    geo = E.build_geometry(st, (256, 256), 4)
    hh = get_hlist(30, [0.005, 0.5])[:, 0]
    h = np.column_stack([hh, hh])
    bands = E.tap_bands(geo, h, np.full(30, fp64)).reshape(-1).astype(np.int32)
    plan = E.Plan(2, 2, geo.N, geo.L, geo.margin, 12, 30, fp64, "cpu")
    for ps in plan.pass_tables[0]:
        table, nb = E.conv_schedule(ps, fp64, bands)
        assert table is not None and (ps.mode == 0 or nb % E.XCDS == 0)
        t = table.numpy().reshape(-1, 2)
        live = t[t[:, 0] >= 0]
        assert np.all(t[t[:, 0] < 0][:, 1] == 0)
        ntt = -(-ps.nout // 64)
        seen = set()
        for j in range(ps.jobs_host.shape[0]):
            ro = int(ps.jobs_host[j, 4])
            ld = E.CONV_LD[fp64]
            if ps.mode == 0:
                ntd = -(-ps.max_ro // ld)
                want = {td + ntd * tt for tt in range(ntt) for td in range(-(-ro // ld))}
            else:
                ntd = -(-ps.inner // ld)
                want = set(range(ntd * ntt * ro))
            got = live[live[:, 0] == j][:, 1]
            assert len(got) == len(set(got.tolist())) == len(want)
            assert set(got.tolist()) == want
            seen.add(j)
        assert seen == set(range(ps.jobs_host.shape[0]))
        if ps.mode == 1:
            per_xcd = [np.sum(t[x::E.XCDS, 0] >= 0) for x in range(E.XCDS)]
            assert max(per_xcd) <= 1.25 * np.mean(per_xcd) + 64
        else:
            assert len(live) == len(t)
