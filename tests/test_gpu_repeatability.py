"""Run-to-run spread of repeated fits on the same inputs.

The reference is a deterministic CPU program.  On the GPU the order of some
float64 additions is not fixed: points of one spread cell are ranked by
atomic counters (their order inside the cell varies), and the spread and
adjoint windows add work items into LDS with atomics.  So two fits of the
same data may differ in the last bits.  The first test bounds that spread
and checks that it never changes the selected bandwidth: float64 form within
1e-12 of itself, float32 form within 1e-6 (a float64 difference in a grid
value can round to a neighbouring float32).  The cross-block sums of the
probe dot products and residual quadratic forms are added in block order, so
on the same inputs they repeat bit for bit (second test)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-6)])
def test_repeated_fits_agree(prec, tol):
    import fastlpr_amd as fl
    rs = np.random.default_rng(11)
    n = 200_000
    x = rs.uniform(0, 1, (n, 2))
    y = np.sin(2 * np.pi * x[:, 0]) * np.cos(2 * np.pi * x[:, 1]) \
        + (0.05 + 0.2 * x[:, 0]) * rs.standard_normal(n)
    hh = fl.get_hlist(30, [0.005, 0.5])[:, 0]
    h = np.column_stack([hh, hh])
    opt = {"order": 2, "N": [128, 128], "precision": prec}
    runs = [fl.cv_fastlpr_pair(x, y, h, opt) for _ in range(3)]
    worst = 0.0
    for other in runs[1:]:
        for a, b in zip(other, runs[0]):
            assert a.gcv_yhat["id1se"] == b.gcv_yhat["id1se"]
            e = max(_rel(a.yhat, b.yhat), _rel(a.gcv_yhat["gcv_m"], b.gcv_yhat["gcv_m"]),
                    _rel(a.gcv_yhat["dof"], b.gcv_yhat["dof"]))
            assert e <= tol, e
            worst = max(worst, e)
    print(f"repeat spread {prec}: {worst:.2e}")


def _fit_state(prec):
    """One 2-D order-2 fit's device state (run_grid's debug buffers)."""
    import torch
    from fastlpr_amd import engine, get_hlist
    rs = np.random.default_rng(23)
    n = 60_000
    x = torch.as_tensor(rs.uniform(0, 1, (n, 2)), device="cuda")
    y = torch.sin(6.0 * x[:, :1]) + 0.1 * torch.as_tensor(rs.standard_normal((n, 1)),
                                                           device="cuda")
    pr = torch.as_tensor(rs.standard_normal((n, 10)), device="cuda")
    hh = get_hlist(8, [0.01, 0.3])[:, 0]
    dbg = {}
    engine.run_grid(x, y.contiguous(), 1, pr, 10, np.column_stack([hh, hh]), 2, (64, 64), 4,
                    prec == "fp64", 1e-8, want_rss=True, debug=dbg)
    torch.cuda.synchronize()
    return dbg


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_ordered_dot_and_quadform_sums_repeat_bitwise(prec):
    """flpr_solve's probe dots and flpr_quadform's residual sums add their
    block partials in block order (no floating-point atomics): two calls on
    the same device inputs give bit-identical sums, equal to the float64
    sums of the products they reduce; a workspace too small for one block
    per group is refused (-3) before any launch."""
    import torch
    from fastlpr_amd import _lib, engine
    from fastlpr_amd.engine import i64_array, ptr
    dbg = _fit_state(prec)
    plan, h = dbg["plan"], dbg["h"]
    dh, nodes = h.shape[0], plan.nodes
    nval, ny = 11, 1
    per_h0 = plan.stage_per_h[0]
    s = engine.stream_handle()
    nsum = engine.sum_part_len(dh * nval)

    def solve_all():
        dot = torch.zeros((dh * nval,), dtype=torch.float64, device="cuda")
        m_y = torch.empty((dh * ny * nodes,), dtype=torch.float64, device="cuda")
        sump = torch.zeros((nsum,), dtype=torch.float64, device="cuda")
        for (a, b, gf, pl, stage0) in dbg["stage0s"]:
            _lib.call("flpr_solve", int(gf), ptr(stage0), per_h0, b - a, nodes, 2, pl.nt, 0,
                      i64_array(pl.s_off), i64_array(pl.t_off), ptr(dbg["key"]),
                      dbg["lam_scale"], dbg["eps_add"], ptr(dbg["shift"]), nval, ny,
                      m_y.data_ptr() + a * ny * nodes * 8, ptr(dbg["q"]),
                      dot.data_ptr() + a * nval * 8, ptr(sump), nsum, s)
        torch.cuda.synchronize()
        assert int(torch.count_nonzero(sump)) == 0   # the workspace is left zero
        return dot, m_y

    def quad_of(m_y, length=None):
        quad = torch.zeros((dh * ny,), dtype=torch.float64, device="cuda")
        L = nsum if length is None else length
        sump = torch.zeros((max(L, 1),), dtype=torch.float64, device="cuda")
        _lib.call("flpr_quadform", _lib.C.byref(dbg["geo"].geom), ptr(m_y), ny, dh,
                  ptr(dbg["gram"]), ptr(quad), ptr(sump), L, s)
        torch.cuda.synchronize()
        return quad

    d1, m1 = solve_all()
    d2, m2 = solve_all()
    assert torch.equal(m1, m2)
    assert torch.equal(d1, d2)
    # the dots are sum_node q[c][node] m_c(h, node); m_c for the probes is the
    # solve's own output when every channel is written
    nval_out = nval
    m_all = torch.empty((dh * nval_out * nodes,), dtype=torch.float64, device="cuda")
    dot_all = torch.zeros((dh * nval,), dtype=torch.float64, device="cuda")
    sump = torch.zeros((nsum,), dtype=torch.float64, device="cuda")
    for (a, b, gf, pl, stage0) in dbg["stage0s"]:
        _lib.call("flpr_solve", int(gf), ptr(stage0), per_h0, b - a, nodes, 2, pl.nt, 0,
                  i64_array(pl.s_off), i64_array(pl.t_off), ptr(dbg["key"]), dbg["lam_scale"],
                  dbg["eps_add"], ptr(dbg["shift"]), nval, nval_out,
                  m_all.data_ptr() + a * nval_out * nodes * 8, ptr(dbg["q"]),
                  dot_all.data_ptr() + a * nval * 8, ptr(sump), nsum, s)
    torch.cuda.synchronize()
    assert torch.equal(dot_all, d1)
    ref = (m_all.view(dh, nval, nodes) * dbg["q"].view(1, nval, nodes)).sum(dim=2).reshape(-1)
    assert _rel(d1.cpu().numpy(), ref.cpu().numpy()) < 1e-12
    # residual quadratic forms: repeatable bit for bit, and equal to the Gram
    # form sum_cell sum_{a<=b} (2 - [a == b]) W_ab m_a m_b
    q1, q2 = quad_of(m1), quad_of(m1)
    assert torch.equal(q1, q2)
    N0, N1 = 64, 64
    mm = m1.view(dh, N0, N1)
    corner = [mm[:, :-1, :-1], mm[:, :-1, 1:], mm[:, 1:, :-1], mm[:, 1:, 1:]]
    W = dbg["gram"].view(10, N0 - 1, N1 - 1)
    ref_q = torch.zeros((dh,), dtype=torch.float64, device="cuda")
    e = 0
    for ia in range(4):
        for ib in range(ia, 4):
            f = 1.0 if ia == ib else 2.0
            ref_q += f * (W[e][None] * corner[ia] * corner[ib]).sum(dim=(1, 2))
            e += 1
    assert _rel(q1.cpu().numpy(), ref_q.cpu().numpy()) < 1e-12
    # a smaller workspace uses fewer blocks: the same sums to rounding
    q3 = quad_of(m1, engine.sum_part_len(dh * ny, 3))
    assert _rel(q3.cpu().numpy(), q1.cpu().numpy()) < 1e-13
    with pytest.raises(_lib.FlprError, match="status -3"):
        quad_of(m1, dh * ny)
