"""Many independent cv_fastlpr problems in shared launches.

The reference fits one problem per ``cv_fastlpr`` call (api.py:167-832).
``cv_fastlpr_batch(xs, ys, h, options)`` returns what
``[cv_fastlpr(x, y, h, options) for x, y in zip(xs, ys)]`` returns, but
runs every stage once per *class* of problems instead of once per problem:
problems whose grid structure, order, response columns, precision and
(filtered) bandwidths agree share one launch per stage, with the problem on
``blockIdx.z`` and its geometry and point range in a device table
(``FlprBatchItem``; the ``flpr_batch_*`` entry points).  Point arrays of a
class hold the problems' rows one after the other; the counting sort runs
once over all of them with per-problem key ranges, so every problem's sorted
points stay in its own row range.  The convolution passes of all problems of
a class go into one job table per pass.

Two phases:

* ``BatchPlan(xs, h, options)`` -- the x side, once per set of designs: the
  statistics of every problem (one launch per dimension count, one host
  read), its geometry and bandwidth tables (host), the grouping into
  classes, and every device table (geometry items, tap bands, convolution
  job and block tables).
* ``BatchPlan.fit(ys)`` -- the response side: the sort, pack, spread,
  centring, taps, convolutions, regularisation key, adjoint sums, solves,
  quadratic forms, selection and the fitted values, queued without a host
  synchronisation, then one read of every problem's selection results.
  ``BatchPlan.capture()`` records that sequence into a HIP graph (every
  buffer is owned by the plan and cleared by kernels, not memsets), after
  which ``fit`` copies the new responses into the plan's input buffer and
  replays the graph.

Cubic GCV / DOF interpolation (2-D / 3-D problems under 5000 points by
default) runs batched too: every channel's grid solutions, the spline
coefficients (flpr_batch_spline_coef) and the sums at the points
(flpr_batch_point_sums).  Problems the batched path does not cover
(``calc_dof=False``, accuracy above 6 or different accuracies for y and S,
``xrange``, a caller ``random_matrix``, grids narrower than the spread
footprint, bandwidth chunking) are fitted by ``cv_fastlpr`` itself, one
call each, in the same ``fit``.
"""
from __future__ import annotations

import math
import warnings
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _lib, api, engine, tune
from ._lib import FlprBatchItem, call, i64_array
from .structures import RegressionOutput

__all__ = ["BatchPlan", "cv_fastlpr_batch"]

# device memory for the moment slabs of one class launch set
CLASS_BUDGET = 48 << 30
MAX_JOBS = 65535


def cv_fastlpr_batch(xs: Sequence, ys: Sequence, h=None, options=None) -> List[RegressionOutput]:
    """``[cv_fastlpr(x, y, h_i, options_i) ...]`` for independent problems,
    computed in shared launches.  h and options are one value for every
    problem or a list with one per problem."""
    plan = BatchPlan(xs, h, options)
    return plan.fit(ys)


def _per_problem(v, P):
    if isinstance(v, (list, tuple)) and len(v) == P and (
            v == [] or isinstance(v[0], (dict, type(None))) or np.ndim(v[0]) >= 1):
        return list(v)
    return [v] * P


@dataclass
class _Problem:
    idx: int
    x: torch.Tensor            # (n, d) float64 on the device
    n: int
    d: int
    opts: dict
    h_given: Optional[np.ndarray]
    order: int
    N: tuple
    fp64: bool
    as_numpy: bool
    x_host: Optional[np.ndarray]
    stats: Optional[engine.Stats] = None
    geo: Optional[engine.Geometry] = None
    hf: Optional[np.ndarray] = None
    warn: Optional[str] = None
    groups: Optional[list] = None
    band_np: Optional[np.ndarray] = None
    batched: bool = True
    acc_n: int = 4


@dataclass
class _Class:
    """Problems that share one launch per stage."""
    probs: List[_Problem]
    d: int
    order: int
    N: tuple
    ny: int
    cplx: bool
    dy: int
    K: int
    variance: bool
    fp64: bool
    agg: int
    dstd: float
    reg: float
    hf: np.ndarray
    groups: list
    geo0: engine.Geometry
    cubic: tuple = (False, False)   # (GCV, DOF) interpolation is cubic
    bufs: Dict[str, torch.Tensor] = field(default_factory=dict)
    tabs: Dict[str, object] = field(default_factory=dict)


class BatchPlan:
    """The x side of a batch of cv_fastlpr problems (module docstring)."""

    def __init__(self, xs: Sequence, h=None, options=None, tune: bool = False):
        """tune: measure the launch shapes of every class on first use
        (fastlpr_amd.tune); choices already made, in this process or in the
        file FASTLPR_AMD_TUNE_FILE names, are applied either way."""
        self.tune = bool(tune)
        P = len(xs)
        if P == 0:
            raise ValueError("cv_fastlpr_batch needs at least one problem")
        hs = _per_problem(h, P)
        opts_l = _per_problem(options, P)
        dev = None
        for x in xs:
            if isinstance(x, torch.Tensor) and x.is_cuda:
                dev = x.device
                break
        self.dev = dev if dev is not None else api._device()
        self.P = P
        self.probs: List[_Problem] = []
        for i, (x, hh, oo) in enumerate(zip(xs, hs, opts_l)):
            self.probs.append(self._problem(i, x, hh, oo))
        # class sets per (response kind, problems of that kind)
        self._kindsets: Dict[tuple, List[_Class]] = {}
        self._graph = None
        self._stats()

    # ------------------------------------------------------------------ set-up
    def _problem(self, i, x, h, options) -> _Problem:
        as_numpy = not (isinstance(x, torch.Tensor) and x.is_cuda)
        x_arr = api._as_column_matrix(x)
        if x_arr.shape[0] == 0:
            raise ValueError("Input x must have at least one sample.")
        x_host = None
        if not isinstance(x_arr, torch.Tensor):
            api._check_finite(x_arr, "Input x", "cv_fastlpr")
            x_host = x_arr
        opts = api.set_defaults(options, order=0, regularization=1e-8, num_dof_sample=10,
                                kernel_type="gaussian", y_type_out="mean", calc_dof=True,
                                gcv_selection="first")
        n, dims = x_arr.shape
        if opts.get("accuracy") is None:
            opts["accuracy"] = max(6 - int(np.ceil(np.log10(n))), 4)
        order = int(opts["order"])
        if order not in (0, 1, 2):
            raise ValueError("Polynomial order must be 0, 1, or 2.")
        if dims > 3:
            raise NotImplementedError("the GPU engine supports 1-3 dimensions")
        if str(opts.get("kernel_type", "gaussian")).lower() not in ("gauss", "gaussian", "normal"):
            raise NotImplementedError("the GPU engine implements the Gaussian kernel")
        api._check_accuracy(int(opts["accuracy"]), dims)
        h_given = api._prepare_bandwidths(h, dims, opts, None) if h is not None else None
        N = api._grid_shape(opts, n, dims)
        xd = (x_arr.to(self.dev) if isinstance(x_arr, torch.Tensor)
              else torch.as_tensor(x_arr, device=self.dev)).to(torch.float64).contiguous()
        fp64 = str(opts.get("precision", "fp32")).lower() == "fp64"
        return _Problem(i, xd, n, dims, opts, h_given, order, N, fp64, as_numpy, x_host)

    def _stats(self):
        """x statistics of every problem: one batched launch per dimension
        count, one host read for all of them."""
        by_d: Dict[int, List[_Problem]] = {}
        for p in self.probs:
            by_d.setdefault(p.d, []).append(p)
        outs = []
        for d, ps in by_d.items():
            items, rows = _items(ps, [None] * len(ps))
            x_all = torch.cat([p.x for p in ps], dim=0)
            per = 5 * d + 2
            part = torch.empty((len(ps) * 512 * per,), dtype=torch.float64, device=self.dev)
            out = torch.empty((len(ps) * per,), dtype=torch.float64, device=self.dev)
            call("flpr_batch_xstats", ptr(items), len(ps), ptr(x_all), d, None, 0, ptr(part),
                 ptr(out), engine.stream_handle())
            outs.append((ps, per, out))
        for ps, per, out in outs:
            o = out.cpu().numpy().reshape(len(ps), per)
            for p, row in zip(ps, o):
                d = p.d
                w = row[:5 * d].reshape(d, 5)
                std = np.sqrt(w[:, 3] / np.maximum(w[:, 4] - 1, 1))
                std = np.where(std == 0, 1.0, std)
                p.stats = engine.Stats(w[:, 0].copy(), w[:, 1].copy(), w[:, 2].copy(), std,
                                       int(row[5 * d]), 0)
                if p.stats.bad_x:
                    raise ValueError("Input x contains NaN or Inf values. Please remove or "
                                     "handle them before calling cv_fastlpr.")
                api._check_axes(p.stats, p.opts.get("xrange"))

    def _setup(self, probs: List[_Problem], ny: int, cplx: bool) -> List[_Class]:
        """Geometry, bandwidth tables and classes of the problems probs for
        responses of ny real columns (complex ones as (re, im) pairs)."""
        keyed: Dict[tuple, List[_Problem]] = {}
        for p in probs:
            o = p.opts
            K = int(o.get("num_dof_sample", 10))
            dims, n = p.d, p.n
            lin_default = "linear" if (dims == 1 or n >= 5000) else "cubic"
            gcv_m = str(o.get("gcv_interp_method", lin_default))
            dof_m = str(o.get("dof_interp_method", lin_default))
            acc_n = max(6 - int(np.ceil(np.log10(n))), 4)
            acc = int(o["accuracy"])
            variance = str(o.get("y_type_out", "mean")) == "variance"
            methods_ok = gcv_m in ("linear", "cubic") and dof_m in ("linear", "cubic")
            cubic = (gcv_m == "cubic", dof_m == "cubic")
            if any(cubic) and min(p.N) < 4:
                methods_ok = False    # the single call raises cubic's error
            p.batched = (bool(o.get("calc_dof", True)) and methods_ok
                         and str(o.get("precision", "fp32")).lower() != "mixed"
                         and str(o.get("bandwidth_bracket", "full")).lower() == "full"
                         and acc == acc_n and acc_n <= 6 and o.get("xrange") is None
                         and o.get("random_matrix") is None and K >= 1 and ny + K <= 32
                         and bool(o.get("flag_power2", True))
                         and not (variance and ny != 1))
            if not p.batched:
                continue
            bw = p.h_given if p.h_given is not None else api._prepare_bandwidths(
                None, dims, o, np.column_stack((0.1 * p.stats.std, 1.0 * p.stats.std)))
            p.acc_n = acc_n
            geo, gkey = engine.geometry_for(p.stats, p.N, acc_n)
            if any(geo.geom.G[k] > geo.geom.L[k] for k in range(dims)):
                p.batched = False   # the torus fold of small grids: one call
                continue
            hf, warn, groups, band_np = engine.bandwidth_tables(geo, gkey, bw, n, p.order, p.fp64)
            p.geo, p.hf, p.warn, p.groups, p.band_np = geo, hf, warn, groups, band_np
            agg = engine.AGGREGATIONS.get(str(o.get("gcv_selection", "first")))
            if agg is None:
                raise ValueError(f"Unknown aggregation method: {o.get('gcv_selection')}. "
                                 "Use 'first', 'average', or 'sum'.")
            key = (dims, p.order, tuple(p.N), tuple(geo.L), geo.margin, K, p.fp64, variance,
                   agg, float(o.get("dstd", 0.0)), float(o["regularization"]), cubic,
                   int(o.get("dof_seed", 42)), hf.tobytes(), tuple(groups))
            keyed.setdefault(key, []).append(p)
        classes: List[_Class] = []
        for key, ps in keyed.items():
            dims, order, N, L, margin, K, fp64, variance, agg, dstd, reg, cubic = key[:12]
            plan_ok = True
            for (a, b, gf) in ps[0].groups:
                pl = engine.get_plan(dims, order, N, L, margin, 1 + ny + K, b - a, gf, self.dev)
                plan_ok = plan_ok and pl.chunk == b - a
            if not plan_ok:
                for p in ps:
                    p.batched = False
                continue
            for sub in self._split(ps, ny, K):
                c = _Class(sub, dims, order, N, ny, cplx, ny // 2 if cplx else ny, K, variance,
                           fp64, agg, dstd, reg, sub[0].hf, sub[0].groups, sub[0].geo, cubic)
                self._build_class(c)
                choice = (tune.tune_class(self, c, self._retile) if self.tune
                          else tune.lookup(c))
                if choice is not None and choice != tune.DEFAULT[c.d]:
                    self._retile(c, *choice)
                c.tabs["launch_shape"] = choice or tune.DEFAULT[c.d]
                classes.append(c)
        return classes

    def _retile(self, c: "_Class", tile: int, item_points: int) -> None:
        """Give class c's problems geometries with another spread tile edge
        and points per work item (the autotuner's knobs)."""
        for p in c.probs:
            p.geo, _ = engine.geometry_for(p.stats, p.N, p.acc_n, tile=tile,
                                           item_points=item_points)
        c.geo0 = c.probs[0].geo
        P = len(c.probs)
        items = (FlprBatchItem * P)()
        for j, p in enumerate(c.probs):
            items[j].g = p.geo.geom
            items[j].n = p.n
            items[j].pt = int(c.tabs["pts"][j])
        c.bufs["items"] = torch.as_tensor(np.frombuffer(items, dtype=np.uint8).copy(),
                                          device=self.dev)

    def _split(self, ps: List[_Problem], ny: int, K: int):
        """Cut a class into launch sets that respect the moment-slab budget and
        the job limit of one convolution launch."""
        p0 = ps[0]
        C = 1 + ny + K
        per_prob = 0
        jobs = 0
        for (a, b, gf) in p0.groups:
            pl = engine.get_plan(p0.d, p0.order, p0.N, p0.geo.L, p0.geo.margin, C, b - a, gf,
                                 self.dev)
            esz = 8 if gf else 4
            per_prob += esz * (b - a) * sum(pl.stage_per_h)
            jobs = max(jobs, max(len(t.jobs_host) for t in pl.pass_tables[0]))
        cap = max(1, min(CLASS_BUDGET // max(per_prob, 1), MAX_JOBS // max(jobs, 1)))
        return [ps[i:i + cap] for i in range(0, len(ps), cap)]

    def _build_class(self, c: _Class):
        """Device tables of one class: geometry items, tap bands, job and block
        tables of every convolution pass."""
        dev = self.dev
        d, P = c.d, len(c.probs)
        g0 = c.geo0.geom
        rows = [p.n for p in c.probs]
        pts = np.concatenate([[0], np.cumsum(rows)])
        c.tabs["rows"] = int(pts[-1])
        c.tabs["pts"] = pts
        c.tabs["nmax"] = int(max(rows))
        items = (FlprBatchItem * P)()
        for j, p in enumerate(c.probs):
            items[j].g = p.geo.geom
            items[j].n = p.n
            items[j].pt = int(pts[j])
        it_host = np.frombuffer(items, dtype=np.uint8).copy()
        c.bufs["items"] = torch.as_tensor(it_host, device=dev)
        c.tabs["x"] = torch.cat([p.x for p in c.probs], dim=0).contiguous()
        C = 1 + c.ny + c.K
        cells = int(np.prod([g0.G[k] for k in range(d)]))
        c.tabs["cells"] = cells
        c.tabs["C"] = C
        pad = engine.conv_pad(c.geo0)
        c.tabs["pad"] = pad
        dh = c.hf.shape[0]
        c.tabs["dh"] = dh
        band_all = np.concatenate([p.band_np for p in c.probs]).astype(np.int32)
        c.bufs["bands"] = torch.as_tensor(band_all, device=dev)
        c.tabs["band_np"] = band_all
        c.bufs["h"] = torch.as_tensor(np.ascontiguousarray(c.hf), device=dev)
        c.bufs["hsum"] = torch.as_tensor(c.hf.sum(axis=1), device=dev)
        plan0 = engine.get_plan(d, c.order, c.N, c.geo0.L, c.geo0.margin, C,
                                c.groups[0][1] - c.groups[0][0], c.groups[0][2], dev)
        c.tabs["plan0"] = plan0
        npow, nmax_ax = plan0.npow, plan0.nmax
        per_tap = d * npow * nmax_ax
        c.tabs["per_tap"] = per_tap
        # per precision group: batch passes (every problem's jobs, offset)
        gtabs = []
        grid_stride = C * cells
        for (a, b, gf) in c.groups:
            pl = engine.get_plan(d, c.order, c.N, c.geo0.L, c.geo0.margin, C, b - a, gf, dev)
            st0 = (b - a) * pl.stage_per_h[0]
            inter_stride = {k: pl.chunk * pl.stage_per_h[k] for k in range(1, d)}
            passes = []
            for ps in pl.pass_tables[0]:
                jh = ps.jobs_host
                allj = []
                for j in range(P):
                    t = jh.copy()
                    t[:, 0] += j * (grid_stride if ps.in_stage == d else inter_stride[ps.in_stage])
                    t[:, 1] += j * (st0 if ps.out_stage == 0 else inter_stride[ps.out_stage])
                    t[:, 2] += j * dh * per_tap
                    t[:, 3] += j * dh * d
                    allj.append(t)
                jb = np.concatenate(allj)
                bp = engine.Pass(ps.axis, ps.mode, ps.K, ps.nout, ps.inner, ps.base, ps.max_ro,
                                 torch.as_tensor(jb, device=dev), ps.in_stage, ps.out_stage, jb,
                                 ps.max_np)
                bands_g = band_all[a * d * 2:]
                engine.conv_schedule(bp, gf, bands_g)   # builds and keeps the block table
                passes.append((bp, bands_g))
            gtabs.append((a, b, gf, pl, st0, inter_stride, passes))
        c.tabs["groups"] = gtabs
        c.tabs["probe_pp"] = None
        if any(c.cubic):
            # inverse collocation matrices of every problem's raw axes
            inv = [np.concatenate([engine.cubic_inverse(ax).reshape(-1) for ax in p.geo.axes_raw])
                   for p in c.probs]
            c.tabs["inv_stride"] = int(inv[0].size)
            c.bufs["inv"] = torch.as_tensor(np.concatenate(inv), device=dev)
        if c.variance:
            c.bufs["inv_n"] = torch.as_tensor(
                np.repeat([1.0 / p.n for p in c.probs], rows), device=dev)

    # ------------------------------------------------------------------ fit
    def fit(self, ys: Sequence) -> List[RegressionOutput]:
        """Fit the responses ys (one per problem, as cv_fastlpr takes y)."""
        if len(ys) != self.P:
            raise ValueError(f"expected {self.P} responses, got {len(ys)}")
        yl = []
        for p, y in zip(self.probs, ys):
            ya = y if isinstance(y, torch.Tensor) else np.asarray(y)
            if ya.ndim == 1:
                ya = ya[:, None]
            elif ya.ndim == 2 and ya.shape[0] == 1 and ya.shape[1] > 1:
                ya = ya.T
            elif ya.ndim != 2:
                raise ValueError(f"Response y must be 1D or 2D array. Got {ya.ndim}D array.")
            if ya.shape[0] != p.n:
                raise ValueError(
                    f"Response y must have same number of samples as x. "
                    f"x has {p.n} samples, y has {ya.shape[0]} samples.")
            if not isinstance(ya, torch.Tensor):
                api._check_finite(ya, "Response y", "cv_fastlpr")
            yl.append(ya)
        # responses of different column counts or kinds: one class set each,
        # built on first use and kept
        kinds: Dict[tuple, List[int]] = {}
        for i, ya in enumerate(yl):
            c = bool(ya.is_complex()) if isinstance(ya, torch.Tensor) else bool(np.iscomplexobj(ya))
            kinds.setdefault((ya.shape[1] * (2 if c else 1), c), []).append(i)
        sig = []
        classes: List[_Class] = []
        for (ny, c), idx in kinds.items():
            key = ((ny, c), tuple(idx))
            if key not in self._kindsets:
                self._kindsets[key] = self._setup([self.probs[i] for i in idx], ny, c)
            sig.append(key)
            classes.extend(self._kindsets[key])
        sig = tuple(sig)
        self._last_sig = sig
        dev = self.dev
        ydev = [api._y_channels(y, dev) for y in yl]
        graph = self._graph if (self._graph is not None and self._graph[0] == sig) else None
        if graph is not None:
            for c in classes:
                torch.cat([ydev[p.idx] for p in c.probs], dim=0, out=c.bufs["y_in"])
            graph[1].replay()
            pending = graph[2]
            host = graph[3].cpu().numpy() if graph[3] is not None else None
        else:
            pending = []
            for c in classes:
                y_all = torch.cat([ydev[p.idx] for p in c.probs], dim=0).contiguous()
                pending.append((c, self._run_class(c, y_all)))
            # one read of every class's selection results
            flat = [st["out"].reshape(-1) for (_c, st) in pending]
            host = torch.cat(flat).cpu().numpy() if flat else None
        keep = graph is not None   # replay buffers are overwritten by the next replay
        results: Dict[int, RegressionOutput] = {}
        off = 0
        for c, st in pending:
            nout = st["out"].numel()
            hv = host[off:off + nout].reshape(len(c.probs), -1)
            off += nout
            arrays = self._class_arrays(c, st, hv, keep)
            for j, p in enumerate(c.probs):
                results[p.idx] = self._result(c, st, j, p, hv[j], yl[p.idx], arrays)
        # problems outside the batched path: one call each
        for p, y in zip(self.probs, yl):
            if p.idx not in results:
                x = p.x_host if p.as_numpy and p.x_host is not None else p.x
                results[p.idx] = api.cv_fastlpr(x, y, p.h_given, p.opts)
        return [results[p.idx] for p in self.probs]

    def capture(self, ys: Sequence) -> None:
        """Record the batched stages of a fit of responses like ys into one HIP
        graph; later fit() calls with responses of the same kinds copy them
        into the plan's input buffers and replay the graph.  Runs one eager
        fit first, so that every device table exists before the capture."""
        self._graph = None
        self.fit(ys)
        sig = self._last_sig
        classes = [cl for key in sig for cl in self._kindsets[key]]
        for c in classes:
            c.bufs["y_in"] = torch.empty((c.tabs["rows"], c.ny), dtype=torch.float64,
                                         device=self.dev)
        graph = torch.cuda.CUDAGraph()
        torch.cuda.synchronize()
        with torch.cuda.graph(graph):
            pending = [(c, self._run_class(c, c.bufs["y_in"])) for c in classes]
            flat = [st["out"].reshape(-1) for (_c, st) in pending]
            host_src = torch.cat(flat) if flat else None
        torch.cuda.synchronize()
        self._graph = (sig, graph, pending, host_src)

    def _run_class(self, c: _Class, y_all: torch.Tensor) -> dict:
        """Queue every stage of one class (no host synchronisation)."""
        dev = self.dev
        s = engine.stream_handle()
        f64, i32 = torch.float64, torch.int32
        d, P = c.d, len(c.probs)
        g0 = c.geo0.geom
        gp = _lib.C.byref(g0)
        items = ptr(c.bufs["items"])
        rows, nmax = c.tabs["rows"], c.tabs["nmax"]
        x = c.tabs["x"]
        ny, K, C = c.ny, c.K, c.tabs["C"]
        nval = ny + K
        cells, pad, dh = c.tabs["cells"], c.tabs["pad"], c.tabs["dh"]
        nodes = int(np.prod(c.N))
        rcells = int(np.prod([n_ - 1 for n_ in c.N]))
        nw = (1 << d) * ((1 << d) + 1) // 2
        ntiles = int(np.prod([g0.ntile[k] for k in range(d)]))
        nkeys1 = ntiles * int(np.prod([g0.tile[k] for k in range(d)]))
        nkeys = P * nkeys1
        red_stride = dh * nval + dh * ny
        out_stride = 4 * dh + 5
        st = {}
        # response transform of the variance fit: log(y + 1/n) (api.py:390-396)
        y_raw = y_all[:, 0].contiguous() if c.variance else None
        if c.variance:
            y_all = torch.log(y_all + c.bufs["inv_n"][:, None]).contiguous()
        # zeroed scratch: key counts, spread sums, max keys, reductions, adjoint sums
        # and the workspace of the fixed-order dot / quadratic-form sums
        nsum = engine.sum_part_len(P * dh * nval, -(-max(nodes, rcells) // 256))
        zspec = [(2 * nkeys, i32), (P * C, f64), (P, torch.int64), (P * red_stride, f64),
                 (P * nval * nodes, f64), (P * nw * rcells, f64), (nsum, f64)]
        cnt, sums, key_max, red, q, gram, sump = _zeroed(dev, zspec, s)
        # y statistics: non-finite count and sum of squares per problem
        ystat = torch.empty((P * (2 + ny),), dtype=f64, device=dev)
        ypart = torch.empty((P * 512 * (2 + ny),), dtype=f64, device=dev)
        call("flpr_batch_xstats", items, P, None, 0, ptr(y_all), ny, ptr(ypart), ptr(ystat), s)
        ysq = ystat.view(P, 2 + ny)[:, 2:].contiguous()
        # points: keys, counting sort, packed sorted data
        key = torch.empty((rows,), dtype=i32, device=dev)
        rank = torch.empty((rows,), dtype=i32, device=dev)
        key_start = torch.empty((nkeys + 1,), dtype=i32, device=dev)
        order_idx = torch.empty((rows,), dtype=i32, device=dev)
        call("flpr_batch_points_prepare", gp, items, P, nmax, ptr(x), ptr(key), ptr(cnt[:nkeys]),
             ptr(rank), s)
        call("flpr_points_bin", ptr(key), rows, nkeys, ptr(cnt[:nkeys]), ptr(key_start),
             ptr(cnt[nkeys:]), ptr(rank), ptr(order_idx), s)
        probes, pp = self._probes(c)
        wab = torch.empty((2 * d * rows,), dtype=f64, device=dev)
        vals = torch.empty((nval * rows,), dtype=f64, device=dev)
        rawi = torch.empty((d * rows,), dtype=i32, device=dev)
        rawf = torch.empty((d * rows,), dtype=f64, device=dev)
        call("flpr_batch_points_pack", gp, items, P, nmax, ptr(x), ptr(order_idx), ptr(y_all), ny,
             ptr(probes), K, ptr(wab), ptr(vals), ptr(rawi), ptr(rawf), s)
        # spread, centring
        fcells = int(np.prod([g0.tile[k] + 2 * g0.msp for k in range(d)]))
        part = torch.empty((P * ntiles * C * fcells,), dtype=f64, device=dev)
        grid_stride = C * cells
        grid64 = torch.empty((P * grid_stride + pad,), dtype=f64, device=dev)
        call("flpr_batch_spread", gp, items, P, nmax, ptr(key_start), ptr(wab), ptr(vals), 1,
             nval, ptr(part), ptr(grid64), grid_stride, ptr(sums), s)
        del part
        any32 = any(not gf for (_a, _b, gf) in c.groups)
        shift = torch.empty((P * C,), dtype=f64, device=dev)
        grid32 = (torch.empty((P * grid_stride + pad,), dtype=torch.float32, device=dev)
                  if any32 else None)
        call("flpr_batch_grid_center", P, ptr(grid64), grid_stride, C, cells, ptr(sums),
             ptr(shift), ptr(grid32), s)
        # taps of every problem (its own grid axis), split strips
        plan0 = c.tabs["plan0"]
        npow, nmax_ax, per_tap = plan0.npow, plan0.nmax, c.tabs["per_tap"]
        tap64 = torch.empty((P * dh * per_tap,), dtype=f64, device=dev)
        tap32 = torch.empty((P * dh * per_tap,), dtype=torch.float32, device=dev) if any32 else None
        call("flpr_batch_taps", gp, items, P, ptr(c.bufs["h"]), dh, npow, nmax_ax, ptr(tap64),
             ptr(tap32), s)
        bands = c.bufs["bands"]
        strips, strip_row = (engine.split_strips(tap32, bands, P * dh, d, npow, nmax_ax, s)
                             if any32 else (None, 0))
        # convolution passes, regularisation key per problem
        if c.order == 0:
            lam_scale, eps_add = float(c.reg), float(np.finfo(float).eps)
        else:
            lam_scale, eps_add = 1e-6, 0.0
        stage0s = []
        for (a, b, gf, pl, st0, inter_stride, passes) in c.tabs["groups"]:
            fdt = f64 if gf else torch.float32
            src_grid = grid64 if gf else grid32
            tap_ptr = (tap64.data_ptr() + a * per_tap * 8) if gf else \
                (strips.data_ptr() + a * strip_row * 2)
            band_ptr = bands.data_ptr() + a * d * 2 * 4
            stage0 = torch.empty((P * st0 + pad,), dtype=fdt, device=dev)
            inter = {k: torch.empty((P * inter_stride[k] + pad,), dtype=fdt, device=dev)
                     for k in range(1, d)}
            for (bp, bands_g) in passes:
                src = src_grid if bp.in_stage == d else inter[bp.in_stage]
                dst = stage0 if bp.out_stage == 0 else inter[bp.out_stage]
                engine.launch_conv(bp, gf, ptr(src), ptr(dst), tap_ptr, nmax_ax, band_ptr,
                                   bands_g, s)
            del inter
            if c.order == 0:
                offs, absval = [pl.s_off[0]], 0
            else:
                offs, absval = pl.diag_off, 1
            call("flpr_batch_max_moment", int(gf), P, ptr(stage0), st0, pl.stage_per_h[0], b - a,
                 i64_array(offs), len(offs), pl.nodes, absval, ptr(key_max), s)
            stage0s.append((a, b, gf, pl, st0, stage0))
        m_stride = dh * ny * nodes
        per_h0 = plan0.stage_per_h[0]
        if not any(c.cubic):
            # adjoint sums, solves (+ dots), residual quadratic forms
            apart = torch.empty((P * engine.adjoint_workspace(g0, d, ntiles, max(nval, nw)),),
                                dtype=f64, device=dev)
            call("flpr_batch_adjoint", gp, items, P, nmax, ptr(key_start), ptr(rawi), ptr(rawf),
                 ptr(vals), nval, ptr(apart), ptr(q), ptr(gram), s)
            del apart
            m_y = torch.empty((P * m_stride,), dtype=f64, device=dev)
            for (a, b, gf, pl, st0, stage0) in stage0s:
                call("flpr_batch_solve", int(gf), P, ptr(stage0), st0, per_h0, b - a, nodes,
                     c.order, pl.nt, i64_array(pl.s_off), i64_array(pl.t_off), ptr(key_max),
                     lam_scale, eps_add, ptr(shift), nval, ny,
                     m_y.data_ptr() + a * ny * nodes * 8, m_stride, ptr(q),
                     red.data_ptr() + a * nval * 8, red_stride, ptr(sump), nsum, s)
            call("flpr_batch_quadform", gp, P, ptr(m_y), m_stride, ny, dh, ptr(gram),
                 red.data_ptr() + dh * nval * 8, red_stride, ptr(sump), nsum, s)
        else:
            # every channel's grid solutions, then the sums at the points
            # (spline or multilinear), as cv_fastlpr's cubic route
            a_stride = dh * nval * nodes
            m_all = torch.empty((P * a_stride,), dtype=f64, device=dev)
            for (a, b, gf, pl, st0, stage0) in stage0s:
                call("flpr_batch_solve", int(gf), P, ptr(stage0), st0, per_h0, b - a, nodes,
                     c.order, pl.nt, i64_array(pl.s_off), i64_array(pl.t_off), ptr(key_max),
                     lam_scale, eps_add, ptr(shift), nval, nval,
                     m_all.data_ptr() + a * nval * nodes * 8, a_stride, None, None, 0, None, 0,
                     s)
            coef = torch.empty_like(m_all)
            work = torch.empty_like(m_all) if d > 1 else coef
            call("flpr_batch_spline_coef", gp, P, ptr(c.bufs["inv"]), c.tabs["inv_stride"],
                 ptr(m_all), dh * nval, ptr(work), ptr(coef), s)
            del work
            nblk = max(1, min(64, -(-nmax // 256)))
            ppart = torch.empty((P * (-(-(dh * nval) // 8)) * nblk * 8,), dtype=f64, device=dev)
            call("flpr_batch_point_sums", gp, items, P, nmax, ptr(x), ptr(y_all), ny, ptr(probes),
                 K, ptr(m_all), ptr(coef), dh, int(c.cubic[0]), int(c.cubic[1]), ptr(ppart),
                 ptr(red), red_stride, s)
            del coef, ppart
            m_y = m_all.view(P, dh, nval, nodes)[:, :, :ny, :].contiguous().view(-1)
            (ysq,) = _zeroed(dev, [(P * ny, f64)], s)
        # selection per problem, fitted values at the selected bandwidth
        out = torch.empty((P * out_stride,), dtype=f64, device=dev)
        sel = torch.empty((P,), dtype=i32, device=dev)
        call("flpr_batch_select", items, P, ptr(red), red_stride, dh, nval, ny, K, ptr(ysq),
             ptr(pp), ptr(c.bufs["hsum"]), int(c.cplx), c.agg, c.dstd, ptr(out), out_stride,
             ptr(sel), s)
        yhat = torch.empty((rows, ny), dtype=f64, device=dev)
        resid = torch.empty((rows, ny), dtype=f64, device=dev)
        grid_v = part_v = None
        if c.variance:
            grid_v = torch.empty((P * nodes,), dtype=f64, device=dev)
            part_v = torch.empty((P * 1024,), dtype=f64, device=dev)
        call("flpr_batch_fit_output", gp, items, P, nmax, ptr(x), ptr(m_y), m_stride, ny,
             ptr(sel), ptr(y_raw if c.variance else y_all), int(c.variance), ptr(yhat),
             ptr(resid), ptr(grid_v), ptr(part_v), out.data_ptr() + (4 * dh + 3) * 8,
             out_stride, s)
        # non-finite count of y into the spare slot of out
        out.view(P, out_stride)[:, 4 * dh + 4].copy_(ystat.view(P, 2 + ny)[:, 1])
        st.update(out=out, m_y=m_y, yhat=yhat, resid=resid, grid_v=grid_v, stage0s=stage0s,
                  y_all=y_all)
        return st

    def _probes(self, c: _Class):
        """The probe matrix (at least nmax rows; every problem reads its
        leading rows) and p'p of every problem's rows, on the device."""
        if c.tabs.get("probe_pp") is None:
            seed = int(c.probs[0].opts.get("dof_seed", 42))
            rows = []
            base = None
            for p in c.probs:
                pb, pph, _ppd = api._probes(p.n, c.K, seed, None, self.dev)
                rows.append(pph)
            base = api._probe_matrix(c.tabs["nmax"], c.K, seed, self.dev)[1]
            c.tabs["probe_pp"] = torch.as_tensor(np.concatenate(rows), device=self.dev)
            c.tabs["probe_base"] = base
        return c.tabs["probe_base"], c.tabs["probe_pp"]

    def _class_arrays(self, c: _Class, st: dict, hv: np.ndarray, copy: bool) -> dict:
        """The device arrays the results of class c slice, made with a few
        launches for the whole class (not one per problem): the fitted values
        and residuals (copied once when a later replay would overwrite them),
        and, at each problem's selected bandwidth, its grid values and its
        S_00 row (gathered per precision group).  hv: the class's host rows
        of the selection results."""
        dh = c.tabs["dh"]
        P = len(c.probs)
        nodes = int(np.prod(c.N))
        # a problem without a finite score raises in _result; its index is
        # clamped here only so that the gathers stay inside the buffers
        sel = np.clip(np.nan_to_num(hv[:, 4 * dh]), 0, dh - 1).astype(np.int64)
        dev = self.dev
        yhat, resid = st["yhat"], st["resid"]
        if copy:
            yhat, resid = yhat.clone(), resid.clone()
        jj = torch.arange(P, device=dev)
        if c.variance:
            grid = st["grid_v"].view(P, 1, nodes)
            if copy:
                grid = grid.clone()
        else:
            # advanced indexing: a gathered copy, independent of the buffer
            grid = st["m_y"].view(P, dh, c.ny, nodes)[jj, torch.as_tensor(sel, device=dev)]
        s0 = torch.empty((P, nodes), dtype=torch.float64, device=dev)
        for (a, b, gf, pl, st0, stage0) in st["stage0s"]:
            js = np.nonzero((sel >= a) & (sel < b))[0]
            if js.size == 0:
                continue
            per_h = pl.stage_per_h[0]
            off0 = pl.s_off[0]
            blk = stage0[:P * st0].view(P, st0)[:, :(b - a) * per_h].view(P, b - a, per_h)
            jt = torch.as_tensor(js, device=dev)
            rows = blk[jt, torch.as_tensor(sel[js] - a, device=dev), off0:off0 + nodes]
            s0[jt] = rows.to(torch.float64)
        if c.order == 0:
            s0 += float(np.finfo(float).eps)
        return {"yhat": yhat, "resid": resid, "grid": grid, "s0": s0}

    def _result(self, c: _Class, st: dict, j: int, p: _Problem, hv: np.ndarray, y_in,
                arrays: dict):
        """RegressionOutput of problem j of class c from its host row of the
        selection results (the layout of cv_fastlpr's readback) and the class
        arrays of _class_arrays."""
        dh = c.tabs["dh"]
        if hv[4 * dh + 4] > 0:
            api._check_finite(y_in, "Response y", "cv_fastlpr")
        if int(hv[4 * dh + 2]) == 0:
            raise ValueError(
                "All GCV scores are NaN or Inf. Regression failed for all bandwidths. "
                "Try: (1) Use larger bandwidth h, (2) Increase grid points, "
                "(3) Check for data quality issues")
        if p.warn:
            warnings.warn(p.warn, stacklevel=3)
        scores, errs = hv[:dh].copy(), hv[dh:2 * dh].copy()
        dof_all, dof_se = hv[2 * dh:3 * dh].copy(), hv[3 * dh:4 * dh].copy()
        sel, idmin = int(hv[4 * dh]), int(hv[4 * dh + 1])
        hf = c.hf
        gcv_result = {"bandwidths": hf, "gcv_m": scores, "gcv_sd": errs, "idmin": idmin,
                      "id1se": sel, "h1se": hf[sel], "hmin": hf[idmin],
                      "dof": dof_all, "dof_stderr": dof_se}
        r0, r1 = int(c.tabs["pts"][j]), int(c.tabs["pts"][j + 1])
        out = arrays["yhat"][r0:r1]
        resid_r = arrays["resid"][r0:r1]
        dy = c.dy
        if c.cplx:
            yhat = torch.view_as_complex(out.reshape(p.n, dy, 2))
            resid = torch.view_as_complex(resid_r.reshape(p.n, dy, 2))
        else:
            yhat, resid = out, resid_r
        if dy == 1:
            yhat = yhat[:, 0]
        resid = resid.reshape(-1)
        d_norm = None
        if c.variance:
            d_norm = 1.0 / float(hv[4 * dh + 3])
        grid_vals = arrays["grid"][j]
        # S_00 at the selected bandwidth (gathered per class)
        s0 = arrays["s0"][j].reshape(c.N)
        geo = p.geo
        interp = api.GridInterpolant(geo, geo.axes_raw, grid_vals, c.N, c.cplx, dy, p.as_numpy,
                                     dy == 1)

        def conv(t):
            return t.cpu().numpy() if p.as_numpy else t

        metadata = {"options": p.opts, "order": c.order, "bandwidths_tested": hf,
                    "fitted_at_samples": conv(yhat), "residuals": conv(resid),
                    "mean": geo.mean, "std": geo.std, "dof": float(dof_all[sel]),
                    "dof_stderr": float(dof_se[sel]), "dof_shared": False, "batched": True}
        return RegressionOutput(yhat=conv(yhat), grid=tuple(geo.axes_raw), h=hf[sel],
                                gcv_yhat=gcv_result, fpp_yhat=interp, metadata=metadata,
                                dof=float(dof_all[sel]), dof_stderr=float(dof_se[sel]),
                                d=d_norm, s0=conv(s0))


def ptr(t):
    return None if t is None else t.data_ptr()


def _items(ps, geos):
    """Device FlprBatchItem table (geometry left zero where geos[j] is None)."""
    P = len(ps)
    items = (FlprBatchItem * P)()
    pt = 0
    for j, p in enumerate(ps):
        if geos[j] is not None:
            items[j].g = geos[j]
        items[j].n = p.n
        items[j].pt = pt
        pt += p.n
    dev = ps[0].x.device
    t = torch.as_tensor(np.frombuffer(items, dtype=np.uint8).copy(), device=dev)
    return t, pt


def _zeroed(dev, specs, stream):
    """Zero-filled buffers from one allocation, cleared by flpr_fill_zero (a
    kernel, so that a captured graph clears them on every replay)."""
    offs, sizes, total = [], [], 0
    for (cnt, dt) in specs:
        offs.append(total)
        nb = int(cnt) * dt.itemsize
        sizes.append(nb)
        total += (nb + 255) // 256 * 256
    arena = torch.empty((max(total, 256),), dtype=torch.uint8, device=dev)
    call("flpr_fill_zero", arena.data_ptr(), arena.numel(), stream)
    out = []
    for (cnt, dt), o, nb in zip(specs, offs, sizes):
        out.append(arena[o:o + nb].view(dt) if cnt else None)
    return out
