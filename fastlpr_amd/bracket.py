"""The learned bandwidth bracket: ``cv_fastlpr(..., {'bandwidth_bracket': 'learned'})``.

The reference sweeps every bandwidth of the list (api.py:664-740).  Here a
small MLP reads a 64-number summary of the data's spread grid and proposes
where the GCV minimum lies, so that only a bracket of ``BRACKET`` consecutive
bandwidths around the proposal is convolved with every channel and solved:

1. features: a spread of the channels (1, y, y^2) at the fit's geometry and
   ``flpr_bracket_features`` (multi-scale increments of the local mean of y,
   noise and density summaries);
2. the MLP (``flpr_mlp_forward``, 64 -> 128 -> 128 -> 2, GELU, bf16 MFMA with
   float32 accumulation) gives the centre and width of the bracket in
   log10 h;
3. lambda = 1e-6 max |S_kk| over every node and EVERY bandwidth of the list
   (regression.py:243-267) from a pass that convolves the ones channel only,
   so the bracket's GCV scores equal the full sweep's at those bandwidths;
4. the full fit on the bracket's bandwidths with that lambda, GCV and the
   1-SE choice among them.

When the chosen bandwidth sits on an inner edge of the bracket (the minimum
may lie outside it), or the data fall outside what the model was trained on
(another dimension count, cubic GCV interpolation, fewer than ``BRACKET``
bandwidths), the full sweep runs instead: the sweep stays the fallback and
the judge.

The model is trained on the GPU by the same kernels (``flpr_mlp_backward``,
``flpr_mlp_loss``, ``flpr_mlp_adam``) on GCV curves the engine itself
computes (``train``): synthetic 2-D data sets, n log-uniform in [1e4, 1e6],
responses that are sums of 1-5 plane sinusoids with frequencies U(1, 8) plus
noise of standard deviation U(0.01, 0.5), labelled by the full 30-bandwidth
GCV sweep in [0.005, 0.5] (order 2, grid 256 x 256).  The weights ship in
``bracket_mlp_2d.npz`` beside this file.
"""
from __future__ import annotations

import math
import os
import time
import warnings
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib, engine
from ._lib import FlprBatchItem, call

BRACKET = 5
IN, HID, OUT = 64, 128, 2
MODEL_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bracket_mlp_2d.npz")
# the bracket's GCV minimum may be this much above the full sweep's and still
# count as a hit (the configuration's acceptance rule)
HIT_TOL = 0.02


def ptr(t):
    return None if t is None else t.data_ptr()


# ---------------------------------------------------------------------------
# features
# ---------------------------------------------------------------------------

def features(xs: Sequence[torch.Tensor], ys: Sequence[torch.Tensor], N=(256, 256),
             stats: Optional[List[engine.Stats]] = None) -> torch.Tensor:
    """(P, 64) float32 features of 2-D problems (x (n, 2), y (n,) float64 on
    the device), computed at each problem's cv_fastlpr geometry for the grid
    N; one launch per stage for all problems."""
    sg = spread_grids(xs, ys, N, stats)
    return grid_features(sg)


def grid_features(sg: dict) -> torch.Tensor:
    """flpr_bracket_features on grids laid out as spread_grids returns them
    (grid_stride elements per problem, channel 0 = sum w, channel 1 = sum w y)."""
    P, G0, G1 = sg["P"], sg["G"][0], sg["G"][1]
    dev = sg["grid"].device
    f64 = torch.float64
    fpart = torch.empty((P * 57 * 64 * 8,), dtype=f64, device=dev)
    feat = torch.empty((P, IN), dtype=torch.float32, device=dev)
    call("flpr_bracket_features", P, sg["grid"].data_ptr(), sg.get("stride", 2 * G0 * G1), G0, G1,
         ptr(sg["floor"]), ptr(sg["npts"]), ptr(sg["ysq"]), ptr(sg["dx"]), ptr(fpart), ptr(feat),
         engine.stream_handle())
    return feat


def grid_spacing(geo: engine.Geometry) -> float:
    """Geometric mean grid spacing of a 2-D geometry in z-scored units."""
    return math.sqrt(abs(geo.axes_scaled[0][1] - geo.axes_scaled[0][0]) *
                     abs(geo.axes_scaled[1][1] - geo.axes_scaled[1][0]))


def spread_grids(xs: Sequence[torch.Tensor], ys: Sequence[torch.Tensor], N=(256, 256),
                 stats: Optional[List[engine.Stats]] = None) -> dict:
    """The spread grids of the channels (1, y) of every problem at its
    cv_fastlpr geometry (flpr_batch_* spread path), with the per-problem
    floor (1e-3 of the mean of channel 0), point count, sum of y^2 and grid
    spacing the features read.  Returns a dict with grid (P x 2 x G0 x G1
    float64), floor, npts, ysq, dx (P each), G and P."""
    dev = xs[0].device
    P = len(xs)
    s = engine.stream_handle()
    f64, i32 = torch.float64, torch.int32
    ns = [int(x.shape[0]) for x in xs]
    pts = np.concatenate([[0], np.cumsum(ns)])
    rows = int(pts[-1])
    nmax = int(max(ns))
    x_all = torch.cat(xs, dim=0).contiguous()
    y_all = torch.cat([y.reshape(-1) for y in ys], dim=0).contiguous()
    if stats is None:
        stats = _xstats(xs, x_all, pts)
    geos = []
    for st, n in zip(stats, ns):
        acc = max(6 - int(np.ceil(np.log10(n))), 4)
        geos.append(engine.geometry_for(st, tuple(N), acc)[0])
    g0 = geos[0].geom
    if any(tuple(g.geom.G[k] for k in range(2)) != tuple(g0.G[k] for k in range(2)) or
           g.geom.msp != g0.msp for g in geos):
        raise ValueError("bracket features: problems of one call must share the grid structure")
    items = (FlprBatchItem * P)()
    for j, g in enumerate(geos):
        items[j].g = g.geom
        items[j].n = ns[j]
        items[j].pt = int(pts[j])
    it = engine.upload(np.frombuffer(bytes(items), dtype=np.uint8).copy(), torch.uint8, dev)
    gp = _lib.C.byref(g0)
    ntiles = int(g0.ntile[0] * g0.ntile[1])
    nkeys1 = ntiles * int(g0.tile[0] * g0.tile[1])
    nkeys = P * nkeys1
    cnt = torch.zeros((2 * nkeys,), dtype=i32, device=dev)
    sums = torch.zeros((P * 2,), dtype=f64, device=dev)
    key = torch.empty((rows,), dtype=i32, device=dev)
    rank = torch.empty((rows,), dtype=i32, device=dev)
    key_start = torch.empty((nkeys + 1,), dtype=i32, device=dev)
    order = torch.empty((rows,), dtype=i32, device=dev)
    call("flpr_batch_points_prepare", gp, ptr(it), P, nmax, ptr(x_all), ptr(key), ptr(cnt[:nkeys]),
         ptr(rank), s)
    call("flpr_points_bin", ptr(key), rows, nkeys, ptr(cnt[:nkeys]), ptr(key_start),
         ptr(cnt[nkeys:]), ptr(rank), ptr(order), s)
    wab = torch.empty((4 * rows,), dtype=f64, device=dev)
    vals = torch.empty((rows,), dtype=f64, device=dev)
    call("flpr_batch_points_pack", gp, ptr(it), P, nmax, ptr(x_all), ptr(order), ptr(y_all), 1,
         None, 0, ptr(wab), ptr(vals), None, None, s)
    cells = int(g0.G[0] * g0.G[1])
    fcells = int((g0.tile[0] + 2 * g0.msp) * (g0.tile[1] + 2 * g0.msp))
    part = torch.empty((P * ntiles * 2 * fcells,), dtype=f64, device=dev)
    grid = torch.empty((P * 2 * cells,), dtype=f64, device=dev)
    call("flpr_batch_spread", gp, ptr(it), P, nmax, ptr(key_start), ptr(wab), ptr(vals), 1, 1,
         ptr(part), ptr(grid), 2 * cells, ptr(sums), s)
    del part
    floor = (sums.view(P, 2)[:, 0] * (1e-3 / cells)).contiguous()
    npts = engine.upload(np.asarray(ns, dtype=np.float64), f64, dev)
    ysq = torch.stack([(y.reshape(-1) * y.reshape(-1)).sum() for y in ys]).contiguous()
    dx = engine.upload(np.asarray([grid_spacing(g) for g in geos]), f64, dev)
    return {"grid": grid, "floor": floor, "npts": npts, "ysq": ysq, "dx": dx,
            "G": (int(g0.G[0]), int(g0.G[1])), "P": P, "stride": 2 * cells}


def selector(mdl: "BracketModel", ysq_dev: torch.Tensor, info: dict):
    """The select_h hook of engine.run_grid for one 2-D fit.  The hook queues
    the features of the fit's own spread grid (channels 0 and 1 before
    centring, ysq_dev = sum y^2 on the device) and the MLP's forward pass,
    and returns a callable; the engine queues the list-wide regularisation
    pass, then calls it.  The callable reads the bracket centre and width
    (one host read, by which time the device is busy with that pass) and
    returns the BRACKET indices around the centre; info receives them.  The
    hook returns None (full sweep) when the list is no longer than the
    bracket."""
    def select(geo, grid64, cells, sums, h):
        if len(geo.N) != 2 or h.shape[0] <= BRACKET:
            return None
        dev = grid64.device
        sg = {"grid": grid64, "stride": 2 * cells,
              "G": (int(geo.geom.G[0]), int(geo.geom.G[1])), "P": 1,
              "floor": (sums[:1] * (1e-3 / cells)).contiguous(),
              "npts": engine.device_table(np.asarray([float(info["n"])]), torch.float64, dev),
              "ysq": ysq_dev.reshape(-1)[:1].contiguous(),
              "dx": engine.device_table(np.asarray([grid_spacing(geo)]), torch.float64, dev)}
        out = mdl.queue(grid_features(sg))

        def pick():
            center, width = mdl.read(out)[0]
            idx = choose(float(center), h)
            info.update(indices=idx, n_full=int(h.shape[0]), center=float(center),
                        width=float(width), hlist=h.copy(), fallback=False)
            return idx
        return pick
    return select


def _xstats(xs, x_all, pts) -> List[engine.Stats]:
    P = len(xs)
    d = int(xs[0].shape[1])
    items = (FlprBatchItem * P)()
    for j in range(P):
        items[j].n = int(pts[j + 1] - pts[j])
        items[j].pt = int(pts[j])
    dev = x_all.device
    it = engine.upload(np.frombuffer(bytes(items), dtype=np.uint8).copy(), torch.uint8, dev)
    per = 5 * d + 2
    part = torch.empty((P * 512 * per,), dtype=torch.float64, device=dev)
    out = torch.empty((P * per,), dtype=torch.float64, device=dev)
    call("flpr_batch_xstats", ptr(it), P, ptr(x_all), d, None, 0, ptr(part), ptr(out),
         engine.stream_handle())
    o = out.cpu().numpy().reshape(P, per)
    res = []
    for row in o:
        w = row[:5 * d].reshape(d, 5)
        std = np.sqrt(w[:, 3] / np.maximum(w[:, 4] - 1, 1))
        std = np.where(std == 0, 1.0, std)
        res.append(engine.Stats(w[:, 0].copy(), w[:, 1].copy(), w[:, 2].copy(), std, 0, 0))
    return res


# ---------------------------------------------------------------------------
# the model
# ---------------------------------------------------------------------------

class BracketModel:
    """Weights, feature and target standardisation, and the forward pass."""

    def __init__(self, params: np.ndarray, fmean: np.ndarray, fstd: np.ndarray,
                 tmean: np.ndarray, tstd: np.ndarray, meta: Optional[dict] = None):
        self.params = np.asarray(params, dtype=np.float32)
        self.fmean = np.asarray(fmean, dtype=np.float32)
        self.fstd = np.asarray(fstd, dtype=np.float32)
        self.tmean = np.asarray(tmean, dtype=np.float64)
        self.tstd = np.asarray(tstd, dtype=np.float64)
        self.meta = dict(meta or {})
        self._dev: Dict[str, tuple] = {}

    @classmethod
    def load(cls, path: str = MODEL_FILE) -> "BracketModel":
        if not os.path.exists(path):
            raise FileNotFoundError(f"bracket model not found at {path}; train it with "
                                    "fastlpr_amd.bracket.train()")
        z = np.load(path, allow_pickle=False)
        meta = {k[5:]: z[k] for k in z.files if k.startswith("meta_")}
        return cls(z["params"], z["fmean"], z["fstd"], z["tmean"], z["tstd"], meta)

    def save(self, path: str = MODEL_FILE) -> None:
        extra = {f"meta_{k}": np.asarray(v) for k, v in self.meta.items()}
        np.savez(path, params=self.params, fmean=self.fmean, fstd=self.fstd, tmean=self.tmean,
                 tstd=self.tstd, **extra)

    def _on(self, dev):
        key = str(dev)
        hit = self._dev.get(key)
        if hit is None:
            hit = (torch.as_tensor(self.params, device=dev),
                   torch.as_tensor(self.fmean, device=dev), torch.as_tensor(self.fstd, device=dev))
            self._dev[key] = hit
        return hit

    def predict(self, feat: torch.Tensor) -> np.ndarray:
        """(P, 2) bracket centre and width in log10 h for raw features (P, 64)."""
        return self.read(self.queue(feat))

    def queue(self, feat: torch.Tensor) -> torch.Tensor:
        """The forward pass for raw features (P, 64), queued on the device:
        the standardised outputs (P, 2) float32, not read back."""
        p, fm, fs = self._on(feat.device)
        X = ((feat - fm) / fs).contiguous()
        return forward(p, X)[-1]

    def read(self, O: torch.Tensor) -> np.ndarray:
        """Bracket centre and width in log10 h from queue()'s outputs (one
        host read)."""
        return O.double().cpu().numpy() * self.tstd + self.tmean


def forward(params: torch.Tensor, X: torch.Tensor):
    """flpr_mlp_forward: returns (Z1, H1, Z2, H2, O)."""
    M = X.shape[0]
    dev = X.device
    Z1 = torch.empty((M, HID), dtype=torch.float32, device=dev)
    H1, Z2, H2 = torch.empty_like(Z1), torch.empty_like(Z1), torch.empty_like(Z1)
    O = torch.empty((M, OUT), dtype=torch.float32, device=dev)
    call("flpr_mlp_forward", ptr(params), M, ptr(X), ptr(Z1), ptr(H1), ptr(Z2), ptr(H2), ptr(O),
         engine.stream_handle())
    return Z1, H1, Z2, H2, O


def backward(params: torch.Tensor, X: torch.Tensor, acts, dO: torch.Tensor) -> torch.Tensor:
    """flpr_mlp_backward: the gradient of sum(O * dO) in the parameter layout."""
    M = X.shape[0]
    Z1, H1, Z2, H2, _O = acts
    grads = torch.empty_like(params)
    work = torch.empty((2 * M * HID,), dtype=torch.float32, device=X.device)
    call("flpr_mlp_backward", ptr(params), M, ptr(X), ptr(Z1), ptr(H1), ptr(Z2), ptr(H2),
         ptr(dO.contiguous()), ptr(grads), ptr(work), engine.stream_handle())
    return grads


_MODEL: Optional[BracketModel] = None


def model() -> BracketModel:
    global _MODEL
    if _MODEL is None:
        _MODEL = BracketModel.load()
    return _MODEL


# ---------------------------------------------------------------------------
# labels and targets from the full GCV sweep
# ---------------------------------------------------------------------------

def targets(gcv: np.ndarray, sel: int, h: np.ndarray) -> Tuple[float, float]:
    """Bracket centre (log10 of the selected bandwidth) and width (log10 span
    of the contiguous run of bandwidths around it whose score is within
    HIT_TOL of the minimum)."""
    lh = np.log10(np.asarray(h, dtype=float))
    fin = np.isfinite(gcv)
    gmin = np.min(gcv[fin])
    ok = fin & (gcv <= gmin * (1.0 + HIT_TOL))
    lo = hi = sel
    while lo > 0 and ok[lo - 1]:
        lo -= 1
    while hi < len(gcv) - 1 and ok[hi + 1]:
        hi += 1
    return float(lh[sel]), float(lh[hi] - lh[lo])


def choose(center: float, h: np.ndarray, k: int = BRACKET) -> np.ndarray:
    """Indices of the k consecutive bandwidths of h (sorted by size along the
    first column's order) around the one nearest to 10^center."""
    lh = np.log10(np.asarray(h, dtype=float)).mean(axis=1)
    dh = len(lh)
    if dh <= k:
        return np.arange(dh)
    c = int(np.argmin(np.abs(lh - center)))
    a = min(max(c - k // 2, 0), dh - k)
    return np.arange(a, a + k)


# ---------------------------------------------------------------------------
# training data: synthetic sums of sinusoids, labelled by the full sweep
# ---------------------------------------------------------------------------

def synthetic(i: int, dev, n_range=(1e4, 1e6)):
    """Data set i: x ~ U([0,1]^2), y = sum of 1-5 plane sinusoids (frequencies
    U(1, 8), random directions, amplitudes U(0.5, 1.5) / sqrt(J), phases
    U(0, 2 pi)) + noise N(0, s^2), s ~ U(0.01, 0.5); n log-uniform in
    n_range.  Drawn on the device from a generator seeded with i."""
    rs = np.random.default_rng(10_000_000 + i)
    n = int(round(10 ** rs.uniform(math.log10(n_range[0]), math.log10(n_range[1]))))
    J = int(rs.integers(1, 6))
    freq = rs.uniform(1, 8, J)
    ang = rs.uniform(0, 2 * math.pi, J)
    amp = rs.uniform(0.5, 1.5, J) / math.sqrt(J)
    ph = rs.uniform(0, 2 * math.pi, J)
    sd = rs.uniform(0.01, 0.5)
    g = torch.Generator(device=dev)
    g.manual_seed(i)
    x = torch.rand((n, 2), generator=g, dtype=torch.float64, device=dev)
    y = torch.zeros((n,), dtype=torch.float64, device=dev)
    for j in range(J):
        k1, k2 = freq[j] * math.cos(ang[j]), freq[j] * math.sin(ang[j])
        y += amp[j] * torch.sin(2 * math.pi * (k1 * x[:, 0] + k2 * x[:, 1]) + ph[j])
    y += sd * torch.randn((n,), generator=g, dtype=torch.float64, device=dev)
    return x, y


def label(indices: Sequence[int], hl: np.ndarray, options: dict, dev, chunk: int = 96,
          log=None):
    """Features and full-sweep GCV curves of the synthetic data sets."""
    from .batch import BatchPlan
    feats, curves, sels = [], [], []
    t0 = time.time()
    for c0 in range(0, len(indices), chunk):
        ids = indices[c0:c0 + chunk]
        data = [synthetic(i, dev) for i in ids]
        xs = [d[0] for d in data]
        ys = [d[1] for d in data]
        feats.append(features(xs, ys, tuple(options["N"])).cpu().numpy())
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            res = BatchPlan(xs, hl, options).fit(ys)
        for r in res:
            curves.append(np.asarray(r.gcv_yhat["gcv_m"], dtype=float))
            sels.append(int(r.gcv_yhat["id1se"]))
        del res, data, xs, ys
        if log:
            log(f"labelled {min(c0 + chunk, len(indices))} / {len(indices)} "
                f"({time.time() - t0:.0f} s)")
    return np.concatenate(feats), np.asarray(curves), np.asarray(sels)


def train(n_datasets: int = 20000, epochs: int = 60, batch: int = 256, lr: float = 1e-3,
          seed: int = 0, path: Optional[str] = MODEL_FILE, log=print,
          N=(256, 256), order: int = 2, n_val: int = 1000) -> BracketModel:
    """Label n_datasets synthetic data sets with the full sweep, then train
    the MLP on the GPU with the project's kernels (bf16 MFMA forward and
    backward, float32 Adam).  The last n_val data sets are held out; the
    model's hit rate on them is stored in its metadata."""
    dev = torch.device("cuda", torch.cuda.current_device())
    h = np.logspace(np.log10(0.005), np.log10(0.5), 30)
    hl = np.column_stack([h, h])
    options = {"order": order, "N": list(N)}
    F, G, S = label(list(range(n_datasets)), hl, options, dev, log=log)
    T = np.array([targets(g, s, h) for g, s in zip(G, S)])
    ntr = n_datasets - n_val
    fmean = F[:ntr].mean(axis=0)
    fstd = F[:ntr].std(axis=0) + 1e-6
    tmean = T[:ntr].mean(axis=0)
    tstd = T[:ntr].std(axis=0) + 1e-6
    Xd = torch.as_tensor(((F - fmean) / fstd).astype(np.float32), device=dev)
    Td = torch.as_tensor(((T - tmean) / tstd).astype(np.float32), device=dev)
    rs = np.random.default_rng(seed)
    params = np.zeros(IN * HID + HID + HID * HID + HID + OUT * HID + OUT, dtype=np.float32)
    o = 0
    for (fo, fi) in ((HID, IN), (HID, HID), (OUT, HID)):
        params[o:o + fo * fi] = rs.standard_normal(fo * fi) * math.sqrt(2.0 / fi)
        o += fo * fi + fo
    assert o == params.size == int(_lib.load().flpr_mlp_param_count())
    p = torch.as_tensor(params, device=dev)
    m = torch.zeros_like(p)
    v = torch.zeros_like(p)
    loss = torch.zeros((1,), dtype=torch.float32, device=dev)
    step = 0
    s = engine.stream_handle()
    for ep in range(epochs):
        perm = torch.as_tensor(rs.permutation(ntr), device=dev)
        for b0 in range(0, ntr - batch + 1, batch):
            idx = perm[b0:b0 + batch]
            X = Xd[idx].contiguous()
            Tb = Td[idx].contiguous()
            acts = forward(p, X)
            dO = torch.empty_like(acts[-1])
            call("flpr_mlp_loss", batch, ptr(acts[-1]), ptr(Tb), 1.0, 0.5, ptr(dO), ptr(loss), s)
            g = backward(p, X, acts, dO)
            step += 1
            lr_t = lr * 0.5 * (1 + math.cos(math.pi * step / (epochs * (ntr // batch))))
            call("flpr_mlp_adam", p.numel(), ptr(p), ptr(g), ptr(m), ptr(v), lr_t, 0.9, 0.999,
                 1e-8, step, 1e-5, s)
        if log and (ep % 10 == 0 or ep == epochs - 1):
            log(f"epoch {ep}: loss {float(loss.item()):.4f}")
    mdl = BracketModel(p.cpu().numpy(), fmean, fstd, tmean, tstd)
    # held-out hit rate: the bracket's best score within HIT_TOL of the sweep's
    pred = mdl.predict(torch.as_tensor(F[ntr:].astype(np.float32), device=dev))
    hits = edge = 0
    for (c, _w), g, sel in zip(pred, G[ntr:], S[ntr:]):
        idx = choose(c, hl)
        gmin = np.nanmin(g)
        best = int(idx[np.nanargmin(g[idx])])
        hits += int(np.nanmin(g[idx]) <= gmin * (1 + HIT_TOL))
        edge += int(best in (idx[0], idx[-1]) and best not in (0, len(g) - 1))
    mdl.meta = {"n_datasets": n_datasets, "n_val": n_val, "epochs": epochs, "order": order,
                "grid": np.asarray(N), "h_range": np.asarray([0.005, 0.5]), "n_h": 30,
                "val_hit_rate": hits / n_val, "val_edge_rate": edge / n_val,
                "final_loss": float(loss.item())}
    if log:
        log(f"held-out hit rate {hits / n_val:.3f}, edge rate {edge / n_val:.3f}")
    if path:
        mdl.save(path)
    return mdl


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="train the learned bandwidth bracket on the GPU")
    ap.add_argument("--datasets", type=int, default=20000)
    ap.add_argument("--epochs", type=int, default=150)
    ap.add_argument("--val", type=int, default=1000)
    ap.add_argument("--out", default=MODEL_FILE)
    a = ap.parse_args()
    m = train(n_datasets=a.datasets, epochs=a.epochs, n_val=a.val, path=a.out,
              log=lambda s: print(s, flush=True))
    print({k: (v.tolist() if hasattr(v, "tolist") else v) for k, v in m.meta.items()})
