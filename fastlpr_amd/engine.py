"""Device pipeline of one fit: the cv_fastlpr / cv_fastkde hot path on the GPU.

Stages (each a C-ABI call into libfastlpr_amd.so, all on the current HIP stream):

  1. flpr_xstats            min / max / mean / M2 of x          (model.py:58-68)
  2. flpr_points_prepare    knot positions, spread cells, tiles (convolution.py:169-199,
                                                                 nufftn_type1.py:140-176)
     flpr_points_bin        counting sort of points by tile
  3. flpr_spread            Gaussian gridding of 1, y, probes   (nufftn_type1.py:520-636)
  4. flpr_taps              kernel x monomial taps per h, axis  (convolution.py:412-512)
     flpr_conv (x d)        separable banded convolution passes (convolution.py:904-951)
  5. flpr_max_moment        global regularisation key           (regression.py:217, 243-267)
  6. flpr_adjoint_interp    probe adjoint on the raw grid       (regression.py:1256-1280)
     flpr_solve             per-node solve + probe dot products (regression.py:162-277)
  7. flpr_interp_rss        residual sums per bandwidth         (api.py:686-723)
  8. flpr_interp_eval       final yhat at the data points       (api.py:760-786)

The host computes only O(d + dh) scalars: the geometry from the x
statistics, the bandwidth filter and band limits, and, after one
device-to-host copy of dh x (ncol + nprobe) sums, the Hutchinson DOF, GCV and
the 1-SE selection.

Layout on the device
--------------------
* spread grid: (C, G_0, ..., G_{d-1}) float64, C = [1] + y columns + probes,
  G_k = 2 E_k + 1 cells covering signed cell positions [-E_k, E_k] (the
  reference's circular FFT grid unwrapped; it never wraps on the data support).
* stage-k moment slabs (after convolving axes k..d-1): for every bandwidth and
  every partial power tuple beta, shape (C_beta, G_0..G_{k-1}, N_k..N_{d-1}),
  where C_beta = C when sum(beta) <= order (the T channels need it) and 1
  (only the ones channel, for S) otherwise.
* stage-0 slabs: (C_alpha, nodes) per bandwidth and monomial alpha, nodes
  row-major over (N_0, ..., N_{d-1}).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

from . import _lib
from ._lib import FlprGeom, call, i64_array

# Spread tile edge per dimension.  In 3-D a tile of 4^3 cells keeps a dense
# cluster's points on several blocks: on 1e7 Gaussian-mixture points (128^3)
# edge 4 fits in 40 ms against 49 ms for edge 8 (3: 46 ms, 5: 41 ms); in 2-D
# edges 4 and 16 are a few percent slower than 8 on 1e6 uniform points.
TILE = {1: 16, 2: 8, 3: 4}
# |u| beyond which the Gaussian taps are dropped: exp(-u^2/2) u^4 relative to the
# peak is 1.7e-14 at u = 9 (float32 path) and 1e-27 at u = 12 (float64 path).
BAND_CUT = {False: 9.0, True: 12.0}
# intermediate (stage >= 1) memory budget per bandwidth chunk, bytes
STAGE_BUDGET = 6 << 30
# Bandwidths narrower than this many grid cells (smallest over the axes) are
# convolved in float64 even in the float32 form: below it the local moment
# matrices of order 1 / 2 are so ill-conditioned that float32 moments move
# GCV by more than 1e-5 (the reference's own complex64 pipeline is off by
# 4e-4 at 0.23 cells and 3.5e-3 at 0.18 cells for order 2).  Order 0 solves a
# scalar and never needs it.
FP64_BELOW_CELLS = {0: 0.0, 1: 0.25, 2: 0.4}


def stream_handle() -> int:
    """The current HIP stream of the current device, as an integer handle
    (torch's raw-stream query: torch.cuda.current_stream() builds a Python
    Stream object, several microseconds per call)."""
    return torch._C._cuda_getCurrentRawStream(torch._C._cuda_getDevice())


class KernelTimer:
    """Optional HIP-event timing of named C-ABI launches on the current stream.

    Disabled by default.  When enabled, every timed call records an event pair
    around the launch; ``collect()`` synchronises and returns per-name lists of
    milliseconds together with the algorithmic work attached to each launch.
    ``start(pool)`` creates ``pool`` event pairs up front (each recorded once
    so that the HIP event exists), so a timed launch only records two events
    and creates none; past the pool, events are created as needed."""

    def __init__(self):
        self.enabled = False
        self._pending = []
        self._pool = []

    def start(self, pool: int = 0):
        self._pool = []
        for _ in range(pool):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            b.record()
            self._pool.append((a, b))
        if pool:
            torch.cuda.synchronize()
        self._pool.reverse()
        self.enabled = True
        self._pending = []

    def stop(self):
        self.enabled = False

    def timed(self, name: str, work, fn, *args):
        """work: the launch's algorithmic work, or a callable returning it
        (evaluated in collect(), outside the timed region)."""
        if not self.enabled:
            return fn(*args)
        if self._pool:
            a, b = self._pool.pop()
        else:
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
        a.record()
        out = fn(*args)
        b.record()
        self._pending.append((name, work, a, b))
        return out

    def collect(self):
        torch.cuda.synchronize()
        res: Dict[str, list] = {}
        for name, work, a, b in self._pending:
            w = work() if callable(work) else work
            res.setdefault(name, []).append((a.elapsed_time(b), w))
        self._pending = []
        self._pool = []
        return res


TIMER = KernelTimer()


def ptr(t: Optional[torch.Tensor]):
    return None if t is None else t.data_ptr()


# Block partials per reduced value of the fixed-order cross-block sums (probe
# dots of flpr_solve, quadratic forms of flpr_quadform): their workspace
# holds one ticket per group and this many partials per value; a launch with
# more blocks than that uses fewer (more nodes per thread).  The workspace
# must be zero before its first launch; every launch leaves it zero.
SUM_BLOCKS = 256


def sum_part_len(values: int, blocks: int = SUM_BLOCKS) -> int:
    """Doubles of a fixed-order sum workspace for up to `values` reduced
    values (groups x values per group) of up to `blocks` partials each,
    tickets included."""
    values = max(int(values), 1)
    return (values + 1) // 2 + values * max(1, min(int(blocks), SUM_BLOCKS))


def zeroed(dev, *specs):
    """Several zero-filled device buffers from one allocation and one fill:
    specs are (count, dtype) pairs (count 0 gives None); every buffer starts
    on a 256-byte boundary."""
    offs, sizes, total = [], [], 0
    for (cnt, dt) in specs:
        offs.append(total)
        nb = int(cnt) * dt.itemsize
        sizes.append(nb)
        total += (nb + 255) // 256 * 256
    arena = torch.zeros((max(total, 256),), dtype=torch.uint8, device=dev)
    out = []
    for (cnt, dt), o, nb in zip(specs, offs, sizes):
        out.append(arena[o:o + nb].view(dt) if nb else None)
    return out


_DEVTAB: Dict[tuple, torch.Tensor] = {}


def device_table(a: np.ndarray, dtype: torch.dtype, dev) -> torch.Tensor:
    """A small constant host table on the device, uploaded once per content
    (bandwidths, tap bands, their sums); callers must not write to it."""
    a = np.ascontiguousarray(a)
    key = (a.dtype.str, a.shape, a.tobytes(), dtype, str(dev))
    t = _DEVTAB.get(key)
    if t is None:
        if len(_DEVTAB) >= 256:
            _DEVTAB.pop(next(iter(_DEVTAB)))
        t = upload(a, dtype, dev)
        _DEVTAB[key] = t
    return t


def upload(a: np.ndarray, dtype: torch.dtype, dev) -> torch.Tensor:
    """Host array -> device tensor through pinned memory, without a stream
    synchronisation (a pageable copy would drain the stream first)."""
    t = torch.from_numpy(np.ascontiguousarray(a)).to(dtype)
    if dev.type != "cuda":
        return t.to(dev)
    return t.pin_memory().to(dev, non_blocking=True)


# ---------------------------------------------------------------------------
# polynomial bookkeeping (design_matrix_optimized.py:30-82, convolution.py:488-497)
# ---------------------------------------------------------------------------

def poly_powers(d: int, order: int) -> List[Tuple[int, ...]]:
    """Monomial exponents in the reference's term order."""
    if order == 0:
        return [(0,) * d]
    out: List[Tuple[int, ...]] = []
    for tot in range(order + 1):
        if d == 1:
            out.append((tot,))
        elif d == 2:
            for px in range(tot, -1, -1):
                out.append((px, tot - px))
        else:
            for px in range(tot, -1, -1):
                for py in range(tot - px, -1, -1):
                    out.append((px, py, tot - px - py))
    return out


def pair_list(nt: int) -> List[Tuple[int, int]]:
    return [(i, j) for i in range(nt) for j in range(i, nt)]


# ---------------------------------------------------------------------------
# static plan: pass structure, slab offsets, conv job tables
# ---------------------------------------------------------------------------

@dataclass
class Pass:
    axis: int
    mode: int            # 0 ROW (contiguous axis), 1 COL (strided axis)
    K: int               # input axis length (G_axis)
    nout: int            # output axis length (N_axis)
    inner: int           # contiguous extent after the axis (mode 1)
    base: int            # tap index = i + base - s
    max_ro: int
    jobs: torch.Tensor   # (nj, 8) int64 on device
    in_stage: int
    out_stage: int
    jobs_host: np.ndarray = None
    max_np: int = 1      # most tap rows (powers) of one job: flpr_conv's max_powers


def spread_work(n: int, d: int, msp: int, C: int, cells: int) -> Tuple[float, float]:
    """Algorithmic work of one flpr_spread call: (flops, bytes).  flops: one
    multiply-add per (point, footprint cell, channel), 2 n (2 msp + 1)^d C;
    bytes: the sorted point data read once (2 d weights and C - 1 values, or
    C values without the ones channel, 8 bytes each) and the grid written
    once (C cells of 8 bytes)."""
    taps = (2 * msp + 1) ** d
    return (2.0 * n * taps * C, 8.0 * n * (2 * d + C) + 8.0 * C * cells)


def conv_flops(ps: "Pass", bands: np.ndarray) -> float:
    """Algorithmic flops of one convolution pass: 2 per (output element, tap
    inside the band whose input cell lies on the grid)."""
    jh = ps.jobs_host
    i = np.arange(ps.nout)
    total = 0.0
    cache: Dict[int, float] = {}
    for rec in jh:
        band, ro = int(rec[3]), int(rec[4])
        per = cache.get(band)
        if per is None:
            lo, hi = int(bands[2 * band]), int(bands[2 * band + 1])
            t0 = np.maximum(lo, i + ps.base - (ps.K - 1))
            t1 = np.minimum(hi, i + ps.base)
            per = float(np.sum(np.maximum(t1 - t0 + 1, 0)))
            cache[band] = per
        total += 2.0 * per * ro * (ps.inner if ps.mode == 1 else 1) * max(1, int(rec[5]))
    return total


CONV_NP = 3   # tap rows (powers) one job of the float32 form applies to its input


def power_runs(entries, limit: int):
    """Jobs of one input slab: entries (a, out_off, ro) in increasing power a,
    cut into runs of consecutive powers with the same ro and a constant,
    positive out_off step, at most limit long (flpr_conv's multi-power jobs:
    the input is loaded and split once for every power of a run).  Returns
    [(a0, out_off0, ro, np, out_stride)]."""
    runs = []
    i = 0
    while i < len(entries):
        a0, o0, r0 = entries[i]
        n, stride = 1, 0
        while n < limit and i + n < len(entries):
            a1, o1, r1 = entries[i + n]
            step = o1 - entries[i + n - 1][1]
            if a1 != a0 + n or r1 != r0 or step <= 0 or (n > 1 and step != stride):
                break
            stride = step
            n += 1
        runs.append((a0, o0, r0, n, stride))
        i += n
    return runs


class Plan:
    """Everything that depends only on (d, order, N, channels, dh chunk, precision)."""

    def __init__(self, d: int, order: int, N: Tuple[int, ...], L: Tuple[int, ...], msp: int,
                 C: int, dh: int, fp64: bool, device):
        self.d, self.order, self.N, self.L, self.msp = d, order, tuple(N), tuple(L), msp
        self.C, self.dh, self.fp64, self.device = C, dh, fp64, device
        self.E = tuple(n // 2 + 1 + msp for n in N)
        self.G = tuple(2 * e + 1 for e in self.E)
        self.nodes = int(np.prod(N))
        self.npow = 2 * order + 1
        self.nmax = int(max(N))
        self.powers = poly_powers(d, order)
        self.nt = len(self.powers)
        # stage slabs: stage k holds partial tuples over axes k..d-1
        self.stage_slabs: List[Dict[Tuple[int, ...], Tuple[int, int]]] = [None] * (d + 1)
        self.stage_per_h: List[int] = [0] * (d + 1)
        for k in range(d, -1, -1):
            if k == d:
                tuples = [()]
            else:
                tuples = []
                for beta in self.stage_slabs[k + 1]:
                    for a in range(2 * order - sum(beta) + 1):
                        tuples.append((a,) + beta)
            off = 0
            slabs = {}
            for beta in tuples:
                cb = C if sum(beta) <= order else 1
                size = cb * self._slab_cells(k)
                slabs[beta] = (off, cb)
                off += size
            self.stage_slabs[k] = slabs
            self.stage_per_h[k] = off
        # chunking of bandwidths for the intermediate stages
        inter = max([self.stage_per_h[k] for k in range(1, d)] + [0])
        esz = 8 if fp64 else 4
        self.chunk = dh if inter == 0 else max(1, min(dh, STAGE_BUDGET // max(1, 2 * inter * esz)))
        self.passes: List[List[Pass]] = []  # one list per chunk start
        self.chunks = [(h0, min(dh, h0 + self.chunk)) for h0 in range(0, dh, self.chunk)]
        self.pass_tables = [self._build_passes(h0, h1) for (h0, h1) in self.chunks]
        # stage-0 offsets for the solve
        s0 = self.stage_slabs[0]
        self.s_off = []
        for i, j in pair_list(self.nt):
            alpha = tuple(self.powers[i][k] + self.powers[j][k] for k in range(d))
            self.s_off.append(s0[alpha][0])
        self.t_off = [s0[p][0] for p in self.powers]
        self.diag_off = []
        for p in self.powers:
            self.diag_off.append(s0[tuple(2 * v for v in p)][0])

    def _slab_cells(self, k: int) -> int:
        cells = 1
        for j in range(self.d):
            cells *= self.G[j] if j < k else self.N[j]
        return cells

    def _build_passes(self, h0: int, h1: int) -> List[Pass]:
        d = self.d
        out: List[Pass] = []
        for k in range(d, 0, -1):
            axis = k - 1
            mode = 0 if axis == d - 1 else 1
            K = self.G[axis]
            nout = self.N[axis]
            inner = 1
            for j in range(axis + 1, d):
                inner *= self.N[j]
            prefix = 1
            for j in range(axis):
                prefix *= self.G[j]
            L, N = self.L[axis], self.N[axis]
            qout0 = -(-(L - N) // 2)
            pad_before = (L - N) // 2 + 1
            base = (qout0 - pad_before) + self.E[axis]
            jobs = []
            max_ro, max_np = 1, 1
            in_per_h = self.stage_per_h[k]
            out_per_h = self.stage_per_h[k - 1]
            limit = 1 if self.fp64 else CONV_NP
            for hl in range(h1 - h0):
                hg = h0 + hl
                # stage d (spread grid) has no bandwidth axis; stage 0 is
                # indexed by the global bandwidth; other stages by chunk-local h
                in_h = 0 if k == d else hl
                out_h = hg if k - 1 == 0 else hl
                for beta, (ioff, _icb) in self.stage_slabs[k].items():
                    ents = []
                    for a in range(2 * self.order - sum(beta) + 1):
                        ooff, ocb = self.stage_slabs[k - 1][(a,) + beta]
                        ents.append((a, out_h * out_per_h + ooff, ocb * prefix))
                    for (a0, ooff, ro, npw, ostride) in power_runs(ents, limit):
                        max_ro, max_np = max(max_ro, ro), max(max_np, npw)
                        tap_off = ((hg * d + axis) * self.npow + a0) * self.nmax
                        jobs.append([in_h * in_per_h + ioff, ooff, tap_off, hg * d + axis, ro,
                                     npw, ostride, 0])
            jh = np.asarray(jobs, dtype=np.int64)
            jt = torch.tensor(jh, device=self.device)
            out.append(Pass(axis, mode, K, nout, inner, base, max_ro, jt, k, k - 1, jh, max_np))
        return out

    def y_pass_tables(self, ny: int):
        """Passes that convolve only ny new value channels (a second fit of the
        same x): input = the ny centred channels of their own spread grid,
        intermediate stages hold only the tuples with |beta| <= order (ny
        channels each, compact), and the final T moments are written into this
        plan's stage-0 slabs at channels 1..ny.  Returns (per-stage compact
        sizes per bandwidth, one pass list per chunk); memoised per ny."""
        memo = self.__dict__.setdefault("_ypass", {})
        if ny in memo:
            return memo[ny]
        d = self.d
        yslabs: List[Dict[Tuple[int, ...], int]] = [None] * (d + 1)
        yper = [0] * (d + 1)
        yslabs[d] = {(): 0}
        for k in range(d - 1, 0, -1):
            off, slabs = 0, {}
            for beta in yslabs[k + 1]:
                for a in range(self.order - sum(beta) + 1):
                    slabs[(a,) + beta] = off
                    off += ny * self._slab_cells(k)
            yslabs[k], yper[k] = slabs, off
        nodes = self.nodes
        tables = []
        for (h0, h1) in self.chunks:
            out: List[Pass] = []
            for k in range(d, 0, -1):
                axis = k - 1
                mode = 0 if axis == d - 1 else 1
                K = self.G[axis]
                nout = self.N[axis]
                inner = int(np.prod(self.N[axis + 1:])) if axis + 1 < d else 1
                prefix = int(np.prod(self.G[:axis])) if axis > 0 else 1
                L, N = self.L[axis], self.N[axis]
                base = (-(-(L - N) // 2) - ((L - N) // 2 + 1)) + self.E[axis]
                ro = ny * prefix
                jobs = []
                max_np = 1
                limit = 1 if self.fp64 else CONV_NP
                for hl in range(h1 - h0):
                    hg = h0 + hl
                    for beta, ioff in yslabs[k].items():
                        in_off = 0 if k == d else hl * yper[k] + ioff
                        ents = []
                        for a in range(self.order - sum(beta) + 1):
                            nb = (a,) + beta
                            if k - 1 == 0:
                                out_off = hg * self.stage_per_h[0] + self.stage_slabs[0][nb][0] \
                                    + nodes
                            else:
                                out_off = hl * yper[k - 1] + yslabs[k - 1][nb]
                            ents.append((a, out_off, ro))
                        for (a0, out_off, _ro, npw, ostride) in power_runs(ents, limit):
                            max_np = max(max_np, npw)
                            tap_off = ((hg * d + axis) * self.npow + a0) * self.nmax
                            jobs.append([in_off, out_off, tap_off, hg * d + axis, ro, npw,
                                         ostride, 0])
                jh = np.asarray(jobs, dtype=np.int64)
                jt = torch.tensor(jh, device=self.device)
                out.append(Pass(axis, mode, K, nout, inner, base, ro, jt, k, k - 1, jh, max_np))
            tables.append(out)
        memo[ny] = (yper, tables)
        return memo[ny]


_PLANS: Dict[tuple, Plan] = {}

# MI355X hands consecutive workgroups of a launch to its 8 XCDs in turn; each
# XCD has its own L2
XCDS = 8
CONV_LT = 64                       # flpr_conv's Toeplitz tile extent
CONV_KC = {False: 32, True: 16}    # its K chunk, float32 / float64
CONV_LD = {False: 128, True: 64}   # its dense tile extent, float32 / float64
SCHED_MIN_TILES = 4096             # launches with fewer (tile, power) pairs keep the
                                   # plain (tiles, jobs) grid
_SCHED: Dict[tuple, tuple] = {}


def conv_schedule(ps: Pass, fp64: bool, bands: np.ndarray):
    """Block table of one flpr_conv launch (flpr_conv_tiles): every tile of
    every job once, no idle blocks (the plain grid launches the largest job's
    tile count for every job).  The work of a tile is its number of K chunks
    (+1 for the epilogue).  ROW passes (one input grid shared by every job):
    tiles heaviest first, so that the long tiles do not form the tail.  COL
    passes: the tiles of the jobs that read the same input slab stay together
    and go to one XCD (their input is then fetched into one L2), groups placed
    heaviest first on the least loaded XCD (a group heavier than an XCD's
    share is cut into contiguous pieces), the XCD lists interleaved so that
    block b runs on XCD b mod 8.  (Measured on the 2-D order-2 workload with
    1e6 points: ROW 0.273 -> 0.266 ms, COL 0.636 -> 0.544 ms.)  Memoised per pass and band table.  Returns
    (device int32 table, number of blocks), or (None, 0) for launches too
    small to gain from it."""
    key = (id(ps), bool(fp64), bands.tobytes())
    hit = _SCHED.get(key)
    if hit is not None and hit[0] is ps:
        return hit[1], hit[2]
    kc = CONV_KC[bool(fp64)]
    jh = ps.jobs_host
    nj = jh.shape[0]
    ntt = -(-ps.nout // CONV_LT)
    ro = jh[:, 4].astype(np.int64)
    tiles = np.zeros(nj, dtype=np.int64)
    stride = _lib.C.c_int32(0)
    cnt = _lib.C.c_int32(0)
    for r in np.unique(ro):
        call("flpr_conv_tiles", ps.mode, int(bool(fp64)), int(r), ps.max_ro, ps.nout, ps.inner,
             _lib.C.addressof(cnt), _lib.C.addressof(stride))
        tiles[ro == r] = int(cnt.value)
    ntd = int(stride.value)
    total_tiles = int(tiles.sum())
    if int((tiles * np.maximum(jh[:, 5], 1)).sum()) < SCHED_MIN_TILES:
        _SCHED[key] = (ps, None, 0)
        return None, 0
    # work of output tile tt of job j: K chunks between the band's first and
    # last input (the kernel's s_begin / s_end), +1
    bnd = bands.reshape(-1, 2)[jh[:, 3]]
    i0 = np.arange(ntt) * CONV_LT
    sb = np.maximum(i0[None, :] + ps.base - bnd[:, 1:2], 0)
    se = np.minimum(i0[None, :] + CONV_LT - 1 + ps.base - bnd[:, 0:1], ps.K - 1)
    tw = np.where(se >= sb, (se - sb) // kc + 2, 1)                  # (nj, ntt)
    # every tile of every job, vectorised over jobs: tile k of job j is local
    # index k in [0, tiles[j])
    cnts = tiles.astype(np.int64)
    jb = np.repeat(np.arange(nj), cnts)
    k = np.arange(int(cnts.sum()), dtype=np.int64) - np.repeat(np.cumsum(cnts) - cnts, cnts)
    if ps.mode == 0:
        nrow = np.repeat(-(-ro // CONV_LD[bool(fp64)]), cnts)
        tt, td = np.divmod(k, nrow)
        loc = td + ntd * tt
    else:
        o, rem = np.divmod(k, ntt * ntd)
        tt, td = np.divmod(rem, ntd)
        loc = td + ntd * (tt + ntt * o)
    wk = tw[jb, tt].astype(np.float64) * np.maximum(jh[jb, 5], 1)
    assert len(loc) == total_tiles
    if ps.mode == 0:
        o = np.argsort(-wk, kind="stable")
        table = np.stack([jb[o], loc[o]], axis=1).astype(np.int32)
        dev_t = torch.tensor(table.reshape(-1), device=ps.jobs.device)
        if len(_SCHED) >= 4096:
            _SCHED.pop(next(iter(_SCHED)))
        _SCHED[key] = (ps, dev_t, len(table))
        return dev_t, len(table)
    # groups: jobs reading the same input slab, in job order
    _, gid = np.unique(jh[jb, 0], return_inverse=True)
    order = np.argsort(gid, kind="stable")
    loc, wk, jb, gid = loc[order], wk[order], jb[order], gid[order]
    share = wk.sum() / XCDS
    starts = np.flatnonzero(np.r_[True, gid[1:] != gid[:-1]])
    ends = np.r_[starts[1:], len(gid)]
    cw = np.r_[0.0, np.cumsum(wk)]
    units = []
    for a, b in zip(starts, ends):
        gw = cw[b] - cw[a]
        pieces = max(1, int(math.ceil(gw / share)))
        step = -(-(b - a) // pieces)
        for p0 in range(a, b, step):
            p1 = min(p0 + step, b)
            units.append((cw[p1] - cw[p0], p0, p1))
    units.sort(key=lambda u: -u[0])
    load = np.zeros(XCDS)
    lists: List[list] = [[] for _ in range(XCDS)]
    for (uw, p0, p1) in units:
        x = int(np.argmin(load))
        load[x] += uw
        lists[x].append(np.arange(p0, p1))
    per = [np.concatenate(l_) if l_ else np.zeros(0, dtype=np.int64) for l_ in lists]
    nmax = max(len(v) for v in per)
    table = np.full((nmax * XCDS, 2), -1, dtype=np.int32)
    table[:, 1] = 0
    for x in range(XCDS):
        if len(per[x]):
            rows = x + XCDS * np.arange(len(per[x]))
            table[rows, 0] = jb[per[x]]
            table[rows, 1] = loc[per[x]]
    dev_t = torch.tensor(table.reshape(-1), device=ps.jobs.device)
    if len(_SCHED) >= 4096:
        _SCHED.pop(next(iter(_SCHED)))
    _SCHED[key] = (ps, dev_t, nmax * XCDS)
    return dev_t, nmax * XCDS


def split_strips(tap32: torch.Tensor, bands: torch.Tensor, dh: int, d: int, npow: int,
                 nmax: int, s) -> Tuple[torch.Tensor, int]:
    """The float32 form's tap operand (flpr_conv_strips) for all dh bandwidth
    rows; returns (table, int16 elements per bandwidth row)."""
    V = int(_lib.load().flpr_conv_strip_len(nmax))
    if V < 1:
        raise _lib.FlprError("flpr_conv_strip_len failed")
    ntap = dh * d * npow
    strips = torch.empty((ntap * 24 * V,), dtype=torch.int16, device=tap32.device)
    call("flpr_conv_strips", ptr(tap32), ptr(bands), ntap, npow, nmax, ptr(strips), s)
    return strips, d * npow * 24 * V


def launch_conv(ps: Pass, gf: bool, src: int, dst: int, tap_ptr: int, nmax: int, band_ptr: int,
                bands_g: np.ndarray, s) -> None:
    """One flpr_conv pass with its block schedule (conv_schedule).  tap_ptr:
    the float64 taps (gf) or the split strip table (float32 form)."""
    table, nblocks = conv_schedule(ps, gf, bands_g)   # (None, 0): plain grid
    TIMER.timed("flpr_conv64" if gf else "flpr_conv",
                (lambda ps=ps, bg=bands_g: conv_flops(ps, bg)),
                call, "flpr_conv", ps.mode, int(gf), src, dst, tap_ptr, nmax, ptr(ps.jobs),
                band_ptr,
                int(ps.jobs.shape[0]), ps.max_ro, ps.K, ps.nout, ps.inner, ps.base, ps.max_np,
                ptr(table), nblocks, s)


def get_plan(d, order, N, L, msp, C, dh, fp64, device) -> Plan:
    key = (d, order, tuple(N), tuple(L), msp, C, dh, fp64, str(device))
    p = _PLANS.get(key)
    if p is None:
        p = Plan(d, order, N, L, msp, C, dh, fp64, device)
        _PLANS[key] = p
    return p


# ---------------------------------------------------------------------------
# geometry from the x statistics
# ---------------------------------------------------------------------------

@dataclass
class Stats:
    xmin: np.ndarray
    xmax: np.ndarray
    mean: np.ndarray
    std: np.ndarray
    bad_x: int = 0
    bad_y: int = 0
    ysq: Optional[np.ndarray] = None
    ysq_dev: Optional[torch.Tensor] = None   # the same sums, left on the device


def x_stats(x: torch.Tensor, y: Optional[torch.Tensor] = None) -> Stats:
    """Per-dimension min, max, mean and std(ddof=1) of x (n, d) float64 on the device,
    counts of non-finite values in x and y, and sum_i y_ic^2 per column of y."""
    return x_stats_read(x_stats_launch(x, y))


def x_stats_launch(x: torch.Tensor, y: Optional[torch.Tensor] = None):
    """Queues the statistics pass of x_stats; x_stats_read waits for it, so
    host work placed between the two overlaps the device."""
    n, d = x.shape
    ny = 0 if y is None else y.shape[1]
    m = 5 * d + 2 + ny
    buf = torch.empty(513 * m, dtype=torch.float64, device=x.device)
    part, out = buf[:512 * m], buf[512 * m:]
    call("flpr_xstats", ptr(x), n, d, ptr(y), ny, ptr(part), ptr(out), stream_handle())
    return out, d


def x_stats_read(pending) -> Stats:
    """The Stats of a pass queued by x_stats_launch (one host read)."""
    out, d = pending
    o = out.cpu().numpy()
    w = o[:5 * d].reshape(d, 5)
    cnt = w[:, 4]
    std = np.sqrt(w[:, 3] / np.maximum(cnt - 1, 1))
    std = np.where(std == 0, 1.0, std)
    return Stats(w[:, 0].copy(), w[:, 1].copy(), w[:, 2].copy(), std, int(o[5 * d]),
                 int(o[5 * d + 1]), o[5 * d + 2:].copy(), out[5 * d + 2:])


def y_stats(y: torch.Tensor) -> torch.Tensor:
    """sum_i y_ic^2 per column and the count of non-finite values of y (n, ny)
    float64, left on the device: out[0] = 0, out[1] = non-finite count,
    out[2 + c] = sum of squares of column c (flpr_xstats with d = 0)."""
    n, ny = y.shape
    part = torch.empty(512 * (2 + ny), dtype=torch.float64, device=y.device)
    out = torch.empty(2 + ny, dtype=torch.float64, device=y.device)
    call("flpr_xstats", None, n, 0, ptr(y), ny, ptr(part), ptr(out), stream_handle())
    return out


def padded_sizes(N: Tuple[int, ...], flag_power2: bool = True) -> Tuple[int, ...]:
    """L = 2^ceil(log2(2N-1)) (convolution.py:390-395)."""
    if not flag_power2:
        raise NotImplementedError("flag_power2=False (L = N + 2) is not supported on the GPU path")
    return tuple(int(2 ** math.ceil(math.log2(2 * n - 1))) for n in N)


@dataclass
class Geometry:
    geom: FlprGeom
    N: Tuple[int, ...]
    L: Tuple[int, ...]
    msp: int                         # spread footprint half-width (cells)
    xmin_s: np.ndarray
    xmax_s: np.ndarray
    axes_raw: List[np.ndarray]
    axes_scaled: List[np.ndarray]
    mean: np.ndarray
    std: np.ndarray
    centre: np.ndarray
    margin: int = 0                  # grid margin beyond the nodes (>= msp)


def build_geometry(st: Stats, N: Tuple[int, ...], accuracy: int, xrange=None,
                   flag_power2: bool = True, kde: bool = False,
                   margin: Optional[int] = None, tile: Optional[int] = None,
                   item_points: int = 0, spread_f32: bool = False) -> Geometry:
    """Knot-mapping constants exactly as the reference computes them in float64
    (model.py:58-72, convolution.py:169-199, nufftn_type1.py:140-176,
    kernel.py:97-156); the extreme points give the min / max of every
    monotone float64 step, so no extra pass over the data is needed.

    margin: cells of spread grid kept beyond the nodes on each side (at least
    the footprint half-width); two geometries with the same margin share one
    grid, so channels spread at different accuracies can be convolved
    together.  tile (spread tile edge in cells, default TILE[d]) and
    item_points (points per spread / adjoint work item, 0 = the library
    default) change how the kernels cut the work, not what they compute
    beyond the order of float64 additions."""
    d = len(N)
    if accuracy > 6:
        raise ValueError("the spread grid geometry takes accuracy <= 6; accuracy > 6 spreads "
                         "onto fine_geometry's oversampled grid")
    msp = int(accuracy)
    if msp < 0:
        raise ValueError("accuracy must be >= 0")
    margin = msp if margin is None else max(msp, int(margin))
    L = padded_sizes(N, flag_power2)
    g = FlprGeom()
    g.d = d
    g.msp = msp
    centre = np.zeros(d)
    xmin_s = np.zeros(d)
    xmax_s = np.zeros(d)
    axes_raw, axes_scaled = [], []
    for k in range(d):
        n_k, L_k = int(N[k]), int(L[k])
        mean, std = float(st.mean[k]), float(st.std[k])
        tmin = (float(st.xmin[k]) - mean) / std
        tmax = (float(st.xmax[k]) - mean) / std
        c = (tmax + tmin) / 2
        lo, hi = tmin - c, tmax - c
        scale = hi - lo
        if scale == 0:
            scale = 1.0
        idx = np.linspace(-0.5, 0.5 - 1 / L_k, L_k)
        q0 = (L_k - n_k) // 2
        q1 = q0 + n_k - 1
        idx0 = float(idx[q0])
        ks = float(idx[q1] - idx[q0])
        kmin_pt = (idx0 + 0.5) - 0.5
        kmax_pt = ((ks + idx0) + 0.5) - 0.5
        kmin = (kmin_pt + kmax_pt) * 0.5 - 0.5
        shift = -0.5 - kmin
        tau = (math.pi * msp) / (L_k * L_k * 1 * (1 - 0.5))
        E = n_k // 2 + 1 + margin
        G = 2 * E + 1
        tile_k = TILE[d] if tile is None else int(tile)
        g.n_grid[k] = n_k
        g.L[k] = L_k
        g.E[k] = E
        g.G[k] = G
        g.tile[k] = tile_k
        g.ntile[k] = -(-G // tile_k)
        g.hx[k] = 2 * np.pi / L_k
        # accuracy 0 bins every point to its nearest cell (nufftn_type1.py:442-516):
        # footprint weight exp(0) = 1
        g.inv4tau[k] = 1.0 / (4 * tau) if msp > 0 else 0.0
        g.mean[k] = mean
        g.std_[k] = std
        g.centre[k] = c
        g.xmin_s[k] = lo
        g.scale_s[k] = scale
        g.xmax_s[k] = hi
        g.ks[k] = ks
        g.idx0[k] = idx0
        g.shift[k] = shift
        ax_s = np.linspace(lo, hi, n_k)
        if kde:
            # cv_fastkde maps the scaled axis back to raw units (api.py:1036-1057)
            ax_r = (ax_s + c) * std + mean
            g0, g1 = float(ax_r[0]), float(ax_r[-1])
        else:
            if xrange is None:
                g0, g1 = float(st.xmin[k]), float(st.xmax[k])
            else:
                g0, g1 = float(xrange[k][0]), float(xrange[k][1])
            ax_r = np.linspace(g0, g1, n_k)
        g.graw0[k] = g0
        g.graw1[k] = g1
        centre[k], xmin_s[k], xmax_s[k] = c, lo, hi
        axes_raw.append(ax_r)
        axes_scaled.append(ax_s)
    g.item_points = int(item_points)
    g.spread_f32 = int(bool(spread_f32))
    return Geometry(g, tuple(N), L, msp, xmin_s, xmax_s, axes_raw, axes_scaled,
                    st.mean.copy(), st.std.copy(), centre, margin)


MAX_FINE_ACCURACY = 12


def fine_margin(N: Tuple[int, ...]) -> int:
    """Grid margin that makes E >= N on every axis: the deconvolved grid of
    accuracy > 6 is not compactly supported, so the unwrapped grid must hold
    every cell a node's kernel window reaches."""
    return max(int(n_) - int(n_) // 2 - 1 for n_ in N)


def fine_geometry(geo: Geometry, accuracy: int) -> FlprGeom:
    """Oversampled spread geometry of accuracy > 6 (kernel.py:97-156): R = 2
    (accuracy <= 11) or 3, Mr = R L, Msp = accuracy,
    tau = pi Msp / (L^2 R (R - 1/2)); the knot mapping is the coarse one."""
    acc = int(accuracy)
    d = len(geo.N)
    if acc > MAX_FINE_ACCURACY:
        raise NotImplementedError(f"accuracy > {MAX_FINE_ACCURACY} is not supported on the GPU")
    R = 2 if acc <= 11 else 3
    f = FlprGeom.from_buffer_copy(geo.geom)
    f.msp = acc
    for k in range(d):
        n_k, L_k = int(geo.N[k]), int(geo.L[k])
        tau = (math.pi * acc) / (L_k * L_k * R * (R - 0.5))
        Mr = R * L_k
        E = R * (n_k // 2 + 1) + acc
        f.L[k] = Mr
        f.hx[k] = 2 * np.pi / Mr
        f.inv4tau[k] = 1.0 / (4 * tau)
        f.E[k] = E
        f.G[k] = 2 * E + 1
        f.tile[k] = TILE[d]
        f.ntile[k] = -(-(2 * E + 1) // TILE[d])
    return f


def spread_fine(geo: Geometry, accuracy: int, x: torch.Tensor, n: int,
                vals_y: Optional[torch.Tensor], ny: int, y_pairs: bool,
                probes: Optional[torch.Tensor], K: int, with_ones: bool,
                grid64: torch.Tensor, sums: torch.Tensor, c0: int, s) -> None:
    """Spread channels (ones?, y, probes) onto the oversampled grid of accuracy
    > 6 (nufftn_type1.py:520-636) and resample them, deconvolved, into
    channels c0.. of the coarse unwrapped grid (nufftn_type1.py:693-770); sums
    receives their totals (flpr_resample)."""
    d = len(geo.N)
    dev = x.device
    gf = fine_geometry(geo, accuracy)
    gfp = _lib.C.byref(gf)
    gp = _lib.C.byref(geo.geom)
    f64, i32 = torch.float64, torch.int32
    ntiles = int(np.prod([gf.ntile[k] for k in range(d)]))
    nkeys = ntiles * int(np.prod([gf.tile[k] for k in range(d)]))
    key = torch.empty((n,), dtype=i32, device=dev)
    cnt = torch.zeros((2 * nkeys,), dtype=i32, device=dev)
    key_start = torch.empty((nkeys + 1,), dtype=i32, device=dev)
    order_idx = torch.empty((n,), dtype=i32, device=dev)
    call("flpr_points_prepare", ptr(x), n, gfp, ptr(key), ptr(cnt[:nkeys]), None, s)
    call("flpr_points_bin", ptr(key), n, nkeys, ptr(cnt[:nkeys]), ptr(key_start),
         ptr(cnt[nkeys:]), None, ptr(order_idx), s)
    del key
    nv = ny + K
    wab = torch.empty((2 * d * n,), dtype=f64, device=dev)
    vals = torch.empty((max(nv, 1) * n,), dtype=f64, device=dev)
    call("flpr_points_pack", gfp, ptr(x), ptr(order_idx), n, ptr(vals_y) if ny else None, ny,
         ptr(probes) if K else None, K, ptr(wab), ptr(vals), None, None, s)
    wo = 1 if with_ones else 0
    Cf = wo + nv
    cells_f = int(np.prod([gf.G[k] for k in range(d)]))
    sums_f = torch.zeros((Cf,), dtype=f64, device=dev)
    if d == 3:
        # the 3-D oversampled footprint (up to 25^3 cells) exceeds the spread's
        # LDS window: rows added into the zeroed grid (flpr_spread_wide)
        fine = torch.zeros((Cf * cells_f,), dtype=f64, device=dev)
        call("flpr_spread_wide", gfp, ptr(key_start), ptr(wab), ptr(vals), n, wo, nv, ptr(fine),
             ptr(sums_f), s)
    else:
        fcells = int(np.prod([gf.tile[k] + 2 * gf.msp for k in range(d)]))
        part = torch.empty((ntiles * Cf * fcells,), dtype=f64, device=dev)
        fine = torch.empty((Cf * cells_f,), dtype=f64, device=dev)
        call("flpr_spread", gfp, ptr(key_start), ptr(wab), ptr(vals), n, wo, nv, ptr(part),
             ptr(fine), ptr(sums_f), s)
        del part
    del wab, vals
    cells = int(np.prod([geo.geom.G[k] for k in range(d)]))
    groups = []
    if wo:
        groups.append((0, 1, 0))
    if ny:
        groups.append((wo, ny, 1 if y_pairs else 0))
    if K:
        groups.append((wo + ny, K, 0))
    need = 0
    for (_f0, nch, pairs) in groups:
        w = _lib.C.c_int64(0)
        call("flpr_resample_work", gfp, gp, nch, pairs, _lib.C.addressof(w))
        need = max(need, int(w.value))
    work = torch.empty((need,), dtype=f64, device=dev)
    for (f0, nch, pairs) in groups:
        call("flpr_resample", gfp, gp, fine.data_ptr() + f0 * cells_f * 8, nch, pairs, ptr(work),
             grid64.data_ptr() + (c0 + f0) * cells * 8, sums.data_ptr() + (c0 + f0) * 8, s)


def periodize(geo: Geometry, grid: torch.Tensor, c0: int, nch: int) -> None:
    """Fold channels c0..c0+nch-1 of a freshly spread unwrapped grid onto the
    reference's L-periodic grid when some axis is narrower than the spread
    footprint (G > L, small N; flpr_grid_periodize); otherwise nothing."""
    g = geo.geom
    d = len(geo.N)
    if all(g.G[k] <= g.L[k] for k in range(d)):
        return
    cells = int(np.prod([g.G[k] for k in range(d)]))
    work = torch.empty((nch * int(np.prod([g.L[k] for k in range(d)])),), dtype=torch.float64,
                       device=grid.device)
    call("flpr_grid_periodize", _lib.C.byref(g), grid.data_ptr() + c0 * cells * 8, nch,
         ptr(work), stream_handle())


def filter_bandwidths(geo: Geometry, h: np.ndarray, n: int):
    """Too-small bandwidths (kernel sum < 10 eps) are dropped; all bad -> Silverman
    (convolution.py:542-605).  The d-dimensional kernel sum is the product of
    1-D sums because the Gaussian factorises."""
    d = len(geo.N)
    sums = np.ones(h.shape[0])
    for k in range(d):
        u = geo.axes_scaled[k][None, :] / h[:, k:k + 1]
        sums *= np.sum(np.exp(-0.5 * u * u), axis=1)
    sums = sums / math.sqrt(2 * math.pi) / np.prod(h, axis=1)
    ihbad = np.abs(sums) < np.finfo(np.float64).eps * 10
    warn = None
    if np.any(ihbad) and not np.all(ihbad):
        warn = f"{int(np.sum(ihbad))} bandwidth(s) are too small for this grid and will be removed"
        h = h[~ihbad]
    elif np.all(ihbad):
        hs = (4.0 / (d + 2) / n) ** (1.0 / (d + 4))
        warn = f"All bandwidths too small. Using Silverman's rule: h = {hs:.4f}"
        h = np.array([[hs] * d])
    return h, ihbad, warn


def conv_pad(geo: Geometry) -> int:
    """Elements flpr_conv may read past a job's input: 256 K + 64 inner + 512 for
    the longest axis and the largest inner extent."""
    d = len(geo.N)
    inner = int(np.prod(geo.N[1:])) if d > 1 else 1
    return 256 * int(max(geo.geom.G[k] for k in range(d))) + 64 * inner + 512


def precision_groups(geo: Geometry, h: np.ndarray, order: int, fp64: bool):
    """Contiguous runs [a, b) of bandwidths with their convolution precision
    (True = float64).  The float32 form sends bandwidths narrower than
    FP64_BELOW_CELLS[order] grid cells to float64."""
    dh = h.shape[0]
    if fp64:
        return [(0, dh, True)]
    d = len(geo.N)
    dx = np.array([abs(geo.axes_scaled[k][1] - geo.axes_scaled[k][0]) if geo.N[k] > 1 else 1.0
                   for k in range(d)])
    cells = np.min(h / dx[None, :], axis=1)
    wide = cells >= FP64_BELOW_CELLS.get(order, 0.0)
    groups = []
    a = 0
    for b in range(1, dh + 1):
        if b == dh or wide[b] != wide[a]:
            groups.append((a, b, not bool(wide[a])))
            a = b
    return groups


def tap_bands(geo: Geometry, h: np.ndarray, fp64) -> np.ndarray:
    """[t_lo, t_hi] per (bandwidth, axis): taps with |u| <= cut, widened by one
    on each side and clipped to the axis; (0, -1) when no tap is inside.  fp64
    is one flag or one flag per bandwidth."""
    d = len(geo.N)
    dh = h.shape[0]
    flags = np.broadcast_to(np.asarray(fp64, dtype=bool), (dh,))
    cut = np.where(flags, BAND_CUT[True], BAND_CUT[False])
    out = np.zeros((dh, d, 2), dtype=np.int32)
    for k in range(d):
        ax = geo.axes_scaled[k]
        inside = np.abs(ax[None, :] / h[:, k:k + 1]) <= cut[:, None]
        anyin = inside.any(axis=1)
        first = np.argmax(inside, axis=1)
        last = len(ax) - 1 - np.argmax(inside[:, ::-1], axis=1)
        out[:, k, 0] = np.where(anyin, np.maximum(0, first - 1), 0)
        out[:, k, 1] = np.where(anyin, np.minimum(len(ax) - 1, last + 1), -1)
    return out


# Host-side tables are pure functions of (statistics, grid, bandwidths, order,
# precision); repeated fits on the same x (mean and variance passes, several y
# columns) reuse them instead of recomputing them between device launches.
_GEOMS: Dict[tuple, Geometry] = {}
_BTABS: Dict[tuple, tuple] = {}
_MEMO_MAX = 32


def _memo_put(table: dict, key, value):
    if len(table) >= _MEMO_MAX:
        table.pop(next(iter(table)))
    table[key] = value
    return value


def geometry_for(st: Stats, N, accuracy: int, xrange=None, flag_power2: bool = True,
                 kde: bool = False, margin: Optional[int] = None,
                 tile: Optional[int] = None, item_points: int = 0, spread_f32: bool = False):
    """build_geometry, memoised on the exact statistics and grid options.
    Returns (geometry, key)."""
    xr = None if xrange is None else tuple(tuple(float(v) for v in r) for r in xrange)
    mg = int(accuracy) if margin is None else max(int(accuracy), int(margin))
    key = (st.xmin.tobytes(), st.xmax.tobytes(), st.mean.tobytes(), st.std.tobytes(),
           tuple(int(v) for v in N), int(accuracy), xr, bool(flag_power2), bool(kde), mg,
           tile, int(item_points), bool(spread_f32))
    geo = _GEOMS.get(key)
    if geo is None:
        geo = _memo_put(_GEOMS, key, build_geometry(st, N, accuracy, xrange, flag_power2, kde,
                                                    mg, tile, item_points, spread_f32))
    return geo, key


def bandwidth_tables(geo: Geometry, gkey, h: np.ndarray, n: int, order: int, fp64: bool):
    """Filtered bandwidths, warning, precision groups and tap bands, memoised."""
    h = np.ascontiguousarray(h, dtype=float)
    key = (gkey, h.shape, h.tobytes(), int(n), int(order), bool(fp64))
    hit = _BTABS.get(key)
    if hit is None:
        hf, _ihbad, warn = filter_bandwidths(geo, h, n)
        groups = precision_groups(geo, hf, order, fp64)
        flags = np.zeros(hf.shape[0], dtype=bool)
        for (a, b, gf) in groups:
            flags[a:b] = gf
        band_np = tap_bands(geo, hf, flags).reshape(-1)
        hit = _memo_put(_BTABS, key, (hf, warn, groups, band_np))
    return hit


# ---------------------------------------------------------------------------
# the fit
# ---------------------------------------------------------------------------

@dataclass
class DeviceFit:
    h: np.ndarray                    # filtered bandwidths (dh, d)
    m_y: torch.Tensor                # (dh, ny, nodes) float64
    rss: Optional[np.ndarray]        # (dh, ncol_out)
    dot: Optional[np.ndarray]        # (dh, K)
    s0_all: "S00Rows"                # S_00 per bandwidth (rows on demand)
    geo: Geometry
    plan: Plan
    warn: Optional[str]
    red: Optional[torch.Tensor] = None   # device (dh, nval) dots then (dh, ny) quad forms
    key: Optional[torch.Tensor] = None   # global regularisation key (key_only)
    order_idx: Optional[torch.Tensor] = None  # (n) int32 points grouped by spread cell


class S00Rows:
    """S_00 per bandwidth, (dh, nodes) float64, read from the stage-0 slabs of
    the precision groups: a row (the selected bandwidth) is converted when it
    is indexed, all rows by all() (the density path)."""

    def __init__(self, stage0s, per_h0: int, off0: int, nodes: int, eps_add: float):
        self._parts = [(a, b, st0) for (a, b, _gf, _p, st0) in stage0s]
        self._per_h0, self._off0, self._nodes, self._eps = per_h0, off0, nodes, eps_add

    def _rows(self, st0, n):
        return st0.view(n, self._per_h0)[:, self._off0:self._off0 + self._nodes]

    def __getitem__(self, i: int) -> torch.Tensor:
        for (a, b, st0) in self._parts:
            if a <= i < b:
                r = self._rows(st0, b - a)[i - a].to(torch.float64)
                return r + self._eps if self._eps else r
        raise IndexError(i)

    def all(self) -> torch.Tensor:
        rows = [self._rows(st0, b - a).to(torch.float64) for (a, b, st0) in self._parts]
        out = rows[0] if len(rows) == 1 else torch.cat(rows, dim=0)
        return out + self._eps if self._eps else out


def adjoint_workspace(g, d: int, ntiles: int, nch: int) -> int:
    """Doubles of flpr_adjoint's workspace: per tile and channel, its LDS
    window (tile + 5 cells per axis) or the 2^d corner sums of its cells."""
    wcells = int(np.prod([g.tile[k] + 5 for k in range(d)]))
    corners = int(np.prod([g.tile[k] for k in range(d)])) << d
    return ntiles * nch * max(wcells, corners)


def run_grid(x: torch.Tensor, vals_y: Optional[torch.Tensor], ny: int,
             probes: Optional[torch.Tensor], K: int, h: np.ndarray, order: int,
             N: Tuple[int, ...], accuracy: int, fp64: bool, regularization: float,
             xrange=None, flag_power2: bool = True, want_rss: bool = False,
             stats: Optional[Stats] = None, with_ones_eps: bool = True,
             debug: Optional[dict] = None, readback: bool = True,
             moments_only: bool = False, kde: bool = False,
             keep_all: bool = False, accuracy_y: Optional[int] = None,
             y_pairs: bool = False, share_out: Optional[dict] = None,
             key_only: bool = False, key_in: Optional[torch.Tensor] = None,
             mixed: bool = False, select_h=None) -> DeviceFit:
    """Spread, convolve and solve for all bandwidths, plus the probe dot products and
    the residual sums.

    x: (n, d) float64 on the device; vals_y: (n, ny) float64 or None (complex y is
    passed as interleaved (re, im) columns); probes: (n, K) float64 or None.
    want_rss: also return sum_i (y_ic - yhat_ic(h))^2 per bandwidth and real column.
    stats: x statistics from x_stats(x, vals_y) when the caller already has them.
    with_ones_eps adds float64 eps to S for order 0 (regression.py:127; the
    calc_dof=False path of the reference omits it).
    debug: when a dict is given, the intermediate device buffers are stored in it.
    readback: copy the dots / residual sums to the host (rss, dot); when False
    they stay on the device in DeviceFit.red and nothing waits for the device.
    moments_only: stop after the convolution (S_00 per bandwidth in s0_all, the
    density path of cv_fastkde); kde selects cv_fastkde's raw grid axes.
    keep_all: return the grid solutions of every channel (y columns and probes,
    DeviceFit.m_y is (dh, ny + K, nodes)) instead of the adjoint sums, for
    callers that interpolate them at the points themselves (cubic).
    accuracy_y: spread accuracy of the y columns when it differs from the one
    of the ones channel and the probes (the reference's cv_fastlpr with DOF
    estimation spreads S and the probes at the sample-size accuracy and only
    y at the option's, api.py:569-576 / regression.py:86-87, 756-758).
    Accuracies above 6 spread onto the oversampled grid and deconvolve
    (spread_fine); y_pairs says that the y columns are (re, im) pairs of
    complex responses, which that deconvolution must keep together.
    key_only: with moments_only, also reduce the global regularisation key
    over every bandwidth of h (DeviceFit.key, a device uint64) and stop;
    select_h: a callable (geo, grid64, cells, sums, h) -> None or a callable
    returning indices (called after the list-wide regularisation pass is
    queued),
    called once the spread grid is complete and before it is centred (so
    channels 0 and 1 are the raw sums of the weights and of y): with
    indices, only those bandwidths of the filtered list h are convolved
    with every channel and solved, and the regularisation key is reduced
    over the whole list from the ones channel alone (the learned bracket).
    mixed: the precision form (float32 spread with compensated summation,
    float32 solves with one float64 refinement step, flpr_solve_refined);
    fp64 must be False.
    key_in: use this key instead of reducing one over h (a fit on a subset of
    bandwidths whose lambda must be the full list's, regression.py:243-267).
    share_out: when a dict is given (linear interpolation paths only), it
    receives the x-side state of the fit (sorted points, the ones channel,
    the moment slabs, lambda, the interpolation Gram matrices) for
    run_grid_y.
    """
    n, d = x.shape
    dev = x.device
    s = stream_handle()
    st = stats if stats is not None else x_stats(x, vals_y)
    acc_s = int(accuracy)
    acc_y = acc_s if (accuracy_y is None or ny == 0) else int(accuracy_y)
    fine_s, fine_y = acc_s > 6, acc_y > 6
    if d == 3 and (fine_s or fine_y):
        # the oversampled grid is deconvolved along three axes: the float32
        # convolution of that grid cancels to 6.6e-5 of yhat (12^3 nodes, order
        # 1, accuracy 9), so these fits convolve in float64 in every form
        fp64, mixed = True, False
    split_acc = acc_y != acc_s
    if fine_s or fine_y:
        margin = max(fine_margin(N), min(acc_s, 6), min(acc_y, 6))
    else:
        margin = max(acc_s, acc_y) if split_acc else None
    # with accuracy > 6 the ones channel is not spread onto this grid: msp 0
    geo, gkey = geometry_for(st, N, 0 if fine_s else acc_s, xrange, flag_power2, kde, margin,
                             spread_f32=mixed and not fp64)
    refine = mixed and not fp64 and not fine_s
    nval = ny + K
    C = 1 + nval
    g = geo.geom
    gp = _lib.C.byref(g)
    f64, i32 = torch.float64, torch.int32

    # bandwidth tables (host, memoised) and the sizes of the small zeroed buffers
    h, warn, groups, band_np = bandwidth_tables(geo, gkey, np.asarray(h, dtype=float), n, order,
                                                fp64)
    dh = h.shape[0]
    nodes_all = int(np.prod(geo.N))
    rcells = int(np.prod([n_ - 1 for n_ in geo.N]))
    nw = (1 << d) * ((1 << d) + 1) // 2
    need_q = (K > 0 or want_rss) and not keep_all and not moments_only
    nq_out = dh * ny if (want_rss and ny > 0 and not keep_all) else 0

    # --- points: knots, cell keys, counting sort, sorted SoA data
    ntiles = int(np.prod([g.ntile[k] for k in range(d)]))
    nkeys = ntiles * int(np.prod([g.tile[k] for k in range(d)]))
    nsum = sum_part_len(dh * nval) if need_q else 0
    cnt, sums, key_max, red, q, gram, sump = zeroed(
        dev, (2 * nkeys, i32), (C, f64), (1, torch.int64),
        (dh * nval + nq_out if need_q else 0, f64), (nval * nodes_all if need_q else 0, f64),
        (nw * rcells if (need_q and want_rss) else 0, f64), (nsum, f64))
    # cell keys, their starts, the sorted order and the ranks: one allocation
    ints = torch.empty((3 * n + nkeys + 1,), dtype=i32, device=dev)
    key, order_idx, rank = ints[:n], ints[n:2 * n], ints[2 * n:3 * n]
    key_start = ints[3 * n:]
    key_count, key_cursor = cnt[:nkeys], cnt[nkeys:]
    call("flpr_points_prepare", ptr(x), n, gp, ptr(key), ptr(key_count), ptr(rank), s)
    call("flpr_points_bin", ptr(key), n, nkeys, ptr(key_count), ptr(key_start), ptr(key_cursor),
         ptr(rank), ptr(order_idx), s)
    del rank
    wab = torch.empty((2 * d * n,), dtype=f64, device=dev)
    vals = torch.empty((max(nval, 1) * n,), dtype=f64, device=dev)
    rawi = torch.empty((d * n,), dtype=i32, device=dev)
    rawf = torch.empty((d * n,), dtype=f64, device=dev)
    call("flpr_points_pack", gp, ptr(x), ptr(order_idx), n, ptr(vals_y), ny, ptr(probes), K,
         ptr(wab), ptr(vals), ptr(rawi), ptr(rawf), s)

    # --- spread (float64 accumulation), centring, float32 copy
    cells = int(np.prod([g.G[k] for k in range(d)]))
    # partial windows of the widest footprint spread onto this grid
    wmsp = max(geo.msp, acc_y if (split_acc and not fine_y) else 0)
    fcells = int(np.prod([g.tile[k] + 2 * wmsp for k in range(d)]))
    part = (torch.empty((ntiles * C * fcells,), dtype=f64, device=dev) if not fine_s
            else None)
    # convolution inputs are read in whole tiles past each job's end (the
    # values there are discarded or zeroed): pad them
    pad = conv_pad(geo)
    grid64 = torch.empty((C * cells + pad,), dtype=f64, device=dev)
    if fine_s:
        # every channel through the oversampled grid (acc_y == acc_s here)
        spread_fine(geo, acc_s, x, n, vals_y, ny, y_pairs, probes, K, True, grid64, sums, 0, s)
    else:
        TIMER.timed("flpr_spread", spread_work(n, d, geo.msp, C, cells), call, "flpr_spread", gp,
                    ptr(key_start), ptr(wab), ptr(vals), n, 1, nval, ptr(part), ptr(grid64),
                    ptr(sums), s)
    if split_acc and not fine_s and not fine_y:
        # the y columns again, with the footprint and width of their own
        # accuracy, over the channels the first spread wrote
        geo_y, _ = geometry_for(st, N, int(accuracy_y), xrange, flag_power2, kde, margin,
                                spread_f32=mixed and not fp64)
        gy = _lib.C.byref(geo_y.geom)
        wab_y = torch.empty((2 * d * n,), dtype=f64, device=dev)
        vals_yp = torch.empty((ny * n,), dtype=f64, device=dev)
        call("flpr_points_pack", gy, ptr(x), ptr(order_idx), n, ptr(vals_y), ny, None, 0,
             ptr(wab_y), ptr(vals_yp), None, None, s)
        sums[1:1 + ny].zero_()
        call("flpr_spread", gy, ptr(key_start), ptr(wab_y), ptr(vals_yp), n, 0, ny, ptr(part),
             ptr(grid64) + cells * 8, ptr(sums) + 8, s)
        del wab_y, vals_yp
    if not fine_s:
        periodize(geo, grid64, 0, C)
    if fine_y and not fine_s:
        spread_fine(geo, acc_y, x, n, vals_y, ny, y_pairs, None, 0, False, grid64, sums, 1, s)
    del part
    if select_h is not None:
        pick = select_h(geo, grid64, cells, sums, h)
        if pick is not None:
            # lambda of the whole list first: the device computes it while
            # the host reads the selector's result
            regularisation_key(geo, gkey, h, order, n, fp64, grid64, cells, key_max, s)
            idx = pick()
            key_in = key_max
            h, _w, groups, band_np = bandwidth_tables(geo, gkey, h[np.asarray(idx)], n, order,
                                                      fp64)
            dh = h.shape[0]
    any32 = any(not gf for (_a, _b, gf) in groups)
    shift = torch.empty((C,), dtype=f64, device=dev)
    grid32 = torch.empty((C * cells + pad,), dtype=torch.float32, device=dev) if any32 else None
    call("flpr_grid_center", ptr(grid64), C, cells, ptr(sums), ptr(shift), ptr(grid32), s)

    # --- taps and bands (uploaded from pinned host memory without a stream sync)
    plan0 = get_plan(d, order, geo.N, geo.L, geo.margin, C, groups[0][1] - groups[0][0],
                     groups[0][2], dev)
    npow, nmax = plan0.npow, plan0.nmax
    per_tap = d * npow * nmax
    h_dev = device_table(h, f64, dev)
    tap64 = torch.empty((dh * per_tap,), dtype=f64, device=dev)
    tap32 = torch.empty((dh * per_tap,), dtype=torch.float32, device=dev) if any32 else None
    call("flpr_taps", gp, ptr(h_dev), dh, npow, nmax, float(n), ptr(tap64), ptr(tap32), s)
    bands = device_table(band_np, i32, dev)
    strips, strip_row = (split_strips(tap32, bands, dh, d, npow, nmax, s) if any32
                         else (None, 0))

    # --- global regularisation key (regression.py:217, 243-267)
    if order == 0:
        lam_scale = float(regularization)
        eps_add = float(np.finfo(float).eps) if with_ones_eps else 0.0
    else:
        lam_scale, eps_add = 1e-6, 0.0

    # --- convolution passes per precision group (bandwidth chunks share the
    # intermediate buffers); every group feeds the one global key
    stage0s = []
    for (a, b, gf) in groups:
        plan = get_plan(d, order, geo.N, geo.L, geo.margin, C, b - a, gf, dev)
        fdt = f64 if gf else torch.float32
        src_grid = grid64 if gf else grid32
        tap_ptr = (tap64.data_ptr() + a * per_tap * 8) if gf else \
            (strips.data_ptr() + a * strip_row * 2)
        band_ptr = bands.data_ptr() + a * d * 2 * 4
        bands_g = band_np[a * d * 2:b * d * 2]
        stage0 = torch.empty(((b - a) * plan.stage_per_h[0],), dtype=fdt, device=dev)
        inter = {k: torch.empty((plan.chunk * plan.stage_per_h[k] + pad,), dtype=fdt, device=dev)
                 for k in range(1, d)}
        for ci in range(len(plan.chunks)):
            for ps in plan.pass_tables[ci]:
                src = src_grid if ps.in_stage == d else inter[ps.in_stage]
                dst = stage0 if ps.out_stage == 0 else inter[ps.out_stage]
                launch_conv(ps, gf, ptr(src), ptr(dst), tap_ptr, nmax, band_ptr, bands_g, s)
        del inter
        if key_in is None and (not moments_only or key_only):
            if order == 0:
                offs, absval = [plan.s_off[0]], 0
            else:
                offs, absval = plan.diag_off, 1
            call("flpr_max_moment", int(gf), ptr(stage0), plan.stage_per_h[0], b - a,
                 i64_array(offs), len(offs), plan.nodes, absval, ptr(key_max), s)
        stage0s.append((a, b, gf, plan, stage0))

    if key_in is not None:
        key_max = key_in
    nodes = plan0.nodes
    per_h0 = plan0.stage_per_h[0]
    off0 = plan0.s_off[0]
    if key_only:
        return DeviceFit(h, None, None, None, None, geo, plan0, warn, None, key=key_max)
    if moments_only:
        s0_all = S00Rows(stage0s, per_h0, off0, nodes, eps_add if order == 0 else 0.0)
        return DeviceFit(h, None, None, None, s0_all, geo, plan0, warn, None,
                         order_idx=order_idx)

    # --- adjoint sums, solve (+ dots), residual quadratic form
    # red: one float64 readback buffer, probe / y dots (dh, nval) then
    # quadratic forms (dh, ny); q, gram: adjoint sums (all zeroed above)
    dot = red[:dh * nval] if need_q else None
    if need_q:
        apart = torch.empty((adjoint_workspace(g, d, ntiles, max(nval, nw)),), dtype=f64,
                            device=dev)
        call("flpr_adjoint", gp, ptr(key_start), ptr(rawi), ptr(rawf), ptr(vals), n, nval,
             ptr(apart), ptr(q), ptr(gram), s)
        del apart
    nout = nval if keep_all else ny
    nyc = max(nout, 1)
    m_y = torch.empty((dh * nyc * nodes,), dtype=f64, device=dev)
    for (a, b, gf, plan, stage0) in stage0s:
        m_ptr = (m_y.data_ptr() + a * nyc * nodes * 8) if nout > 0 else None
        d_ptr = (dot.data_ptr() + a * nval * 8) if dot is not None else None
        if refine and not gf:
            call("flpr_solve_refined", ptr(stage0), per_h0, b - a, nodes, order, plan.nt,
                 i64_array(plan.s_off), i64_array(plan.t_off), ptr(key_max), lam_scale, eps_add,
                 ptr(shift), nval, nout, m_ptr, ptr(q), d_ptr, ptr(sump), nsum, s)
        else:
            call("flpr_solve", int(gf), ptr(stage0), per_h0, b - a, nodes, order, plan.nt,
                 int(fine_s), i64_array(plan.s_off), i64_array(plan.t_off), ptr(key_max),
                 lam_scale, eps_add, ptr(shift), nval, nout, m_ptr, ptr(q), d_ptr, ptr(sump),
                 nsum, s)
    quad = None
    if nq_out:
        quad = red[dh * nval:]
        call("flpr_quadform", gp, ptr(m_y), ny, dh, ptr(gram), ptr(quad), ptr(sump), nsum, s)
    s0_all = S00Rows(stage0s, per_h0, off0, nodes, eps_add if order == 0 else 0.0)
    if debug is not None:
        debug.update(grid64=grid64, stage0=stage0s[0][4], stage0s=stage0s, key=key_max, q=q,
                     dot=dot, m_y=m_y, plan=plan0, geo=geo, h=h, order_idx=order_idx,
                     key_start=key_start, lam_scale=lam_scale, eps_add=eps_add, shift=shift,
                     gram=gram, quad=quad, groups=groups)
    if share_out is not None and not keep_all and gram is not None:
        share_out.update(
            geo=geo, x=x, n=n, order=order, order_idx=order_idx, key_start=key_start, wab=wab,
            rawi=rawi, rawf=rawf, grid64=grid64, sums=sums, cells=cells, pad=pad,
            stage0s=stage0s, key_max=key_max, gram=gram, s0_all=s0_all, h=h, warn=warn,
            tap64=tap64, strips=strips, strip_row=strip_row, per_tap=per_tap, nmax=nmax,
            bands=bands, band_np=band_np,
            lam_scale=lam_scale, eps_add=eps_add, fine_s=fine_s, acc_s=acc_s, acc_y=acc_y,
            split_acc=split_acc, margin=margin, st=st, xrange=xrange, flag_power2=flag_power2,
            kde=kde, N=tuple(N), plan0=plan0, any32=any32, mixed=mixed, refine=refine)
    if not readback:
        return DeviceFit(h, m_y.view(dh, nyc, nodes), None, None, s0_all, geo, plan0, warn, red)
    red_h = None if red is None else red.cpu().numpy()
    dot_h = None if red_h is None else red_h[:dh * nval].reshape(dh, nval)
    rss_h = None
    if quad is not None:
        quad_h = red_h[dh * nval:].reshape(dh, ny)
        ysq = st.ysq if (st.ysq is not None and len(st.ysq) == ny) else None
        if ysq is None:
            ysq = (vals_y * vals_y).sum(dim=0).cpu().numpy()
        rss_h = ysq[None, :] - 2.0 * dot_h[:, :ny] + quad_h
    probe_dot = None if (dot_h is None or K == 0) else dot_h[:, ny:]
    return DeviceFit(h, m_y.view(dh, nyc, nodes), rss_h, probe_dot, s0_all, geo, plan0, warn, red)


def regularisation_key(geo: Geometry, gkey, h: np.ndarray, order: int, n: int, fp64: bool,
                       grid64: torch.Tensor, cells: int, key_max: torch.Tensor, s) -> None:
    """The global regularisation key of the bandwidth list h (regression.py:
    217, 243-267: the largest |S_kk| over nodes, diagonal moments and every
    bandwidth, or the largest S for order 0) from channel 0 (the ones channel)
    of a spread grid alone: a one-channel convolution of every bandwidth.
    Reduces into key_max (zeroed by the caller)."""
    d = len(geo.N)
    dev = grid64.device
    f64, i32 = torch.float64, torch.int32
    hf, _w, groups, band_np = bandwidth_tables(geo, gkey, np.asarray(h, dtype=float), n, order,
                                               fp64)
    dh = hf.shape[0]
    pad = conv_pad(geo)
    any32 = any(not gf for (_a, _b, gf) in groups)
    g32 = None
    if any32:
        g32 = torch.empty((cells + pad,), dtype=torch.float32, device=dev)
        g32[:cells].copy_(grid64[:cells])
    plan0 = get_plan(d, order, geo.N, geo.L, geo.margin, 1, groups[0][1] - groups[0][0],
                     groups[0][2], dev)
    npow, nmax = plan0.npow, plan0.nmax
    per_tap = d * npow * nmax
    h_dev = device_table(hf, f64, dev)
    tap64 = torch.empty((dh * per_tap,), dtype=f64, device=dev)
    tap32 = torch.empty((dh * per_tap,), dtype=torch.float32, device=dev) if any32 else None
    call("flpr_taps", _lib.C.byref(geo.geom), ptr(h_dev), dh, npow, nmax, float(n), ptr(tap64),
         ptr(tap32), s)
    bands = device_table(band_np, i32, dev)
    strips, strip_row = (split_strips(tap32, bands, dh, d, npow, nmax, s) if any32
                         else (None, 0))
    for (a, b, gf) in groups:
        plan = get_plan(d, order, geo.N, geo.L, geo.margin, 1, b - a, gf, dev)
        fdt = f64 if gf else torch.float32
        src_grid = grid64 if gf else g32
        tap_ptr = (tap64.data_ptr() + a * per_tap * 8) if gf else \
            (strips.data_ptr() + a * strip_row * 2)
        band_ptr = bands.data_ptr() + a * d * 2 * 4
        bands_g = band_np[a * d * 2:b * d * 2]
        stage0 = torch.empty(((b - a) * plan.stage_per_h[0] + pad,), dtype=fdt, device=dev)
        inter = {k: torch.empty((plan.chunk * plan.stage_per_h[k] + pad,), dtype=fdt, device=dev)
                 for k in range(1, d)}
        for ci in range(len(plan.chunks)):
            for ps in plan.pass_tables[ci]:
                src = src_grid if ps.in_stage == d else inter[ps.in_stage]
                dst = stage0 if ps.out_stage == 0 else inter[ps.out_stage]
                launch_conv(ps, gf, ptr(src), ptr(dst), tap_ptr, nmax, band_ptr, bands_g, s)
        del inter
        if order == 0:
            offs, absval = [plan.s_off[0]], 0
        else:
            offs, absval = plan.diag_off, 1
        call("flpr_max_moment", int(gf), ptr(stage0), plan.stage_per_h[0], b - a,
             i64_array(offs), len(offs), plan.nodes, absval, ptr(key_max), s)


def run_grid_y(sh: dict, x: torch.Tensor, vals_y: torch.Tensor, ny: int,
               y_pairs: bool = False) -> DeviceFit:
    """A second fit of the same x, bandwidths, order, grid, accuracy and
    precision as the run_grid call that filled sh, for new response columns
    vals_y (n, ny), with the residual sums and without probes.  Everything
    that depends only on x is taken from sh: the sorted points, the ones
    channel of the spread grid, the S moments (and the ones channel's T
    moments for the centring), lambda, the interpolation Gram matrices and
    S_00.  The new channels are spread, centred on the ones channel,
    convolved into the T slots of sh's moment slabs (y-only passes), solved
    and reduced; the DeviceFit is the one run_grid(readback=False) returns,
    with (dh, ny) dots and (dh, ny) quadratic forms in red."""
    assert x is sh["x"] and x.shape[0] == sh["n"]
    n, d = x.shape
    dev = x.device
    s = stream_handle()
    geo = sh["geo"]
    g = geo.geom
    gp = _lib.C.byref(g)
    f64 = torch.float64
    cells, pad = sh["cells"], sh["pad"]
    order = sh["order"]
    h = sh["h"]
    dh = h.shape[0]
    C2 = 1 + ny
    # --- the new columns in sorted order, spread next to the first fit's ones channel
    vals = torch.empty((ny * n,), dtype=f64, device=dev)
    acc_s, acc_y = sh["acc_s"], sh["acc_y"]
    ntiles = int(np.prod([g.ntile[k] for k in range(d)]))
    grid64 = torch.empty((C2 * cells + pad,), dtype=f64, device=dev)
    grid64[:cells].copy_(sh["grid64"][:cells])
    nodes = sh["plan0"].nodes
    nsum = sum_part_len(dh * ny)
    sums, red, q, sump = zeroed(dev, (C2, f64), (2 * dh * ny, f64), (ny * nodes, f64),
                                (nsum, f64))
    sums[:1].copy_(sh["sums"][:1])
    if acc_y > 6:
        call("flpr_points_pack", gp, ptr(x), ptr(sh["order_idx"]), n, ptr(vals_y), ny, None, 0,
             None, ptr(vals), None, None, s)
        spread_fine(geo, acc_y, x, n, vals_y, ny, y_pairs, None, 0, False, grid64, sums, 1, s)
    else:
        if sh["split_acc"]:
            geo_y, _ = geometry_for(sh["st"], sh["N"], acc_y, sh["xrange"], sh["flag_power2"],
                                    sh["kde"], sh["margin"], spread_f32=sh["mixed"])
            gy = _lib.C.byref(geo_y.geom)
            wab = torch.empty((2 * d * n,), dtype=f64, device=dev)
            call("flpr_points_pack", gy, ptr(x), ptr(sh["order_idx"]), n, ptr(vals_y), ny,
                 None, 0, ptr(wab), ptr(vals), None, None, s)
        else:
            gy, wab = gp, sh["wab"]
            call("flpr_points_pack", gp, ptr(x), ptr(sh["order_idx"]), n, ptr(vals_y), ny,
                 None, 0, None, ptr(vals), None, None, s)
        fcells = int(np.prod([g.tile[k] + 2 * acc_y for k in range(d)]))
        part = torch.empty((ntiles * ny * fcells,), dtype=f64, device=dev)
        TIMER.timed("flpr_spread", spread_work(n, d, acc_y, ny, cells), call, "flpr_spread", gy,
                    ptr(sh["key_start"]), ptr(wab), ptr(vals), n, 0, ny, ptr(part),
                    grid64.data_ptr() + cells * 8, sums.data_ptr() + 8, s)
        del part
        periodize(geo, grid64, 1, ny)
    shift = torch.empty((C2,), dtype=f64, device=dev)
    any32 = sh["any32"]
    grid32 = torch.empty((C2 * cells + pad,), dtype=torch.float32, device=dev) if any32 else None
    call("flpr_grid_center", ptr(grid64), C2, cells, ptr(sums), ptr(shift), ptr(grid32), s)
    # --- T moments of the new channels into the first fit's slabs
    per_tap = sh["per_tap"]
    bands = sh["bands"]
    for (a, b, gf, plan, stage0) in sh["stage0s"]:
        fdt = f64 if gf else torch.float32
        src_grid = grid64 if gf else grid32
        esz = 8 if gf else 4
        tap_ptr = (sh["tap64"].data_ptr() + a * per_tap * 8) if gf else \
            (sh["strips"].data_ptr() + a * sh["strip_row"] * 2)
        band_ptr = bands.data_ptr() + a * d * 2 * 4
        bands_g = sh["band_np"][a * d * 2:b * d * 2]
        yper, tables = plan.y_pass_tables(ny)
        inter = {k: torch.empty((plan.chunk * yper[k] + pad,), dtype=fdt, device=dev)
                 for k in range(1, d)}
        for ci in range(len(plan.chunks)):
            for ps in tables[ci]:
                src = (src_grid.data_ptr() + cells * esz) if ps.in_stage == d \
                    else ptr(inter[ps.in_stage])
                dst = stage0 if ps.out_stage == 0 else inter[ps.out_stage]
                launch_conv(ps, gf, src, ptr(dst), tap_ptr, sh["nmax"], band_ptr, bands_g, s)
        del inter
    # --- adjoint sums of the new columns, solve (+ dots), residual quadratic form
    plan0 = sh["plan0"]
    nodes = plan0.nodes
    per_h0 = plan0.stage_per_h[0]
    nw = (1 << d) * ((1 << d) + 1) // 2
    apart = torch.empty((adjoint_workspace(g, d, ntiles, max(ny, nw)),), dtype=f64, device=dev)
    call("flpr_adjoint", gp, ptr(sh["key_start"]), ptr(sh["rawi"]), ptr(sh["rawf"]), ptr(vals), n,
         ny, ptr(apart), ptr(q), None, s)
    del apart
    m_y = torch.empty((dh * ny * nodes,), dtype=f64, device=dev)
    for (a, b, gf, plan, stage0) in sh["stage0s"]:
        if sh["refine"] and not gf:
            call("flpr_solve_refined", ptr(stage0), per_h0, b - a, nodes, order, plan.nt,
                 i64_array(plan.s_off), i64_array(plan.t_off), ptr(sh["key_max"]),
                 sh["lam_scale"], sh["eps_add"], ptr(shift), ny, ny,
                 m_y.data_ptr() + a * ny * nodes * 8, ptr(q), red.data_ptr() + a * ny * 8,
                 ptr(sump), nsum, s)
        else:
            call("flpr_solve", int(gf), ptr(stage0), per_h0, b - a, nodes, order, plan.nt,
                 int(sh["fine_s"]), i64_array(plan.s_off), i64_array(plan.t_off),
                 ptr(sh["key_max"]), sh["lam_scale"], sh["eps_add"], ptr(shift), ny, ny,
                 m_y.data_ptr() + a * ny * nodes * 8, ptr(q), red.data_ptr() + a * ny * 8,
                 ptr(sump), nsum, s)
    call("flpr_quadform", gp, ptr(m_y), ny, dh, ptr(sh["gram"]), red.data_ptr() + dh * ny * 8,
         ptr(sump), nsum, s)
    return DeviceFit(h, m_y.view(dh, ny, nodes), None, None, sh["s0_all"], geo, plan0,
                     sh["warn"], red)


def lcv(geo: Geometry, x: torch.Tensor, dens: torch.Tensor, k0h: np.ndarray,
        order: Optional[torch.Tensor] = None):
    """LCV score and standard error per bandwidth (flpr_lcv): dens (dh, nodes)
    float64 on the device; order (n) int32, the points grouped by spread cell
    (faster density reads), or None for their stored order.  Returns a (2 dh,)
    device tensor [scores, stderrs]."""
    n = x.shape[0]
    dh = dens.shape[0]
    dev = x.device
    k0 = upload(np.asarray(k0h, dtype=float), torch.float64, dev)
    part = torch.empty((1024 * dh * 3,), dtype=torch.float64, device=dev)
    out = torch.empty((2 * dh,), dtype=torch.float64, device=dev)
    call("flpr_lcv", _lib.C.byref(geo.geom), ptr(x), ptr(order) if order is not None else None,
         n, ptr(dens.contiguous()), dh, ptr(k0),
         ptr(part), ptr(out), stream_handle())
    return out


# ---------------------------------------------------------------------------
# cubic interpolation (RegularGridInterpolator 'cubic': not-a-knot cubic
# splines along every axis, api.py:588-594, 651-657)
# ---------------------------------------------------------------------------

def _nak_knots(ax: np.ndarray) -> np.ndarray:
    """Knot vector of scipy's make_interp_spline(k=3) (not-a-knot): N + 4 knots."""
    return np.r_[(ax[0],) * 4, ax[2:-2], (ax[-1],) * 4]


def _cubic_basis(t: np.ndarray, v: np.ndarray, N: int):
    """The four nonzero cubic B-splines at v (de Boor's recurrence), and the
    index of the first; v inside [t[3], t[N]]."""
    l = np.clip(np.searchsorted(t, v, side="right") - 1, 3, N - 1)
    basis = np.zeros((v.size, 4))
    basis[:, 0] = 1.0
    left = np.zeros((v.size, 4))
    right = np.zeros((v.size, 4))
    for j in range(1, 4):
        left[:, j] = v - t[l + 1 - j]
        right[:, j] = t[l + j] - v
        saved = np.zeros(v.size)
        for r in range(j):
            temp = basis[:, r] / (right[:, r + 1] + left[:, j - r])
            basis[:, r] = saved + right[:, r + 1] * temp
            saved = left[:, j - r] * temp
        basis[:, j] = saved
    return basis, l - 3


_CUBIC: Dict[tuple, torch.Tensor] = {}
_CUBIC_HOST: Dict[tuple, np.ndarray] = {}


def cubic_inverse(ax: np.ndarray) -> np.ndarray:
    """Inverse of the not-a-knot collocation matrix A[i, j] = B_j(x_i) of one
    raw axis (float64, N x N): coefficients = A^-1 values.  Memoised on the
    axis' end points and length."""
    key = (float(ax[0]), float(ax[-1]), len(ax))
    hit = _CUBIC_HOST.get(key)
    if hit is None:
        N = len(ax)
        if N < 4:
            raise ValueError("cubic interpolation needs at least 4 grid nodes per dimension")
        t = _nak_knots(ax)
        basis, first = _cubic_basis(t, ax, N)
        A = np.zeros((N, N))
        for r in range(4):
            A[np.arange(N), first + r] += basis[:, r]
        hit = np.linalg.inv(A)
        if len(_CUBIC_HOST) >= 4096:
            _CUBIC_HOST.clear()
        _CUBIC_HOST[key] = hit
    return hit


def cubic_tables(geo: Geometry, dev) -> torch.Tensor:
    """The inverse collocation matrices of every axis, concatenated (A_0^-1,
    A_1^-1, ... row-major), on the device."""
    key = (tuple(float(a[0]) for a in geo.axes_raw), tuple(float(a[-1]) for a in geo.axes_raw),
           tuple(geo.N), str(dev))
    hit = _CUBIC.get(key)
    if hit is None:
        host = np.concatenate([cubic_inverse(ax).reshape(-1) for ax in geo.axes_raw])
        hit = upload(host, torch.float64, dev)
        if len(_CUBIC) >= _MEMO_MAX:
            _CUBIC.clear()
        _CUBIC[key] = hit
    return hit


def spline_coefficients(geo: Geometry, grids: torch.Tensor) -> torch.Tensor:
    """B-spline coefficients of grids (G, nodes) float64: A_k^-1 applied along
    every axis (flpr_batch_spline_coef, one problem)."""
    inv = cubic_tables(geo, grids.device)
    G, nodes = grids.shape
    grids = grids.contiguous()
    coef = torch.empty_like(grids)
    work = torch.empty_like(grids) if len(geo.N) > 1 else coef
    call("flpr_batch_spline_coef", _lib.C.byref(geo.geom), 1, ptr(inv), inv.numel(), ptr(grids),
         G, ptr(work), ptr(coef), stream_handle())
    return coef


def one_item(geo: Geometry, n: int, dev) -> torch.Tensor:
    """A one-problem FlprBatchItem table on the device (rows [0, n))."""
    it = _lib.FlprBatchItem()
    it.g = geo.geom
    it.n = int(n)
    it.pt = 0
    return upload(np.frombuffer(bytes(it), dtype=np.uint8).copy(), torch.uint8, dev)


def point_sums(geo: Geometry, x: torch.Tensor, y: torch.Tensor, ny: int, probes: torch.Tensor,
               K: int, m: torch.Tensor, dh: int, gcv_cubic: bool, dof_cubic: bool,
               red: torch.Tensor) -> None:
    """Residual sums and probe dots at the points for grid solutions m (dh,
    ny + K, nodes) (flpr_batch_point_sums, one problem), written into red in
    flpr_select's layout; cubic parts use the spline of m."""
    check_points(geo, x)
    n = x.shape[0]
    nodes = m.shape[-1]
    coef = None
    if gcv_cubic or dof_cubic:
        coef = spline_coefficients(geo, m.reshape(-1, nodes))
    nval = ny + K
    ngroups = -(-(dh * nval) // 8)
    nblk = max(1, min(64, -(-n // 256)))
    part = torch.empty((ngroups * nblk * 8,), dtype=torch.float64, device=x.device)
    call("flpr_batch_point_sums", _lib.C.byref(geo.geom), ptr(one_item(geo, n, x.device)), 1, n,
         ptr(x), ptr(y.contiguous()), ny, ptr(probes), K, ptr(m.contiguous()), ptr(coef), dh,
         int(gcv_cubic), int(dof_cubic), ptr(part), ptr(red), red.numel(), stream_handle())


def check_points(geo: Geometry, x: torch.Tensor) -> None:
    """Query points must be (n, d) for a d-dimensional grid: the interpolation
    kernels read d coordinates per point, so any other shape would read past
    the end of the buffer."""
    d = len(geo.N)
    if x.ndim != 2 or x.shape[1] != d:
        raise ValueError(f"points must have shape (n, {d}) for a {d}-dimensional grid, "
                         f"got {tuple(x.shape)}")


def spline_eval(geo: Geometry, x: torch.Tensor, coef: torch.Tensor) -> torch.Tensor:
    """(n, G) values of the cubic splines with coefficients coef (G, nodes) at x (n, d)."""
    check_points(geo, x)
    n = x.shape[0]
    G = coef.shape[0]
    out = torch.empty((n, G), dtype=torch.float64, device=x.device)
    call("flpr_spline_eval", _lib.C.byref(geo.geom), ptr(x), n, ptr(coef), G, ptr(out),
         stream_handle())
    return out


AGGREGATIONS = {"first": 0, "average": 1, "sum": 2}


def select(fit: DeviceFit, ysq_dev: torch.Tensor, pp_dev: torch.Tensor, n: int, ny: int,
           K: int, cplx: bool, aggregation: str, dstd: float):
    """Hutchinson DOF, GCV scores and the 1-SE choice on the device (flpr_select).

    Returns (out, sel): out (4 dh + 5) float64 on the device holding scores,
    stderrs, dof, dof stderrs, then selected index, argmin index, number of
    finite scores and two spare slots; sel (1,) int32 on the device."""
    if aggregation not in AGGREGATIONS:
        raise ValueError(f"Unknown aggregation method: {aggregation}. "
                         "Use 'first', 'average', or 'sum'.")
    dh = fit.h.shape[0]
    dev = fit.m_y.device
    hsum = device_table(fit.h.sum(axis=1), torch.float64, dev)
    out = torch.empty((4 * dh + 5,), dtype=torch.float64, device=dev)
    sel = torch.empty((1,), dtype=torch.int32, device=dev)
    call("flpr_select", ptr(fit.red), dh, ny + K, ny, K, ptr(ysq_dev), ptr(pp_dev), ptr(hsum),
         n, int(cplx), AGGREGATIONS[aggregation], float(dstd), ptr(out), ptr(sel),
         stream_handle())
    return out, sel


def fit_output(geo: Geometry, x: torch.Tensor, m_all: torch.Tensor, ncol: int,
               sel: Optional[torch.Tensor], yref: torch.Tensor, variance: bool,
               fac_out: Optional[torch.Tensor] = None):
    """yhat (n, ncol) and residuals at the selected bandwidth (flpr_fit_output);
    variance: also the back-transformed grid (nodes) and the factor 1/d
    (written to fac_out, a 1-element float64 device view, when given).
    Returns (yhat, resid, grid or None, fac tensor or None)."""
    n = x.shape[0]
    dev = x.device
    nodes = int(np.prod(geo.N))
    yhat = torch.empty((n, ncol), dtype=torch.float64, device=dev)
    resid = torch.empty((n, ncol), dtype=torch.float64, device=dev)
    grid = part = None
    fac = fac_out
    if variance:
        grid = torch.empty((nodes,), dtype=torch.float64, device=dev)
        part = torch.empty((1024,), dtype=torch.float64, device=dev)
        if fac is None:
            fac = torch.empty((1,), dtype=torch.float64, device=dev)
    m_all = m_all.contiguous()
    call("flpr_fit_output", _lib.C.byref(geo.geom), ptr(x), n, ptr(m_all), ncol, ptr(sel),
         ptr(yref.contiguous()), int(variance), ptr(yhat), ptr(resid), ptr(grid), ptr(part),
         ptr(fac), stream_handle())
    return yhat, resid, grid, fac


def interp_eval(geo: Geometry, x: torch.Tensor, m: torch.Tensor, ncol: int,
                yvar: Optional[torch.Tensor] = None):
    """Multilinear interpolation of (ncol, nodes) float64 grid values at x (n, d)."""
    check_points(geo, x)
    n = x.shape[0]
    out = torch.empty((n, ncol), dtype=torch.float64, device=x.device)
    es = torch.zeros((1,), dtype=torch.float64, device=x.device) if yvar is not None else None
    call("flpr_interp_eval", _lib.C.byref(geo.geom), ptr(x), n, ptr(m.contiguous()), ncol,
         ptr(out), ptr(yvar), ptr(es), stream_handle())
    return out, es
