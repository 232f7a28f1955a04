"""Benchmark: cv_fastlpr fits/sec on the 2-D order-2 heteroscedastic workload.

Workload (one "step"): the heteroscedastic pair on N = 1e6 points in [0,1]^2,
y = sin(2 pi x1) cos(2 pi x2) + (0.05 + 0.2 x1) N(0,1) (float32 inputs),
grid 256 x 256, 30 log-spaced bandwidths in [0.005, 0.5], order 2, GCV with
10 Hutchinson probes:
    pass 1  cv_fastlpr(x, y)                          (mean)
    pass 2  cv_fastlpr(x, (y - yhat)^2, variance)     (variance)
so one step is two fits; value = fits per second = 2 K / time.  Inputs are
resident in HBM before the timed region; outputs stay on the device.  The
step runs the pair as one cv_fastlpr_pair call, whose second pass reuses the
first pass's probe sums (the DOF depends on x, h and the probes only); every
other stage of both passes, including the first pass's probe channels, is
computed in every step.  The same pair as two separate cv_fastlpr calls is
timed too and reported beside it (two_call_fits_per_s).

The JSON line carries the chosen form (float32 convolution on bf16 MFMA with
exact three-term operand splits, float64 spread and solves), the plain form
(float64 end to end) for comparison, the maximum relative yhat difference
between them, a roofline entry for the dominant kernel (the float32 form's
flpr_conv launches, measured with HIP events on the launch stream during a
second pass of the same K steps: event records add host and queue work to every
launch, so the headline steps run without them) and a CPU baseline (the oracle restatement, single thread, on a
bounded sample: the mean pass with 2 and with 4 of the 30 bandwidths after a
warm-up call, extrapolated linearly to 30, and the same on all the host's
cores).  Further legs, outside the headline value: the spread's roofline
entry (HIP events on the spread launches of that second pass), the precision
form "mixed" (float32 spread with compensated summation, float32 solves with
one float64 refinement step) as one pair call with its yhat difference from
the float64 form, the learned bandwidth bracket (mean pass with
bandwidth_bracket="learned" against the full sweep, and whether both select
the same bandwidth), a cold call (caches cleared: probe generation and every
plan rebuilt), and the batched 512-problem workload as problems per second,
batched against one cv_fastlpr call per problem and against its
captured-graph replay.  The default of 50 timed steps (about 0.1 s) keeps a
short host-side disturbance from dominating the launch-bound step time.
After the warm-up the process's garbage is collected and the surviving
objects are frozen (gc.freeze), outside every timed region: a full
collection over the objects of the imports and of the first calls takes
about 35 ms on the bench host, and would otherwise land in one timed pass or
another at random (+0.7 ms per step in a pass of 50).  A long-running caller
holds those objects in the oldest generation the same way.
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOAD = {"workload": "cv_fastlpr 2-D order-2 heteroscedastic pair (mean + variance pass)",
            "n_points": 1_000_000, "grid": [256, 256], "bandwidths": 30,
            "h_range": [0.005, 0.5], "order": 2, "dof_probes": 10, "inputs": "float32",
            "seed": 0}
PEAK_MFMA_FP32 = 157.3e12
PEAK_MFMA_FP64 = 78.6e12
PEAK_MFMA_BF16 = 2.5e15
PEAK_FP64_VALU = 78.6e12          # MI355X float64 vector FMA peak
PEAK_HBM = 8.0e12
# the float32 form's convolution computes each float32 product as six bf16
# MFMA products (hi*hi, hi*mid, mid*hi, hi*lo, lo*hi, mid*mid)
SPLIT_PRODUCTS = 6
COPY_BYTES = 256 << 20            # calibration copy: 256 MiB read + 256 MiB written
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "flpr_conv_traffic_2d_o2_n1e6.txt")


def make_inputs(n: int, seed: int = 0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, (n, 2)).astype(np.float32)
    s = 0.05 + 0.2 * x[:, 0]
    y = (np.sin(2 * np.pi * x[:, 0]) * np.cos(2 * np.pi * x[:, 1])
         + s * rng.standard_normal(n)).astype(np.float32)
    return x, y


def hlist():
    from fastlpr_amd import get_hlist
    h = get_hlist(30, [0.005, 0.5])[:, 0]
    return np.column_stack([h, h])


def step(cv, xd, yd, hl, prec):
    """The pair as two cv_fastlpr calls (the drop-in usage)."""
    opt = {"order": 2, "N": [256, 256], "precision": prec}
    r1 = cv(xd, yd, hl, opt)
    r2y = (yd - r1.yhat) ** 2
    r2 = cv(xd, r2y, hl, dict(opt, y_type_out="variance"))
    return r1, r2


def step_pair(pair, xd, yd, hl, prec):
    """The pair as one cv_fastlpr_pair call."""
    return pair(xd, yd, hl, {"order": 2, "N": [256, 256], "precision": prec})


def cpu_baseline(x, y, hl, threads: int = 1):
    """Oracle (NumPy restatement of the reference dataflow, float32 'ref' mode),
    mean pass: one untimed call with 1 bandwidth (FFT plans, probe matrix),
    then timed calls with 2 and 4 of the 30 bandwidths; the per-bandwidth
    slope extrapolates to 30.  threads: BLAS / OpenMP threads and scipy.fft
    workers (1 = the reference's effective threading; the host's core count
    = the best this restatement does on all cores).  About 20 s of CPU work."""
    import scipy.fft as sfft
    from threadpoolctl import threadpool_limits
    from oracle import restate as R
    xs, ys = x.astype(np.float64), y.astype(np.float64)
    opt = {"order": 2, "N": [256, 256]}
    times = {}
    with threadpool_limits(threads), sfft.set_workers(threads):
        R.cv_fastlpr(xs, ys, hl[14:15], opt, mode="ref")
        for k in (2, 4):
            sub = hl[13:13 + k]
            t = time.perf_counter()
            R.cv_fastlpr(xs, ys, sub, opt, mode="ref")
            times[k] = time.perf_counter() - t
    slope = max(times[4] - times[2], 0.0) / 2.0
    per_fit = times[2] + (hl.shape[0] - 2) * slope
    return {"value": 1.0 / per_fit, "unit": "fits/s", "cores": threads, "kind": "port",
            "host_cores": len(os.sched_getaffinity(0)),
            "sample": (f"oracle restatement (reference dataflow, float32 mode) of the mean pass "
                       f"on the same 1e6 points and 256x256 grid, {threads} thread(s) (BLAS and "
                       f"scipy.fft workers; the spread is single-threaded NumPy), with 2 and 4 "
                       f"of the 30 bandwidths ({times[2]:.1f} s, {times[4]:.1f} s) after one "
                       f"untimed warm-up call; extrapolated linearly to 30 bandwidths: "
                       f"{per_fit:.1f} s per fit")}


def mixture_3d(n: int, seed: int = 0):
    """The 3-D workload: n points from a 4-component Gaussian mixture in
    [0,1]^3 (means U(0.2, 0.8), isotropic sd U(0.05, 0.15), equal weights,
    points outside the cube drawn again), y = ||x - 0.5||^2 + 0.1 N(0,1)."""
    rng = np.random.default_rng(seed)
    means = rng.uniform(0.2, 0.8, (4, 3))
    sd = rng.uniform(0.05, 0.15, 4)
    comp = rng.integers(0, 4, n)
    x = means[comp] + sd[comp, None] * rng.standard_normal((n, 3))
    bad = np.any((x < 0) | (x > 1), axis=1)
    while bad.any():
        k = int(bad.sum())
        c = rng.integers(0, 4, k)
        x[bad] = means[c] + sd[c, None] * rng.standard_normal((k, 3))
        bad = np.any((x < 0) | (x > 1), axis=1)
    y = np.sum((x - 0.5) ** 2, axis=1) + 0.1 * rng.standard_normal(n)
    return x, y


def mixture_3d_leg(steps: int):
    """3-D local linear regression (GCV) and KDE (LCV) on 1e7 mixture points,
    128^3 grid, 16 bandwidths in [0.01, 0.3], float32 form: one untimed call
    of each, then `steps` timed calls of each.  Returns the JSON record."""
    import torch
    import fastlpr_amd as fl
    n = 10_000_000
    x, y = mixture_3d(n)
    hh = fl.get_hlist(16, [0.01, 0.3])[:, 0]
    hl = np.column_stack([hh, hh, hh])
    dev = torch.device("cuda", 0)
    xd, yd = torch.as_tensor(x, device=dev), torch.as_tensor(y, device=dev)
    opt = {"order": 1, "N": [128, 128, 128]}
    out = {"workload": "3-D local linear (GCV) + KDE (LCV), 1e7 Gaussian-mixture points, "
                       "128^3 grid, 16 bandwidths in [0.01, 0.3]",
           "n_points": n, "grid": [128, 128, 128], "bandwidths": 16, "form": "float32",
           "steps": steps}
    for name, fn in (("regression", lambda: fl.cv_fastlpr(xd, yd, hl, dict(opt))),
                     ("kde", lambda: fl.cv_fastkde(xd, hl, {"N": [128, 128, 128]}))):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t) * 1e3 / steps
        out[name + "_ms_per_fit"] = ms
        out[name + "_fits_per_s"] = 1e3 / ms
        print(f"[bench] 3-D mixture {name}: {ms:.1f} ms per fit", file=sys.stderr, flush=True)
    return out


def calibration_copy():
    """A plain streaming kernel of known size, three times: torch.neg over
    256 MiB of float32 (256 MiB read, 256 MiB written), the reference for the
    FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    a = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        torch.neg(a, out=b)
    torch.cuda.synchronize()
    del a, b


def uneven_problem(i: int):
    """Problem i of the batched uneven workload: default_rng(1000 + i) draws
    n = round(10^U(3, 6)) and the order integers(0, 3); i < 256 is 1-D on 1024
    nodes, else 2-D on 128 x 128; 20 bandwidths log-spaced in [0.02, 0.5];
    every tenth problem has complex y = (a + ib) f(x), a, b ~ N(0, 1);
    f = sin(2 pi x1) (+ cos(2 pi x2) in 2-D) + 0.1 N(0, 1)."""
    rng = np.random.default_rng(1000 + i)
    n = int(round(10 ** rng.uniform(3, 6)))
    order = int(rng.integers(0, 3))
    d = 1 if i < 256 else 2
    x = rng.uniform(0, 1, (n, d))
    f = np.sin(2 * np.pi * x[:, 0]) + (np.cos(2 * np.pi * x[:, 1]) if d == 2 else 0.0)
    f = f + 0.1 * rng.standard_normal(n)
    if i % 10 == 0:
        a, b = rng.standard_normal(2)
        f = (a + 1j * b) * f
    hh = np.logspace(np.log10(0.02), np.log10(0.5), 20)
    return x, f, np.column_stack([hh] * d), {"order": order, "N": 1024 if d == 1 else [128, 128]}


def uneven_batch_leg(steps: int):
    """512 uneven problems: one cv_fastlpr call each (looped) against the
    batched BatchPlan (autotuned launch shapes; the x side and the tuning are
    set up before the timed region, as for a plan reused over many response
    sets) and its captured-graph replay.  Inputs resident on the device."""
    import torch
    import fastlpr_amd as fl
    P = 512
    xs, ys, hs, os_ = [], [], [], []
    for i in range(P):
        x, f, h, o = uneven_problem(i)
        xs.append(torch.as_tensor(x, device="cuda"))
        ys.append(torch.as_tensor(f, device="cuda"))
        hs.append(h)
        os_.append(o)
    rows = int(sum(x.shape[0] for x in xs))

    def looped():
        return [fl.cv_fastlpr(xs[i], ys[i], hs[i], os_[i]) for i in range(P)]

    looped()
    torch.cuda.synchronize()
    t = time.perf_counter()
    ref = looped()
    torch.cuda.synchronize()
    t_loop = time.perf_counter() - t
    t = time.perf_counter()
    plan = fl.BatchPlan(xs, hs, os_, tune=True)
    plan.fit(ys)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t
    best = float("inf")
    for _ in range(max(steps, 1)):
        torch.cuda.synchronize()
        t = time.perf_counter()
        got = plan.fit(ys)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t)
    diff = 0.0
    same_sel = True
    for a, b in zip(got, ref):
        ya, yb = a.yhat, b.yhat
        diff = max(diff, float((ya - yb).abs().max() / yb.abs().max()))
        same_sel = same_sel and a.gcv_yhat["id1se"] == b.gcv_yhat["id1se"]
    plan.capture(ys)
    gbest = float("inf")
    for _ in range(max(steps, 1)):
        torch.cuda.synchronize()
        t = time.perf_counter()
        plan.fit(ys)
        torch.cuda.synchronize()
        gbest = min(gbest, time.perf_counter() - t)
    classes = [c for v in plan._kindsets.values() for c in v]
    return {"workload": "512 uneven problems (1-D and 2-D, 1e3-1e6 points), "
                        "one result per problem as cv_fastlpr",
            "points": rows, "classes": len(classes),
            "batched_problems_per_s": P / best, "batched_ms": best * 1e3,
            "graph_replay_problems_per_s": P / gbest, "graph_replay_ms": gbest * 1e3,
            "looped_problems_per_s": P / t_loop, "looped_ms": t_loop * 1e3,
            "plan_setup_and_tuning_ms": t_setup * 1e3,
            "batched_vs_looped_yhat_max_rel_diff": diff, "same_selected_h": same_sel,
            "launch_shapes": sorted({str(c.tabs.get("launch_shape")) for c in classes})}


N_CU, N_SIMD, N_XCD = 256, 1024, 8     # MI355X: 8 XCDs of 32 CUs, 4 SIMDs per CU


def _kname(name: str) -> str:
    return name.replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "")


def _trace_summary(db, f, kernels=None):
    """Kernel-trace database: per kernel, calls, total / mean / min / max
    duration, share of the traced kernel time, and the launch resources."""
    rows = db.execute("select name, duration, vgpr_count, accum_vgpr_count, lds_size, "
                      "workgroup_x * workgroup_y * workgroup_z, grid_x * grid_y * grid_z "
                      "from kernels").fetchall()
    agg = {}
    for name, dur, vg, ag, lds, wg, grid in rows:
        a = agg.setdefault(_kname(name), [0, 0.0, float("inf"), 0.0, vg, ag, lds, wg, grid])
        a[0] += 1
        a[1] += dur
        a[2] = min(a[2], dur)
        a[3] = max(a[3], dur)
    total = sum(a[1] for a in agg.values()) or 1.0
    lines = [f"# kernel trace {f}: {len(rows)} dispatches, {total / 1e6:.3f} ms kernel time",
             f"{'kernel':60s} {'calls':>6s} {'total_us':>10s} {'avg_us':>9s} {'min_us':>9s} "
             f"{'max_us':>9s} {'share':>6s}  vgpr agpr  lds  wg  grid"]
    for k, a in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        if kernels and not re.search(kernels, k):
            continue
        lines.append(f"{k[:60]:60s} {a[0]:6d} {a[1] / 1e3:10.1f} {a[1] / a[0] / 1e3:9.2f} "
                     f"{a[2] / 1e3:9.2f} {a[3] / 1e3:9.2f} {100 * a[1] / total:5.1f}%  "
                     f"{a[4]:4d} {a[5]:4d} {a[6]:5d} {a[7]:4d} {a[8]}")
    return lines


def _derived(c, dur_ns):
    """Per-dispatch metrics from the mean counters of one kernel, following
    rocprofiler-sdk's gfx950 expressions (MfmaUtil, VALUBusy,
    OccupancyPercent, LDSBankConflict), which divide by the GRBM_GUI_ACTIVE of
    one XCD; the database holds it summed over the XCDs when GRBM per
    nanosecond exceeds the clock, and is then divided by their number."""
    out = []
    g = c.get("GRBM_GUI_ACTIVE")
    if not g:
        return out
    per_ns = g / dur_ns
    cyc = g / N_XCD if per_ns > 4.0 else g
    out.append(f"GRBM_GUI_ACTIVE/ns {per_ns:.2f} -> {cyc:.0f} cycles per dispatch "
               f"({cyc / dur_ns:.2f} GHz)")
    if "SQ_VALU_MFMA_BUSY_CYCLES" in c:
        out.append(f"MFMA busy {100 * c['SQ_VALU_MFMA_BUSY_CYCLES'] / (cyc * N_SIMD):.1f}%")
    if "SQ_ACTIVE_INST_VALU" in c:
        out.append(f"VALU busy {100 * c['SQ_ACTIVE_INST_VALU'] / N_CU / cyc:.1f}%")
    if "SQ_WAVE_CYCLES" in c:
        occ = 400 * c["SQ_WAVE_CYCLES"] / cyc / N_CU / 32
        out.append(f"occupancy {occ:.1f}% of 32 wave slots per CU "
                   f"({occ * 8 / 100:.2f} waves per SIMD)")
        if "SQ_WAIT_ANY" in c:
            out.append(f"waiting {100 * c['SQ_WAIT_ANY'] / c['SQ_WAVE_CYCLES']:.1f}% of wave cycles")
    if "SQ_LDS_BANK_CONFLICT" in c:
        out.append(f"LDS bank conflict {100 * c['SQ_LDS_BANK_CONFLICT'] / N_CU / cyc:.2f}% "
                   f"of CU cycles")
    if "SQ_LDS_BANK_CONFLICT" in c and c.get("SQ_LDS_IDX_ACTIVE"):
        out.append(f"conflict / LDS active {100 * c['SQ_LDS_BANK_CONFLICT'] / c['SQ_LDS_IDX_ACTIVE']:.1f}%")
    if "SQ_INSTS_VALU_MFMA_MOPS_BF16" in c:
        fl = c["SQ_INSTS_VALU_MFMA_MOPS_BF16"] * 512.0
        out.append(f"bf16 MFMA {fl:.3e} flop per dispatch = {fl / dur_ns / 1e3:.1f} TFLOP/s")
    waves = c.get("SQ_WAVES")
    if waves:
        mix = [f"{n[9:].lower()} {c[n] / waves:.1f}" for n in
               ("SQ_INSTS_VALU", "SQ_INSTS_MFMA", "SQ_INSTS_LDS", "SQ_INSTS_SALU", "SQ_INSTS_VMEM")
               if n in c]
        if mix:
            out.append("instructions per wave: " + ", ".join(mix))
    return out


def pmc_summary(dbs, kernels=None, brief=False):
    """Summaries of rocprofv3 databases (rocpd format).  A kernel-trace
    database gives the per-kernel time table; a --pmc database gives, per
    kernel, the mean of each counter per dispatch, the mean duration and the
    metrics derived from them, and, for FETCH_SIZE / WRITE_SIZE, the HBM
    traffic of flpr_conv per launch scaled by the calibration copy (a plain
    256 MiB device copy, --calibrate-copy), also per kernel as HBM bytes
    and GB/s.  Databases of several --pmc passes over the same program
    merge per kernel.  kernels: a regular expression; only the kernels whose
    name it matches are listed (the calibration copy is always used).  brief:
    the derived metrics without the raw counter lines."""
    import sqlite3
    tot, cnt, dur, res = {}, {}, {}, {}
    lines = []
    for f in dbs:
        db = sqlite3.connect(f)
        rows = db.execute("select kernel_name, counter_name, value, dispatch_id, duration, "
                          "vgpr_count, accum_vgpr_count, lds_block_size, workgroup_size, "
                          "grid_size from counters_collection").fetchall()
        if not rows:
            lines += _trace_summary(db, f, kernels)
            continue
        for name, cname, val, did, d, vg, ag, lds, wg, grid in rows:
            k = _kname(name)
            tot.setdefault(k, {}).setdefault(cname, 0.0)
            tot[k][cname] += val
            cnt.setdefault(k, {}).setdefault(cname, set()).add((f, did))
            dur.setdefault(k, {})[(f, did)] = d
            res[k] = (vg, ag, lds, wg, grid)
    avg = {k: {c: v / len(cnt[k][c]) for c, v in d.items()} for k, d in tot.items()}
    copy = [v for k, v in avg.items() if "neg_kernel" in k]
    scale = {}
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        cal = max((d.get(c, 0.0) for d in copy), default=0.0)
        if cal > 0:
            scale[c] = (COPY_BYTES / 1024.0) / cal
    for k in sorted(avg):
        if kernels and not re.search(kernels, k):
            continue
        dmean = sum(dur[k].values()) / len(dur[k])
        vg, ag, lds, wg, grid = res[k]
        lines.append(f"== {k[:90]}  ({len(dur[k])} dispatches, mean {dmean / 1e3:.2f} us, "
                     f"vgpr {vg} agpr {ag} lds {lds} wg {wg} grid {grid})")
        for c, v in sorted(avg[k].items()):
            if brief:
                break
            unit = " KiB" if c in ("FETCH_SIZE", "WRITE_SIZE") else ""
            lines.append(f"   {c:30s} {v:16.1f}{unit} per dispatch ({len(cnt[k][c])} dispatches)")
        lines += ["   " + t for t in _derived(avg[k], dmean)]
        if len(scale) == 2 and "FETCH_SIZE" in avg[k] and "WRITE_SIZE" in avg[k]:
            byt = (avg[k]["FETCH_SIZE"] * scale["FETCH_SIZE"] +
                   avg[k]["WRITE_SIZE"] * scale["WRITE_SIZE"]) * 1024.0
            lines.append(f"   HBM {byt / 1e6:.1f} MB per dispatch (FETCH x {scale['FETCH_SIZE']:.3f}"
                         f" + WRITE x {scale['WRITE_SIZE']:.3f}) = {byt / dmean:.0f} GB/s = "
                         f"{byt / dmean / (PEAK_HBM / 1e9):.3f} of 8 TB/s")
    out = {}
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        cal = max((d.get(c, 0.0) for d in copy), default=0.0)
        # mean over the float32 form's flpr_conv dispatches (k_conv_split, both
        # passes), as the roofline entry's launch average
        ktot = sum(tot[k][c] for k in tot if k.startswith("k_conv_split") and c in tot[k])
        kcnt = sum(len(cnt[k][c]) for k in cnt if k.startswith("k_conv_split") and c in cnt[k])
        if cal > 0 and kcnt:
            factor = (COPY_BYTES / 1024.0) / cal
            out[c] = (ktot / kcnt, factor)
    return lines, out


def read_traffic():
    """HBM bytes per flpr_conv launch from the committed counter summary."""
    try:
        with open(TRAFFIC_FILE) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the float64 plain form")
    ap.add_argument("--no-two-calls", action="store_true",
                    help="skip timing the pair as two separate cv_fastlpr calls")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched 512-problem leg")
    ap.add_argument("--no-cold", action="store_true",
                    help="skip the cold-cache call")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the timed steps, run a plain 256 MiB streaming kernel (negation) "
                         "three times: the known-size reference for FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--mixture-3d", type=int, default=0, metavar="K",
                    help="instead of the headline, time K fits each of the 3-D regression "
                         "and KDE on 1e7 mixture points (128^3 grid, 16 bandwidths) and "
                         "print that record")
    ap.add_argument("--pmc-summary", nargs="+", metavar="DB",
                    help="summarise rocprofv3 databases (kernel trace or --pmc; no GPU) and exit")
    ap.add_argument("--pmc-kernels", metavar="REGEX",
                    help="with --pmc-summary: list only the kernels whose name matches")
    ap.add_argument("--pmc-brief", action="store_true",
                    help="with --pmc-summary: derived metrics only, no raw counter lines")
    args = ap.parse_args()
    if args.pmc_summary:
        lines, out = pmc_summary(args.pmc_summary, args.pmc_kernels, args.pmc_brief)
        print("\n".join(lines))
        if "FETCH_SIZE" in out and "WRITE_SIZE" in out:
            (f, ff), (w, fw) = out["FETCH_SIZE"], out["WRITE_SIZE"]
            print(f"conv_fetch_kib_per_launch: {f:.1f} (copy scale {ff:.4f})")
            print(f"conv_write_kib_per_launch: {w:.1f} (copy scale {fw:.4f})")
            print(f"traffic_bytes_per_launch: {(f * ff + w * fw) * 1024.0:.0f}")
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the benchmark measures one GPU; the path would shard by bandwidth)",
              file=sys.stderr)
        sys.exit(2)

    if args.mixture_3d > 0:
        print(json.dumps(mixture_3d_leg(args.mixture_3d)), flush=True)
        if args.calibrate_copy:
            calibration_copy()
        return

    import torch
    import fastlpr_amd
    from fastlpr_amd import engine
    cv = fastlpr_amd.cv_fastlpr
    pair = fastlpr_amd.cv_fastlpr_pair
    n = WORKLOAD["n_points"]
    x, y = make_inputs(n, WORKLOAD["seed"])
    hl = hlist()
    dev = torch.device("cuda", 0)
    xd = torch.as_tensor(x, device=dev).to(torch.float64)
    yd = torch.as_tensor(y, device=dev).to(torch.float64)
    print(f"[bench] {torch.cuda.get_device_name(0)}; warmup {args.warmup}", file=sys.stderr,
          flush=True)

    for _ in range(max(args.warmup, 1)):
        step_pair(pair, xd, yd, hl, "fp32")
    torch.cuda.synchronize()
    gc.collect()
    gc.freeze()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r1, r2 = step_pair(pair, xd, yd, hl, "fp32")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # the same K steps again with HIP events around the convolution and
    # spread launches (on their launch stream), for the roofline entries; the
    # events are created before this region
    engine.TIMER.start(pool=16 * args.steps)
    torch.cuda.synchronize()
    te0 = time.perf_counter()
    for _ in range(args.steps):
        step_pair(pair, xd, yd, hl, "fp32")
    torch.cuda.synchronize()
    te1 = time.perf_counter()
    engine.TIMER.stop()
    kt = engine.TIMER.collect()
    ms_step_events = (te1 - te0) * 1e3 / args.steps
    assert r2.metadata["dof_shared"]
    two_call_fits = pair_vs_calls = None
    per_pass = {}
    if not args.no_two_calls:
        # the same pair as two independent calls
        step(cv, xd, yd, hl, "fp32")
        torch.cuda.synchronize()
        tc0 = time.perf_counter()
        for _ in range(args.steps):
            c1, c2 = step(cv, xd, yd, hl, "fp32")
        torch.cuda.synchronize()
        tc1 = time.perf_counter()
        two_call_fits = 2.0 * args.steps / (tc1 - tc0)
        pair_vs_calls = max(float((r1.yhat - c1.yhat).abs().max() / c1.yhat.abs().max()),
                            float((r2.yhat - c2.yhat).abs().max() / c2.yhat.abs().max()))
        # each pass on its own, as separate cv_fastlpr calls
        opt = {"order": 2, "N": [256, 256], "precision": "fp32"}
        r2y = (yd - c1.yhat) ** 2
        for name, yy, oo in (("mean_pass", yd, opt),
                             ("variance_pass", r2y, dict(opt, y_type_out="variance"))):
            torch.cuda.synchronize()
            tp0 = time.perf_counter()
            for _ in range(args.steps):
                cv(xd, yy, hl, oo)
            torch.cuda.synchronize()
            per_pass[name + "_fits_per_s"] = args.steps / (time.perf_counter() - tp0)
    if args.calibrate_copy:
        calibration_copy()
    ms_step = (t1 - t0) * 1e3 / args.steps
    fits_s = 2.0 * args.steps / (t1 - t0)
    print(f"[bench] chosen: {ms_step:.2f} ms/step ({ms_step_events:.2f} with events), "
          f"{fits_s:.1f} fits/s "
          f"(two calls: {two_call_fits} fits/s, yhat difference {pair_vs_calls})",
          file=sys.stderr, flush=True)
    conv = kt.get("flpr_conv", [])
    spr = kt.get("flpr_spread", [])
    spread_entry = None
    if spr:
        sp_ms = float(np.mean([c[0] for c in spr]))
        sp_flop = float(np.mean([c[1][0] for c in spr]))
        sp_bytes = float(np.mean([c[1][1] for c in spr]))
        sp_ach = sp_flop / (sp_ms * 1e-3)
        spread_entry = {"bound": "valu-fp64", "kernel": "flpr_spread (k_spread_cells + gather)",
                        "achieved": sp_ach / 1e12, "peak": PEAK_FP64_VALU / 1e12,
                        "unit": "TFLOP/s", "frac": sp_ach / PEAK_FP64_VALU,
                        "hbm_gb_s": sp_bytes / (sp_ms * 1e-3) / 1e9,
                        "hbm_frac": sp_bytes / (sp_ms * 1e-3) / PEAK_HBM,
                        "avg_launch_ms": sp_ms, "flop_per_launch": sp_flop,
                        "bytes_per_launch": sp_bytes, "launches_timed": len(spr)}
    split_peak = PEAK_MFMA_BF16 / SPLIT_PRODUCTS / 1e12
    conv_ms = float(np.mean([c[0] for c in conv])) if conv else float("nan")
    conv_flop = float(np.mean([c[1] for c in conv])) if conv else float("nan")
    achieved = conv_flop / (conv_ms * 1e-3) / 1e12 if conv else float("nan")
    conv_share = sum(c[0] for c in conv) / (args.steps * ms_step_events) if conv else float("nan")

    extra = {}
    if not args.no_plain:
        for _ in range(1):
            step_pair(pair, xd, yd, hl, "fp64")
        torch.cuda.synchronize()
        tp0 = time.perf_counter()
        ksteps = max(2, args.steps // 3)
        for _ in range(ksteps):
            p1, p2 = step_pair(pair, xd, yd, hl, "fp64")
        torch.cuda.synchronize()
        tp1 = time.perf_counter()
        plain_fits = 2.0 * ksteps / (tp1 - tp0)
        e1 = float((r1.yhat - p1.yhat).abs().max() / p1.yhat.abs().max())
        e2 = float((r2.yhat - p2.yhat).abs().max() / p2.yhat.abs().max())
        extra = {"plain_fp64_fits_per_s": plain_fits,
                 "chosen_vs_plain_yhat_max_rel_err": {"mean_pass": e1, "variance_pass": e2},
                 "selected_h_index": {"mean_pass": int(r1.gcv_yhat["id1se"]),
                                      "variance_pass": int(r2.gcv_yhat["id1se"]),
                                      "plain_mean_pass": int(p1.gcv_yhat["id1se"])}}
        print(f"[bench] plain fp64: {plain_fits:.1f} fits/s; chosen-vs-plain yhat err "
              f"{e1:.2e} / {e2:.2e}", file=sys.stderr, flush=True)
        # the precision form (float32 spread with compensated summation,
        # float32 solves refined once in float64)
        step_pair(pair, xd, yd, hl, "mixed")
        torch.cuda.synchronize()
        tm0 = time.perf_counter()
        for _ in range(ksteps):
            q1, q2 = step_pair(pair, xd, yd, hl, "mixed")
        torch.cuda.synchronize()
        mixed_fits = 2.0 * ksteps / (time.perf_counter() - tm0)
        extra["mixed_precision_fits_per_s"] = mixed_fits
        extra["mixed_vs_plain_yhat_max_rel_err"] = {
            "mean_pass": float((q1.yhat - p1.yhat).abs().max() / p1.yhat.abs().max()),
            "variance_pass": float((q2.yhat - p2.yhat).abs().max() / p2.yhat.abs().max())}
        print(f"[bench] precision form (mixed): {mixed_fits:.1f} fits/s", file=sys.stderr,
              flush=True)
        # the learned bandwidth bracket on the mean pass (drop-in cv_fastlpr)
        ob = {"order": 2, "N": [256, 256], "precision": "fp32"}
        res = {}
        for mode in ("full", "learned"):
            rb = cv(xd, yd, hl, dict(ob, bandwidth_bracket=mode))
            torch.cuda.synchronize()
            tb0 = time.perf_counter()
            for _ in range(ksteps):
                rb = cv(xd, yd, hl, dict(ob, bandwidth_bracket=mode))
            torch.cuda.synchronize()
            res[mode] = (ksteps / (time.perf_counter() - tb0), rb)
        binfo = res["learned"][1].metadata["bracket"]
        hsel = res["learned"][1].h
        extra["learned_bracket"] = {
            "fits_per_s": res["learned"][0], "full_sweep_fits_per_s": res["full"][0],
            "bracket_indices": [int(i) for i in binfo["indices"]],
            "fallback": bool(binfo["fallback"]),
            "same_selected_h": bool(np.allclose(hsel, res["full"][1].h)),
            "call": "cv_fastlpr mean pass, options bandwidth_bracket='learned' vs 'full'"}
        print(f"[bench] learned bracket: {res['learned'][0]:.1f} fits/s vs full sweep "
              f"{res['full'][0]:.1f}", file=sys.stderr, flush=True)

    cold = None
    if not args.no_cold:
        # a first call at this n: the probe matrix is generated on the host
        # (NumPy MT19937, as the reference does on every call) and every plan,
        # geometry and table is rebuilt
        fastlpr_amd.clear_caches()
        torch.cuda.synchronize()
        t = time.perf_counter()
        step_pair(pair, xd, yd, hl, "fp32")
        torch.cuda.synchronize()
        t_pair = time.perf_counter() - t
        fastlpr_amd.clear_caches()
        torch.cuda.synchronize()
        t = time.perf_counter()
        cv(xd, yd, hl, {"order": 2, "N": [256, 256], "precision": "fp32"})
        torch.cuda.synchronize()
        t_one = time.perf_counter() - t
        cold = {"pair_ms": t_pair * 1e3, "fits_per_s_pair": 2.0 / t_pair,
                "single_call_ms": t_one * 1e3, "fits_per_s_single_call": 1.0 / t_one,
                "note": "clear_caches() before each call: probe generation on the host "
                        "(1e6 x 10 normals) and every plan rebuilt"}
        print(f"[bench] cold call: pair {t_pair*1e3:.1f} ms, single {t_one*1e3:.1f} ms",
              file=sys.stderr, flush=True)
    b512 = None
    if not args.no_batch:
        print("[bench] batch: 512 uneven problems, batched vs looped ...", file=sys.stderr,
              flush=True)
        b512 = uneven_batch_leg(3)
        print(f"[bench] batch: batched {b512['batched_problems_per_s']:.0f} problems/s, "
              f"looped {b512['looped_problems_per_s']:.0f}", file=sys.stderr, flush=True)
    cpu = cpu_all = None
    if not args.no_cpu:
        print("[bench] cpu baseline (oracle, 1 thread, bounded sample) ...", file=sys.stderr,
              flush=True)
        cpu = cpu_baseline(x, y, hl)
        ncores = len(os.sched_getaffinity(0))
        print(f"[bench] cpu baseline (oracle, {ncores} threads, bounded sample) ...",
              file=sys.stderr, flush=True)
        cpu_all = cpu_baseline(x, y, hl, threads=ncores)

    line = {
        "metric": "cv_fastlpr fits/sec, 2D N=1e6 order2 30 h",
        "value": fits_s,
        "unit": "fits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": dict(WORKLOAD, form="chosen: float32 convolution (bf16 MFMA, exact 3-term "
                       "splits), float64 spread/solve",
                       fits_per_step=2, call="cv_fastlpr_pair (second pass shares the DOF)"),
        "roofline": {"bound": "mfma", "kernel": "flpr_conv (float32 form, k_conv_split)",
                     "achieved": achieved, "peak": split_peak, "unit": "TFLOP/s",
                     "peak_basis": "bf16 dense MFMA 2.5 PFLOP/s / 6 split products per "
                                   "float32 product (float32 MFMA peak: 157.3)",
                     "frac": achieved / split_peak if conv else None,
                     "traffic": read_traffic(), "avg_launch_ms": conv_ms,
                     "flop_per_launch": conv_flop, "launches_timed": len(conv),
                     "share_of_step_time": conv_share,
                     "timing": "a second pass of the same K steps with HIP events around "
                               "each convolution launch on its stream",
                     "ms_per_step_with_events": ms_step_events},
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
        "roofline_spread": spread_entry,
        "cold_call": cold,
        "uneven_512_batched": b512,
    }
    line.update(extra)
    line["two_call_fits_per_s"] = two_call_fits
    line.update(per_pass)
    line["pair_vs_two_calls_yhat_max_rel_diff"] = pair_vs_calls
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
