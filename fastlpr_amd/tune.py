"""Launch-shape autotuner for batched problems (BatchPlan).

Two knobs decide how the point kernels cut their work; neither changes what
is computed beyond the order of float64 additions:

* ``tile``        spread / adjoint tile edge in cells (FlprGeom.tile): the
                  LDS window a block accumulates, and the blocks per grid;
* ``item_points`` points per spread / adjoint work item when a cell holds
                  many (FlprGeom.item_points): more points per item means
                  fewer LDS flushes, fewer items means fewer waves.

For every class of a plan the tuner times the class's whole batched stage
sequence (``BatchPlan._run_class`` on a synthetic response, HIP events on
the launch stream, best of three after a warm-up) for each candidate and
keeps the fastest.  Choices are memoised per class signature (dimension,
grid, order, channels, cubic flags, problem count and a log2 bucket of the
point count) in memory and, when ``FASTLPR_AMD_TUNE_FILE`` names a file, in
that JSON file, so a later process reuses them without measuring.
"""
from __future__ import annotations

import json
import math
import os
from typing import Dict, Optional, Tuple

import torch

CANDIDATES = {
    1: [(t, ip) for t in (16, 32, 64) for ip in (8, 32, 128)],
    2: [(t, ip) for t in (8, 16) for ip in (16, 64)],
    3: [(8, 16), (8, 64)],
}
DEFAULT = {1: (16, 0), 2: (8, 0), 3: (8, 0)}

_CHOICES: Dict[str, Tuple[int, int]] = {}
_LOADED = False


def _file() -> Optional[str]:
    return os.environ.get("FASTLPR_AMD_TUNE_FILE") or None


def _load():
    global _LOADED
    if _LOADED:
        return
    _LOADED = True
    f = _file()
    if f and os.path.exists(f):
        try:
            with open(f) as fh:
                for k, v in json.load(fh).items():
                    _CHOICES[k] = (int(v[0]), int(v[1]))
        except (OSError, ValueError):
            pass


def _save():
    f = _file()
    if not f:
        return
    try:
        os.makedirs(os.path.dirname(os.path.abspath(f)), exist_ok=True)
        with open(f, "w") as fh:
            json.dump({k: list(v) for k, v in sorted(_CHOICES.items())}, fh, indent=1)
    except OSError:
        pass


def signature(c) -> str:
    """Class signature the choice is kept under."""
    rows = int(c.tabs["rows"])
    return "d{}_N{}_o{}_C{}_cub{}{}_P{}_r{}".format(
        c.d, "x".join(str(v) for v in c.N), c.order, c.tabs["C"], int(c.cubic[0]),
        int(c.cubic[1]), len(c.probs), int(math.log2(max(rows, 1))))


def lookup(c) -> Optional[Tuple[int, int]]:
    _load()
    return _CHOICES.get(signature(c))


def tune_class(plan, c, retile) -> Tuple[int, int]:
    """Measure every candidate (tile, item_points) for class c of plan and
    return the fastest; retile(c, tile, item_points) rebuilds the class's
    geometry table for a candidate."""
    _load()
    sig = signature(c)
    hit = _CHOICES.get(sig)
    if hit is not None:
        return hit
    dev = plan.dev
    y = torch.zeros((int(c.tabs["rows"]), c.ny), dtype=torch.float64, device=dev)
    best, best_ms = DEFAULT[c.d], float("inf")
    timings = {}
    for cand in CANDIDATES[c.d]:
        retile(c, *cand)
        plan._run_class(c, y)            # warm-up (block tables, LDS configs)
        ms = float("inf")
        for _ in range(3):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            plan._run_class(c, y)
            b.record()
            b.synchronize()
            ms = min(ms, a.elapsed_time(b))
        timings[cand] = ms
        if ms < best_ms:
            best, best_ms = cand, ms
    retile(c, *best)
    _CHOICES[sig] = best
    c.tabs["tune_ms"] = {f"{t}x{ip}": round(v, 3) for (t, ip), v in timings.items()}
    _save()
    return best


def choices() -> Dict[str, Tuple[int, int]]:
    """Every choice made or loaded in this process."""
    _load()
    return dict(_CHOICES)
