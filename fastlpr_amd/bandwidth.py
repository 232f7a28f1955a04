"""Bandwidth candidates (bandwidth.py:22-137 of the reference)."""
from __future__ import annotations

from typing import Callable, Sequence, Union

import numpy as np


def get_hlist(n: Union[Sequence[int], int], range, spacing: Callable | None = None) -> np.ndarray:
    """Bandwidth grid of shape (prod(n), d).

    Per dimension ``spacing(lo, hi, n_k)`` (default log-spaced, a single value
    when n_k == 1); the Cartesian product is flattened with the first
    dimension varying fastest (MATLAB ndgrid order), as bandwidth.py:126-136.
    """
    r = np.asarray(range, dtype=float)
    if r.ndim == 1:
        r = r.reshape(1, 2)
    d = r.shape[0]
    counts = [int(n)] * d if np.isscalar(n) else [int(v) for v in n]
    if len(counts) != d:
        raise ValueError("n must match the number of rows in range.")
    if spacing is None:
        def spacing(lo, hi, num):
            return np.logspace(np.log10(lo), np.log10(hi), num)
    axes = [spacing(lo, hi, k) if k > 1 else np.array([lo], dtype=float)
            for (lo, hi), k in zip(r, counts)]
    mesh = np.meshgrid(*axes, indexing="ij")
    return np.column_stack([m.ravel("F") for m in mesh])
