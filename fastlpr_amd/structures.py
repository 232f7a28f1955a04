"""Result containers with the reference's field names (structures.py:24-71)."""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, Optional

import numpy as np


@dataclass
class RegressionOutput:
    """Output of cv_fastlpr (structures.py:24-50).

    yhat        fitted values at the data points
    grid        evaluation grid (tuple of raw-space axes)
    h           selected bandwidth (d,)
    gcv_yhat    dict: bandwidths, gcv_m, gcv_sd, idmin, id1se, h1se, hmin, dof, dof_stderr
    fpp_yhat    interpolant for prediction at new points
    metadata    options, order, bandwidths_tested, fitted_at_samples, residuals, mean, std, ...
    dof, dof_stderr   DOF at the selected bandwidth
    d           variance normalisation factor (y_type_out='variance'), else None
    s0          zero-th kernel moment on the grid at the selected bandwidth
    """

    yhat: np.ndarray
    grid: tuple
    h: np.ndarray
    gcv_yhat: Optional[Dict[str, object]] = None
    fpp_yhat: Optional[object] = None
    metadata: Optional[Dict[str, object]] = None
    dof: Optional[float] = None
    dof_stderr: Optional[float] = None
    d: Optional[float] = None
    s0: Optional[np.ndarray] = None


@dataclass
class KDEOutput:
    """Output of cv_fastkde (structures.py:53-71)."""

    fhat: np.ndarray
    grid: tuple
    h: np.ndarray
    lcv: Optional[Dict[str, object]] = None
    fpp: Optional[object] = None
    metadata: Optional[Dict[str, object]] = None
