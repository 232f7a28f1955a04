"""Write the golden fixtures under tests/golden/ by running the reference fastLPR.

Run in the development container only (the reference checkout is not present
where the GPU tests run):

    FASTLPR_QUIET=1 python tests/golden/make_golden.py /root/reference/fastLPR_py/src [names...]

Each fixture stores the inputs (or, for the large case, the seeded recipe and a
hash of the inputs), the options, and the reference outputs of
``cv_fastlpr`` / ``cv_fastkde`` (and ``fastlpr_interval`` / ``fastlpr_predict``): yhat (or a seeded subsample of it plus its
hash), gcv_m, gcv_sd, dof, dof_stderr, idmin, id1se, the selected bandwidth,
the grid values at the selected bandwidth, s0 and the variance factor d.
These end-to-end compositions (probe stream x interpolation x float32
rounding x output-grid offset) are what only the reference itself can pin.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import warnings

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _hash(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def recipe_inputs(name: str):
    """Seeded inputs; every case is regenerated identically by tests/golden/recipes.py."""
    sys.path.insert(0, HERE)
    import recipes
    return recipes.make(name)


def run_lpr(fl, name, store_inputs=True, n_sub=None):
    x, y, h, opt = recipe_inputs(name)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        r = fl.cv_fastlpr(x, y, h, dict(opt))
    g = r.gcv_yhat
    out = dict(
        kind="lpr", name=name, options=json.dumps(opt), hlist=np.asarray(h, float),
        gcv_m=np.asarray(g["gcv_m"]), gcv_sd=np.asarray(g["gcv_sd"]),
        dof=np.asarray(g["dof"]), dof_stderr=np.asarray(g["dof_stderr"]),
        idmin=int(g["idmin"]), id1se=int(g["id1se"]), h1se=np.asarray(g["h1se"]),
        grid_values=np.asarray(r.fpp_yhat.values if hasattr(r.fpp_yhat, "values") else np.nan),
        s0=np.asarray(r.s0), d=np.nan if r.d is None else float(r.d),
        x_hash=_hash(x), y_hash=_hash(y), yhat_hash=_hash(np.asarray(r.yhat)),
    )
    if store_inputs:
        out.update(x=x, y=y, yhat=np.asarray(r.yhat))
    else:
        idx = np.random.default_rng(1).choice(len(x), size=n_sub, replace=False)
        idx.sort()
        out.update(yhat_idx=idx, yhat_sub=np.asarray(r.yhat)[idx])
    return out


def run_kde(fl, name):
    x, _, h, opt = recipe_inputs(name)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        r = fl.cv_fastkde(x, h, dict(opt))
    lc = r.lcv
    return dict(kind="kde", name=name, options=json.dumps(opt), hlist=np.asarray(h, float),
                x=x, fhat=np.asarray(r.fhat), lcv_m=np.asarray(lc["lcv_m"]),
                lcv_sd=np.asarray(lc["lcv_sd"]), idmax=int(lc["idmax"]),
                id1se=int(lc["id1se"]), h1se=np.asarray(lc["h1se"]),
                grid=np.stack([np.asarray(g) for g in r.grid]) if len({len(g) for g in r.grid}) == 1
                else np.asarray(np.concatenate(r.grid)))


def run_interval(fl, name):
    """Mean fit, variance fit of the squared residuals, fastlpr_interval
    (confidence and prediction, alpha 0.05) and fastlpr_predict at new points."""
    import recipes
    x, y, h, opt = recipe_inputs(name)
    xn = recipes.predict_points(name)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        mu = fl.cv_fastlpr(x, y, h, dict(opt))
        sg = fl.cv_fastlpr(x, (y - mu.yhat) ** 2, h, dict(opt, y_type_out="variance"))
        ci = fl.fastlpr_interval(mu, sg, 0.05, type="confidence")
        pi = fl.fastlpr_interval(mu, sg, 0.05, type="prediction")
        pred = fl.fastlpr_predict(mu, xn)
    return dict(kind="interval", name=name, options=json.dumps(opt), hlist=np.asarray(h, float),
                x=x, y=y, x_new=xn, pred=np.asarray(pred), ci=np.asarray(ci),
                pi=np.asarray(pi), mu_grid=np.asarray(mu.fpp_yhat.values),
                sigma_grid=np.asarray(sg.fpp_yhat.values), s0=np.asarray(mu.s0),
                h_mu=np.asarray(mu.h), h_sigma=np.asarray(sg.h))


def main(ref_src: str, only=None):
    sys.path.insert(0, ref_src)
    os.environ.setdefault("FASTLPR_QUIET", "1")
    import fastlpr as fl
    import scipy
    from fastlpr.backend_selection import select_accumulation_backend
    meta = dict(numpy=np.__version__, scipy=scipy.__version__,
                backend=select_accumulation_backend())
    sys.path.insert(0, HERE)
    import recipes
    def wanted(name):
        return not only or name in only

    for name in recipes.LPR_SMALL + recipes.LPR_CUBIC + recipes.LPR_ACC + recipes.LPR_ACC_HI:
        if not wanted(name):
            continue
        out = run_lpr(fl, name)
        out["meta"] = json.dumps(meta)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        print("wrote", name)
    for name in recipes.LPR_LARGE:
        if not wanted(name):
            continue
        out = run_lpr(fl, name, store_inputs=False, n_sub=4096)
        out["meta"] = json.dumps(meta)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        print("wrote", name)
    for name in recipes.IVL:
        if not wanted(name):
            continue
        out = run_interval(fl, name)
        out["meta"] = json.dumps(meta)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        print("wrote", name)
    for name in recipes.KDE + recipes.KDE_ACC_HI:
        if not wanted(name):
            continue
        out = run_kde(fl, name)
        out["meta"] = json.dumps(meta)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        print("wrote", name)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/fastLPR_py/src", sys.argv[2:])
