"""fastlpr_amd: MI355X engine for the fastLPR cv_fastlpr / cv_fastkde hot path.

Public API (same names and semantics as the reference package ``fastlpr``):
``cv_fastlpr``, ``cv_fastkde``, ``get_hlist``, ``fastlpr_predict``,
``fastlpr_interval``, ``RegressionOutput``, ``KDEOutput``; plus
``cv_fastlpr_pair`` (the heteroscedastic mean + variance fit in one call) and
``cv_fastlpr_batch`` / ``BatchPlan`` (many independent problems in shared
launches, optionally replayed from a captured HIP graph).
"""
from .bandwidth import get_hlist
from .structures import KDEOutput, RegressionOutput

__version__ = "0.1.0"


def __getattr__(name):
    # the API pulls in torch and the HIP library lazily
    if name in ("cv_fastlpr", "cv_fastkde", "clear_caches", "GridInterpolant",
                "fastlpr_predict", "fastlpr_interval", "cv_fastlpr_pair"):
        from . import api
        return getattr(api, name)
    if name in ("cv_fastlpr_batch", "BatchPlan"):
        from . import batch
        return getattr(batch, name)
    raise AttributeError(name)


__all__ = ["cv_fastlpr", "cv_fastkde", "get_hlist", "RegressionOutput", "KDEOutput",
           "clear_caches", "fastlpr_predict", "fastlpr_interval", "cv_fastlpr_pair",
           "cv_fastlpr_batch", "BatchPlan"]
