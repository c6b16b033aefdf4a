"""Flash-SD-KDE on MI355X: Gaussian KDE and one-step empirical SD-KDE with the
pairwise sums in hand-written gfx950 HIP kernels (libflashsdkde.so).

Import surface of the reference package (flash_sd_kde/__init__.py:1-15).
"""

from flash_sd_kde.estimator import FlashSDKDE
from flash_sd_kde.kde import (
    emp_sd_kde_fit_transform,
    kde_eval,
    kde_eval_linearized,
    kde_eval_linearized_nonfused,
)

__all__ = [
    "FlashSDKDE",
    "emp_sd_kde_fit_transform",
    "kde_eval",
    "kde_eval_linearized",
    "kde_eval_linearized_nonfused",
]
