"""sklearn-style estimator over the gfx950 kernels.

Interface of the reference's FlashSDKDE (flash_sd_kde/estimator.py:22-157):
the constructor fields and their defaults, the fitted attributes bandwidth_,
n_features_in_ and X_fit_, fit / score_samples / score, ValueError for bad
input or an unfitted estimator, and score_samples = log(density + eps) as a
numpy array (:114). mode also accepts "sdkde" for "sd_kde".

How it is organised here:
* input checks and the bandwidth rules are module-level and shared by fit
  and score_samples; fit checks everything before it sets any attribute, so
  a failed fit leaves the estimator as it was;
* the evaluation set (the training points, or for SD-KDE their debiased
  copy, computed on the GPU) is made once in fit and moved to the device on
  the first score_samples call, so repeated scoring uploads only the queries;
* the dimension picks the kernels: 16 features run the reference's 16-D
  path with the chosen backends, 1 feature (mode="kde") the 1-D kernels, and
  any other dimension the zero-padded 16-D kernels of flash_sd_kde.ndim
  (exact: padding leaves distances unchanged). The reference accepts only
  16 features for SD-KDE and 1 or 16 for KDE.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, Literal

import numpy as np
import torch

from flash_sd_kde import ndim
from flash_sd_kde.bandwidth import silverman_bandwidth_1d, silverman_bandwidth_nd
from flash_sd_kde.kde import emp_sd_kde_fit_transform, kde_eval
from globals import DEFAULT_EPS, DEFAULT_KDE_BACKEND, DEFAULT_PRECISION_MODE, ND_FEATURES

FlashMode = Literal["sd_kde", "kde"]
BandwidthSpec = float | Literal["silverman"]

_CANONICAL_MODE = {"sd_kde": "sd_kde", "sdkde": "sd_kde", "kde": "kde"}


def _silverman(x: np.ndarray) -> float:
    """The reference's rule of thumb: 1-D for one feature, else the N-D form."""
    return silverman_bandwidth_1d(x[:, 0]) if x.shape[1] == 1 else silverman_bandwidth_nd(x)


BANDWIDTH_RULES: dict[str, Callable[[np.ndarray], float]] = {"silverman": _silverman}


def as_samples(X) -> np.ndarray:
    """(n_samples, n_features) float32 array, or ValueError."""
    arr = np.asarray(X, dtype=np.float32)
    if arr.ndim != 2:
        raise ValueError(f"Expected X with shape (n_samples, n_features) as in sklearn, "
                         f"got an array of {arr.ndim} dimension(s).")
    if arr.shape[0] == 0:
        raise ValueError("X must contain at least one sample.")
    if arr.shape[1] == 0:
        raise ValueError("X must contain at least one feature.")
    return arr


def resolve_bandwidth(spec: BandwidthSpec, x: np.ndarray) -> float:
    """A positive float, or a rule name of BANDWIDTH_RULES applied to x."""
    if isinstance(spec, str):
        rule = BANDWIDTH_RULES.get(spec)
        if rule is None:
            raise ValueError(f"Unsupported bandwidth={spec!r}: give a positive float or one of "
                             f"{sorted(BANDWIDTH_RULES)}.")
        return float(rule(x))
    h = float(spec)
    if not h > 0.0:
        raise ValueError(f"bandwidth must be positive, got {h}.")
    return h


def canonical_mode(mode: str) -> str:
    try:
        return _CANONICAL_MODE[mode]
    except KeyError:
        raise ValueError(f"Unsupported mode={mode!r}; expected 'sd_kde' (or 'sdkde') or 'kde'.") \
            from None


@dataclass
class FlashSDKDE:
    """KDE / one-step empirical SD-KDE estimator on MI355X.

    fit(X) keeps the training points (mode="kde") or their SD-debiased copy
    (mode="sd_kde"); score_samples(X) returns log(density + eps) per sample;
    score(X) their mean.
    """

    bandwidth: BandwidthSpec = "silverman"
    mode: FlashMode = "sd_kde"
    device: str | torch.device = "cuda"
    precision_mode: str = DEFAULT_PRECISION_MODE
    kde_backend: str = DEFAULT_KDE_BACKEND
    emp_score_backend: str | None = None
    use_precomputed_norms: bool = True
    autotune: bool = True
    eps: float = DEFAULT_EPS

    bandwidth_: float | None = field(default=None, init=False)
    n_features_in_: int | None = field(default=None, init=False)
    X_fit_: np.ndarray | None = field(default=None, init=False, repr=False)
    _eval_set: torch.Tensor | np.ndarray | None = field(default=None, init=False, repr=False)

    # ---- sklearn interface ----
    def fit(self, X, y=None) -> "FlashSDKDE":
        del y
        x = as_samples(X)
        h = resolve_bandwidth(self.bandwidth, x)
        mode = canonical_mode(self.mode)
        eval_set = self._debiased(x, h) if mode == "sd_kde" else x
        self.bandwidth_, self.n_features_in_, self.X_fit_ = h, x.shape[1], x
        self._eval_set = eval_set
        return self

    def score_samples(self, X) -> np.ndarray:
        if self._eval_set is None:
            raise ValueError("Estimator is not fitted. Call fit(X) before score_samples(X).")
        q = as_samples(X)
        if q.shape[1] != self.n_features_in_:
            raise ValueError(f"X has {q.shape[1]} features, but the estimator was fitted with "
                             f"{self.n_features_in_}.")
        density = self._density(self._device_eval_set(), q)
        return torch.log(density + self.eps).detach().cpu().numpy()

    def score(self, X, y=None) -> float:
        del y
        return float(np.mean(self.score_samples(X)))

    # ---- evaluation-set construction and density, by dimension ----
    def _debiased(self, x: np.ndarray, h: float):
        if x.shape[1] == ND_FEATURES:
            return emp_sd_kde_fit_transform(
                x, h, device=self.device, precision_mode=self.precision_mode,
                emp_score_backend=self.emp_score_backend,
                use_precomputed_norms=self.use_precomputed_norms, autotune=self.autotune,
                eps=self.eps)
        return ndim.sd_kde_fit_transform_nd(x, h, device=self.device, eps=self.eps)

    def _device_eval_set(self) -> torch.Tensor:
        dev = torch.device(self.device)
        s = self._eval_set
        if not (isinstance(s, torch.Tensor) and s.device.type == dev.type):
            s = torch.as_tensor(np.ascontiguousarray(s), dtype=torch.float32, device=dev)
            self._eval_set = s
        return s

    def _density(self, data: torch.Tensor, q: np.ndarray) -> torch.Tensor:
        d = self.n_features_in_
        if d == ND_FEATURES or d == 1:
            return kde_eval(data[:, 0] if d == 1 else data, q[:, 0] if d == 1 else q,
                            self.bandwidth_, device=self.device,
                            precision_mode=self.precision_mode, kde_backend=self.kde_backend,
                            use_precomputed_norms=self.use_precomputed_norms,
                            autotune=self.autotune)
        return ndim.kde_eval_nd(data, q, self.bandwidth_, device=self.device)
