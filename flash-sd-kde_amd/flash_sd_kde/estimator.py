"""sklearn-style estimator (reference flash_sd_kde/estimator.py:22-157).

FlashSDKDE(bandwidth, mode, device, precision_mode, kde_backend,
emp_score_backend, use_precomputed_norms, autotune, eps) with fit /
score_samples / score. score_samples returns log(density + eps) as numpy,
exactly like the reference (:114); the precision forms in
flash_sd_kde.precision give log-densities by log-sum-exp instead.
mode accepts "sd_kde" (reference spelling) and "sdkde" as an alias.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Literal

import numpy as np
import torch

from flash_sd_kde.bandwidth import silverman_bandwidth_1d, silverman_bandwidth_nd
from flash_sd_kde.kde import emp_sd_kde_fit_transform, kde_eval
from globals import DEFAULT_EPS, DEFAULT_KDE_BACKEND, DEFAULT_PRECISION_MODE, ND_FEATURES

FlashMode = Literal["sd_kde", "kde"]
BandwidthSpec = float | Literal["silverman"]
_MODE_ALIASES = {"sd_kde": "sd_kde", "sdkde": "sd_kde", "kde": "kde"}


@dataclass
class FlashSDKDE:
    """KDE / one-step empirical SD-KDE estimator on MI355X.

    fit(X) stores the training data (mode="kde") or its SD-debiased copy
    (mode="sd_kde", 16 features); score_samples(X) returns per-sample log
    densities; score(X) their mean.
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
    _fit_eval_data: torch.Tensor | np.ndarray | None = field(default=None, init=False, repr=False)
    _is_fitted: bool = field(default=False, init=False, repr=False)

    def fit(self, X: np.ndarray, y=None) -> "FlashSDKDE":
        del y
        x_fit = self._validate_input(X)
        self.n_features_in_ = x_fit.shape[1]
        self.bandwidth_ = self._resolve_bandwidth(x_fit)
        self.X_fit_ = x_fit
        mode = self._validate_mode()
        if mode == "sd_kde":
            if self.n_features_in_ != ND_FEATURES:
                raise ValueError(
                    f'mode="sd_kde" currently requires {ND_FEATURES} features, '
                    f"but got {self.n_features_in_}."
                )
            self._fit_eval_data = emp_sd_kde_fit_transform(
                x_fit, self.bandwidth_, device=self.device, precision_mode=self.precision_mode,
                emp_score_backend=self.emp_score_backend,
                use_precomputed_norms=self.use_precomputed_norms, autotune=self.autotune,
                eps=self.eps)
        elif self.n_features_in_ == 1:
            self._fit_eval_data = x_fit[:, 0].astype(np.float32, copy=False)
        else:
            self._fit_eval_data = x_fit
        self._is_fitted = True
        return self

    def score_samples(self, X: np.ndarray) -> np.ndarray:
        self._check_is_fitted()
        queries = self._validate_input(X)
        if queries.shape[1] != self.n_features_in_:
            raise ValueError(
                f"X has {queries.shape[1]} features, but estimator was fitted with "
                f"{self.n_features_in_} features."
            )
        query_eval = queries[:, 0].astype(np.float32, copy=False) if self.n_features_in_ == 1 else queries
        density = kde_eval(self._fit_eval_data, query_eval, self.bandwidth_, device=self.device,
                           precision_mode=self.precision_mode, kde_backend=self.kde_backend,
                           use_precomputed_norms=self.use_precomputed_norms,
                           autotune=self.autotune)
        return torch.log(density + self.eps).detach().cpu().numpy()

    def score(self, X: np.ndarray, y=None) -> float:
        del y
        return float(np.mean(self.score_samples(X)))

    def _validate_input(self, X: np.ndarray) -> np.ndarray:
        arr = np.asarray(X, dtype=np.float32)
        if arr.ndim != 2:
            raise ValueError(
                "Expected X with shape (n_samples, n_features), consistent with sklearn."
            )
        if arr.shape[0] == 0:
            raise ValueError("X must contain at least one sample.")
        if arr.shape[1] == 0:
            raise ValueError("X must contain at least one feature.")
        return arr

    def _resolve_bandwidth(self, X: np.ndarray) -> float:
        if isinstance(self.bandwidth, str):
            if self.bandwidth != "silverman":
                raise ValueError(
                    f"Unsupported bandwidth='{self.bandwidth}'. Use float or 'silverman'."
                )
            if X.shape[1] == 1:
                return float(silverman_bandwidth_1d(X[:, 0]))
            return float(silverman_bandwidth_nd(X))
        bw = float(self.bandwidth)
        if bw <= 0:
            raise ValueError("bandwidth must be positive.")
        return bw

    def _validate_mode(self) -> str:
        if self.mode not in _MODE_ALIASES:
            raise ValueError(
                f"Unsupported mode='{self.mode}'. Expected one of ['sd_kde', 'kde']."
            )
        return _MODE_ALIASES[self.mode]

    def _check_is_fitted(self) -> None:
        if not self._is_fitted:
            raise ValueError("Estimator is not fitted. Call fit(X) before score_samples(X).")
