"""Bandwidth selection by held-out likelihood.

select_bandwidth follows the reference's benchmark helper
(benchmarks/mnist_fashion_pca16_ood.py:63-92, `_select_bandwidth`): for each
multiplier, h = base_h * mult, densities of the validation points under the
KDE of the training points via kde_eval, score = mean(log(density + eps))
computed on the host in float32 as there, and the first strictly best h wins
(ties keep the earlier multiplier; base_h is returned if every score is -inf
or NaN).

validation_loglik_grid is the exact form of the same score: each validation
log-density by log-sum-exp in the library's log-density kernel (FP64 result),
so densities far below eps (common in 16-D, where the Gaussian normaliser is
(2 pi)^-8 h^-16) are not floored. select_bandwidth_lse picks the best h on it.
"""

from __future__ import annotations

from typing import Sequence, Tuple

import numpy as np
import torch

from flash_sd_kde import _ops
from flash_sd_kde.kde import kde_eval
from globals import DEFAULT_EPS, DEFAULT_KDE_BACKEND, DEFAULT_PRECISION_MODE


def select_bandwidth(train, val, base_h: float, multipliers: Sequence[float], *,
                     device="cuda", precision_mode: str = DEFAULT_PRECISION_MODE,
                     kde_backend: str = DEFAULT_KDE_BACKEND, use_precomputed_norms: bool = True,
                     autotune: bool = True, eps: float = DEFAULT_EPS,
                     return_scores: bool = False):
    """(best_h, best_ll) over h = base_h * mult; with return_scores also the
    list of (h, ll) in multiplier order."""
    best_h, best_ll = base_h, -np.inf
    scores = []
    for mult in multipliers:
        h = base_h * mult
        dens = kde_eval(train, val, h, device=device, precision_mode=precision_mode,
                        kde_backend=kde_backend, use_precomputed_norms=use_precomputed_norms,
                        autotune=autotune)
        dens_np = dens.detach().cpu().numpy()
        ll = float(np.mean(np.log(dens_np + eps)))
        scores.append((h, ll))
        if ll > best_ll:
            best_ll, best_h = ll, h
    return (best_h, best_ll, scores) if return_scores else (best_h, best_ll)


def validation_loglik_grid(train, val, bandwidths: Sequence[float], *, device="cuda",
                           precision: str = "fp32") -> np.ndarray:
    """Mean validation log-density per bandwidth (float64 numpy, (B,)), each
    log-density exact by log-sum-exp (16-D points)."""
    dev = torch.device(device)
    x = _ops.as_points16(train, dev)
    q = _ops.as_points16(val, dev)
    out = np.empty(len(bandwidths))
    for b, h in enumerate(bandwidths):
        out[b] = float(_ops.kde16_logdensity(x, q, float(h), precision=precision).mean())
    return out


def select_bandwidth_lse(train, val, base_h: float, multipliers: Sequence[float], *,
                         device="cuda", precision: str = "fp32") -> Tuple[float, float]:
    hs = [base_h * m for m in multipliers]
    ll = validation_loglik_grid(train, val, hs, device=device, precision=precision)
    k = int(np.argmax(ll))
    return hs[k], float(ll[k])
