"""Bandwidth rules (host, numpy).

`silverman_bandwidth` / `silverman_bandwidth_nd` reproduce the reference's
rules (flash_sd_kde/reference.py:24-36 and :171-184), which the estimator's
bandwidth="silverman" uses:
    h = 0.9 * mean_d(min(std_d, IQR_d / 1.34)) * n^(-1/(d+4))   (1-D: n^(-1/5)).
`scott_mean_std_bandwidth` is the rule the benchmark workloads are stated in:
    h = mean_d(std_d) * (4/(d+2))^(1/(d+4)) * n^(-1/(d+4)).
"""

from __future__ import annotations

import numpy as np


def silverman_bandwidth(data: np.ndarray) -> float:
    x = np.asarray(data, dtype=np.float32)
    if x.ndim != 1:
        raise ValueError("data must be 1D for silverman_bandwidth.")
    n = x.size
    if n == 0:
        raise ValueError("data must contain at least one element.")
    std_dev = float(np.std(x))
    iqr = float(np.percentile(x, 75) - np.percentile(x, 25))
    sigma = min(std_dev, iqr / 1.34)
    if sigma <= 0:
        sigma = 1.0
    return float(0.9 * sigma * n ** (-1.0 / 5.0))


def silverman_bandwidth_1d(data: np.ndarray) -> float:
    return silverman_bandwidth(data)


def silverman_bandwidth_nd(data: np.ndarray) -> float:
    x = np.asarray(data, dtype=np.float32)
    if x.ndim != 2:
        raise ValueError("data must be 2D for silverman_bandwidth_nd.")
    n, d = x.shape
    if n == 0:
        raise ValueError("data must contain at least one element.")
    std_per_dim = np.std(x, axis=0)
    q75, q25 = np.percentile(x, [75, 25], axis=0)
    sigma = float(np.mean(np.minimum(std_per_dim, (q75 - q25) / 1.34)))
    if sigma <= 0:
        sigma = 1.0
    return float(0.9 * sigma * n ** (-1.0 / (d + 4)))


def scott_mean_std_bandwidth(data: np.ndarray) -> float:
    x = np.asarray(data, dtype=np.float64)
    if x.ndim == 1:
        x = x[:, None]
    n, d = x.shape
    if n == 0:
        raise ValueError("data must contain at least one element.")
    sigma = float(np.mean(np.std(x, axis=0)))
    return float(sigma * (4.0 / (d + 2.0)) ** (1.0 / (d + 4.0)) * n ** (-1.0 / (d + 4.0)))
