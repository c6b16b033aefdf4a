"""Seeded synthetic workloads (host, numpy).

Gaussian mixtures are sampled the way the reference's error suite does it
(experiments/error_suite_16d/truth.py:57-64, datasets.py:34-42): component
means from default_rng(seed), then a fresh default_rng(seed) draws the
component labels and the unit normals; queries use seed + 1.

Workloads of the benchmark and the parity tests:
  plain_normal_16d   X_train (n, 16), X_query (m, 16) i.i.d. N(0, I) from one
                     default_rng(seed), train first (example.py:26,59-60 style).
  gmm_isotropic      K components, means ~ N(0, scale^2 I), sigma = 1, equal weights.
  gmm_full_cov       K components with full covariances (eigenvalues
                     log-uniform [0.1, 3], random orthogonal rotation),
                     means ~ N(0, scale^2 I).
  uneven_sizes       dataset sizes n_i = round(exp(U[log lo, log hi])).
"""

from __future__ import annotations

import numpy as np

D = 16


def plain_normal_16d(n: int, m: int, seed: int = 0):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, D)).astype(np.float32)
    q = rng.standard_normal((m, D)).astype(np.float32)
    return x, q


def gmm_means(k: int, scale: float, seed: int, d: int = D) -> np.ndarray:
    return np.random.default_rng(seed).normal(size=(k, d)) * scale


def gmm_isotropic(n: int, *, k: int = 8, mean_scale: float = 3.0, sigma: float = 1.0,
                  seed: int = 0, means: np.ndarray | None = None) -> np.ndarray:
    """n points of an equal-weight isotropic mixture; `seed` picks both the
    means (unless given) and the sample stream."""
    if means is None:
        means = gmm_means(k, mean_scale, seed)
    k = means.shape[0]
    rng = np.random.default_rng(seed)
    comp = rng.choice(k, size=n, p=np.full(k, 1.0 / k))
    x = rng.normal(size=(n, means.shape[1])) * sigma + means[comp]
    return x.astype(np.float32)


def gmm_isotropic_split(n_train: int, n_query: int, *, k: int = 8, mean_scale: float = 3.0,
                        seed: int = 0):
    """Train from seed, queries from seed + 1, same mixture (means from seed)."""
    means = gmm_means(k, mean_scale, seed)
    return (gmm_isotropic(n_train, seed=seed, means=means),
            gmm_isotropic(n_query, seed=seed + 1, means=means), means)


def random_rotation(rng: np.random.Generator, d: int = D) -> np.ndarray:
    a = rng.normal(size=(d, d))
    q, r = np.linalg.qr(a)
    return q * np.sign(np.diag(r))[None, :]


def gmm_full_cov_params(k: int = 32, *, mean_scale: float = 4.0, seed: int = 0,
                        eig_lo: float = 0.1, eig_hi: float = 3.0):
    rng = np.random.default_rng(seed)
    means = rng.normal(size=(k, D)) * mean_scale
    chol = np.empty((k, D, D))
    for c in range(k):
        ev = np.exp(rng.uniform(np.log(eig_lo), np.log(eig_hi), size=D))
        rot = random_rotation(rng)
        chol[c] = rot * np.sqrt(ev)[None, :]  # cov = rot diag(ev) rot^T = chol chol^T
    return means, chol


def gmm_full_cov(n: int, *, params, seed: int = 0) -> np.ndarray:
    """n points of the equal-weight full-covariance mixture `params`."""
    means, chol = params
    k = means.shape[0]
    rng = np.random.default_rng(seed)
    comp = rng.choice(k, size=n, p=np.full(k, 1.0 / k))
    z = rng.normal(size=(n, D))
    out = np.empty((n, D), dtype=np.float32)
    for c in range(k):
        sel = np.nonzero(comp == c)[0]
        out[sel] = (z[sel] @ chol[c].T + means[c]).astype(np.float32)
    return out


def gmm_full_cov_logpdf(x: np.ndarray, params) -> np.ndarray:
    """Exact log-density of the equal-weight full-covariance mixture (FP64)."""
    means, chol = params
    k = means.shape[0]
    x = np.asarray(x, dtype=np.float64)
    logs = np.empty((k, x.shape[0]))
    for c in range(k):
        y = np.linalg.solve(chol[c], (x - means[c]).T)  # L^-1 (x - mu)
        logdet = 2.0 * np.sum(np.log(np.abs(np.diag(np.linalg.qr(chol[c])[1]))))
        logs[c] = -0.5 * np.sum(y * y, axis=0) - 0.5 * logdet - 0.5 * D * np.log(2 * np.pi)
    mx = logs.max(axis=0)
    return mx + np.log(np.exp(logs - mx).sum(axis=0)) - np.log(k)


def uneven_sizes(count: int = 16, lo: int = 2048, hi: int = 262144, seed: int = 0) -> list[int]:
    rng = np.random.default_rng(seed)
    return [int(round(v)) for v in np.exp(rng.uniform(np.log(lo), np.log(hi), size=count))]
