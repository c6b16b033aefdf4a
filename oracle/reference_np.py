"""FP32 numpy restatement of the reference's CPU functions (test oracle).

Each function follows the reference's operation order so that outputs agree
with it to FP32 rounding; large inputs are processed in row blocks (every
output row depends only on its own query row, so blocking does not change
the math). Citations are to the reference's flash_sd_kde/reference.py.
"""

from __future__ import annotations

import math

import numpy as np

EPS = 1e-12


def _f32(a) -> np.ndarray:
    return np.asarray(a, dtype=np.float32)


def silverman_bandwidth(data) -> float:
    """reference.py:24-36 - 0.9 min(std, IQR/1.34) n^(-1/5) (sigma <= 0 -> 1)."""
    x = _f32(data)
    if x.ndim != 1 or x.size == 0:
        raise ValueError("1-D non-empty data required")
    q75, q25 = np.percentile(x, 75), np.percentile(x, 25)
    s = min(float(np.std(x)), float(q75 - q25) / 1.34)
    s = 1.0 if s <= 0 else s
    return float(0.9 * s * x.size ** (-0.2))


def silverman_bandwidth_nd(data) -> float:
    """reference.py:171-184 - 0.9 mean_d(min(std_d, IQR_d/1.34)) n^(-1/(d+4))."""
    x = _f32(data)
    n, d = x.shape
    hi, lo = np.percentile(x, [75, 25], axis=0)
    s = float(np.mean(np.minimum(np.std(x, axis=0), (hi - lo) / 1.34)))
    s = 1.0 if s <= 0 else s
    return float(0.9 * s * n ** (-1.0 / (d + 4)))


def kde_eval_1d_numpy(queries, data, bandwidth: float) -> np.ndarray:
    """reference.py:187-197 - densities, FP32 phi, host-float normalisation."""
    q, x = _f32(queries), _f32(data)
    z = (q[:, None] - x[None, :]) / np.float32(bandwidth)
    phi = np.exp(-0.5 * z * z)
    return (1.0 / (math.sqrt(2.0 * math.pi) * x.size * bandwidth)) * phi.sum(axis=1)


def kde_eval_1d_linearized_numpy(queries, data, bandwidth: float) -> np.ndarray:
    """reference.py:200-213 - phi * (1 + 1/2 - z^2/2)."""
    q, x = _f32(queries), _f32(data)
    z = (q[:, None] - x[None, :]) / np.float32(bandwidth)
    s = z * z
    phi = np.exp(-0.5 * s) * (1.0 + 0.5 - 0.5 * s)
    return (1.0 / (math.sqrt(2.0 * math.pi) * x.size * bandwidth)) * phi.sum(axis=1)


def _dist_nd(q: np.ndarray, x: np.ndarray) -> np.ndarray:
    """max(|q|^2 + |x|^2 - 2 q x^T, 0) in FP32 (reference.py:228-232)."""
    qn = (q * q).sum(axis=1, keepdims=True)
    xn = (x * x).sum(axis=1, keepdims=True).T
    return np.maximum(qn + xn - 2.0 * (q @ x.T), 0.0)


def kde_sums_nd(queries, data, bandwidth: float, *, laplacian: bool = False,
                block: int = 2048) -> np.ndarray:
    """Unnormalised FP32 sums sum_j phi (Laplacian: phi (1 + d/2 - s/2))."""
    q, x = _f32(queries), _f32(data)
    inv_h2 = 1.0 / (bandwidth * bandwidth)
    d = x.shape[1]
    out = np.empty(q.shape[0], dtype=np.float32)
    for a in range(0, q.shape[0], block):
        dist = _dist_nd(q[a:a + block], x)
        if laplacian:
            s = dist * inv_h2
            phi = np.exp(-0.5 * s) * (1.0 + 0.5 * d - 0.5 * s)
        else:
            phi = np.exp(-0.5 * dist * inv_h2)
        out[a:a + block] = phi.sum(axis=1)
    return out


def _nd_norm(bandwidth: float, d: int, n: int) -> float:
    return 1.0 / ((2.0 * math.pi) ** (d / 2.0) * (bandwidth ** d) * n)


def kde_eval_nd_numpy(queries, data, bandwidth: float, *, block: int = 2048) -> np.ndarray:
    """reference.py:216-235 - N-D densities."""
    x = _f32(data)
    return _nd_norm(bandwidth, x.shape[1], x.shape[0]) * kde_sums_nd(queries, x, bandwidth,
                                                                     block=block)


def kde_eval_nd_linearized_numpy(queries, data, bandwidth: float, *, block: int = 2048) -> np.ndarray:
    """reference.py:238-260 - linearized N-D densities."""
    x = _f32(data)
    return _nd_norm(bandwidth, x.shape[1], x.shape[0]) * kde_sums_nd(
        queries, x, bandwidth, laplacian=True, block=block)


def empirical_score_nd_numpy(data, bandwidth: float, *, eps: float = EPS, block: int = 2048):
    """reference.py:263-284 - (pdf_sum (n,), weighted (n,16), score (n,16)), j = i included."""
    x = _f32(data)
    inv_h2 = 1.0 / (bandwidth * bandwidth)
    n = x.shape[0]
    pdf = np.empty(n, dtype=np.float32)
    w = np.empty_like(x)
    for a in range(0, n, block):
        phi = np.exp(-0.5 * _dist_nd(x[a:a + block], x) * inv_h2)
        pdf[a:a + block] = phi.sum(axis=1)
        w[a:a + block] = phi @ x
    score = (w / (pdf[:, None] + eps) - x) * inv_h2
    return pdf, w, score


def score_rows(rows, data, bandwidth: float, *, block: int = 2048):
    """pdf and weighted sums of empirical_score_nd_numpy (reference.py:276-282)
    for a subset of query rows against the whole data set."""
    q, x = _f32(rows), _f32(data)
    inv_h2 = 1.0 / (bandwidth * bandwidth)
    pdf = np.empty(q.shape[0], dtype=np.float32)
    w = np.empty_like(q)
    for a in range(0, q.shape[0], block):
        phi = np.exp(-0.5 * _dist_nd(q[a:a + block], x) * inv_h2)
        pdf[a:a + block] = phi.sum(axis=1)
        w[a:a + block] = phi @ x
    return pdf, w


def empirical_sd_kde_transform_nd_numpy(data, bandwidth: float, *, eps: float = EPS,
                                        block: int = 2048) -> np.ndarray:
    """reference.py:287-295 - x + (h^2/2) score."""
    _, _, score = empirical_score_nd_numpy(data, bandwidth, eps=eps, block=block)
    return _f32(data) + 0.5 * (bandwidth ** 2) * score


def kde_score_eval_1d(x_points, data, bandwidth: float, *, block: int = 4096,
                      eps: float = EPS) -> np.ndarray:
    """reference.py:83-120 - 1-D KDE score d/dx log p at x_points (FP32 blocks)."""
    xq, d = _f32(x_points), _f32(data)
    h = np.float32(bandwidth)
    inv_h = np.float32(1.0) / h
    zs = np.zeros(xq.size, dtype=np.float32)
    dzs = np.zeros(xq.size, dtype=np.float32)
    for a in range(0, d.size, block):
        diff = (xq[:, None] - d[None, a:a + block]) * inv_h
        phi = np.exp(-0.5 * diff * diff)
        zs += phi.sum(axis=1)
        dzs += (-diff * phi).sum(axis=1)
    c = np.float32(1.0 / np.sqrt(2.0 * np.pi))
    pdf = (c / (d.size * h)) * zs
    der = (c / (d.size * h * h)) * dzs
    return der / (pdf + eps)
