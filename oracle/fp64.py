"""FP64 oracle: direct-difference pair sums and log-sum-exp log-densities.

These do not use the norm expansion of the product kernels: distances are
sum_f (q_f - x_f)^2 in FP64, so they check the expansion-and-clamp arithmetic
instead of repeating it. log_density_expansion restates the reference error
suite's FP64 oracle (experiments/error_suite_16d/reference.py:32-82: norm
expansion in FP64, clamp, logsumexp per sample chunk, logaddexp across
chunks, log_const = -log n - d log h - (d/2) log 2 pi).
"""

from __future__ import annotations

import math

import numpy as np


def _sq_dist(q: np.ndarray, x: np.ndarray) -> np.ndarray:
    diff = q[:, None, :] - x[None, :, :]
    return np.einsum("ijk,ijk->ij", diff, diff)


def _blocks(n: int, per: int):
    for a in range(0, n, per):
        yield a, min(n, a + per)


def kde_sums(queries, data, bandwidth: float, *, laplacian: bool = False) -> np.ndarray:
    """sum_j exp(-|q-x|^2/2h^2) (Laplacian: times 1 + d/2 - |q-x|^2/(2h^2)), FP64."""
    q = np.asarray(queries, dtype=np.float64)
    x = np.asarray(data, dtype=np.float64)
    if q.ndim == 1:
        q, x = q[:, None], x[:, None]
    d = x.shape[1]
    out = np.zeros(q.shape[0])
    per_q = max(1, 4_000_000 // max(1, x.shape[0] * d))
    for a, b in _blocks(q.shape[0], per_q):
        s = _sq_dist(q[a:b], x) / (bandwidth * bandwidth)
        phi = np.exp(-0.5 * s)
        if laplacian:
            phi = phi * (1.0 + 0.5 * d - 0.5 * s)
        out[a:b] = phi.sum(axis=1)
    return out


def kde_density(queries, data, bandwidth: float, *, laplacian: bool = False) -> np.ndarray:
    x = np.asarray(data, dtype=np.float64)
    if x.ndim == 1:
        x = x[:, None]
    n, d = x.shape
    norm = 1.0 / ((2.0 * math.pi) ** (d / 2.0) * bandwidth ** d * n)
    return norm * kde_sums(queries, data, bandwidth, laplacian=laplacian)


def log_density(queries, data, bandwidth: float) -> np.ndarray:
    """log p(q) by log-sum-exp over direct FP64 distances."""
    q = np.asarray(queries, dtype=np.float64)
    x = np.asarray(data, dtype=np.float64)
    n, d = x.shape
    log_const = -math.log(n) - d * math.log(bandwidth) - 0.5 * d * math.log(2.0 * math.pi)
    out = np.empty(q.shape[0])
    per_q = max(1, 4_000_000 // max(1, n * d))
    for a, b in _blocks(q.shape[0], per_q):
        e = -0.5 * _sq_dist(q[a:b], x) / (bandwidth * bandwidth)
        mx = e.max(axis=1, keepdims=True)
        out[a:b] = mx[:, 0] + np.log(np.exp(e - mx).sum(axis=1)) + log_const
    return out


def log_density_expansion(queries, data, bandwidth: float, *, query_chunk: int = 1024,
                          sample_chunk: int = 65536) -> np.ndarray:
    """FP64 log-density with the norm expansion and chunked logsumexp/logaddexp
    (the reference error suite's oracle). Scales to n ~ 1e6 with BLAS."""
    q = np.asarray(queries, dtype=np.float64)
    x = np.asarray(data, dtype=np.float64)
    n, d = x.shape
    inv_h2 = 1.0 / (bandwidth * bandwidth)
    log_const = -math.log(n) - d * math.log(bandwidth) - 0.5 * d * math.log(2.0 * math.pi)
    xn_all = (x * x).sum(axis=1)
    out = np.empty(q.shape[0])
    for a, b in _blocks(q.shape[0], query_chunk):
        qq = q[a:b]
        qn = (qq * qq).sum(axis=1, keepdims=True)
        acc = None
        for c0, c1 in _blocks(n, sample_chunk):
            dist = np.maximum(qn + xn_all[None, c0:c1] - 2.0 * (qq @ x[c0:c1].T), 0.0)
            e = -0.5 * dist * inv_h2
            mx = e.max(axis=1)
            lse = mx + np.log(np.exp(e - mx[:, None]).sum(axis=1))
            acc = lse if acc is None else np.logaddexp(acc, lse)
        out[a:b] = acc + log_const
    return out


def score_rows(rows, data, bandwidth: float):
    """(pdf (m,), weighted (m,d)) of `rows` against all of `data`: sum_j phi(r_i, x_j)
    and sum_j phi(r_i, x_j) x_j, FP64 direct differences."""
    r = np.asarray(rows, dtype=np.float64)
    x = np.asarray(data, dtype=np.float64)
    m = r.shape[0]
    pdf = np.zeros(m)
    w = np.zeros((m, x.shape[1]))
    per = max(1, 4_000_000 // max(1, x.shape[0] * x.shape[1]))
    for a, b in _blocks(m, per):
        phi = np.exp(-0.5 * _sq_dist(r[a:b], x) / (bandwidth * bandwidth))
        pdf[a:b] = phi.sum(axis=1)
        w[a:b] = phi @ x
    return pdf, w


def score_sums(data, bandwidth: float):
    """(pdf (n,), weighted (n,d)) over all ordered pairs incl. j = i, FP64 direct:
    score_rows with the data as its own rows."""
    return score_rows(data, data, bandwidth)


def debias(data, pdf, weighted, bandwidth: float, eps: float = 1e-12):
    """x + (h^2/2) (w/(pdf+eps) - x)/h^2 in FP64; returns (debiased, score)."""
    x = np.asarray(data, dtype=np.float64)
    score = (np.asarray(weighted, np.float64) / (np.asarray(pdf, np.float64)[:, None] + eps) - x) / (
        bandwidth * bandwidth)
    return x + 0.5 * bandwidth * bandwidth * score, score


def loo_loglik(data, bandwidths) -> np.ndarray:
    """Leave-one-out mean log-likelihood per bandwidth:
    mean_i log( 1/(n-1) sum_{j != i} N(x_i; x_j, h^2 I) ), FP64, brute force
    over all pairs with a per-row log-sum-exp (the direct sum underflows double
    precision for small h)."""
    x = np.asarray(data, dtype=np.float64)
    n, d = x.shape
    d2 = _sq_dist(x, x)
    np.fill_diagonal(d2, np.inf)
    out = []
    for h in bandwidths:
        e = -0.5 * d2 / (h * h)
        mx = e.max(axis=1, keepdims=True)
        lse = mx[:, 0] + np.log(np.exp(e - mx).sum(axis=1))
        const = -math.log(n - 1) - 0.5 * d * math.log(2.0 * math.pi) - d * math.log(h)
        out.append(float(np.mean(lse + const)))
    return np.asarray(out)


def rel_log_error(logp, logp_ref):
    """|logp - logp_ref| / |logp_ref| (max, median)."""
    logp = np.asarray(logp, np.float64)
    ref = np.asarray(logp_ref, np.float64)
    r = np.abs(logp - ref) / np.maximum(np.abs(ref), 1e-300)
    return float(np.max(r)), float(np.median(r))
