"""GPU checks of the batched and learned-score configurations at their full
sizes (the benchmark runs them too, but asserts nothing):

* batched LOO bandwidth grid: 16 datasets of log-uniform sizes in
  [2048, 262144] (982 723 points), 4-component mixtures (means ~ N(0, 4 I)),
  32 bandwidths log-spaced in [0.05, 2] x the mean-std rule, one persistent
  launch replayed from a HIP graph with the autotuned launch shape. Every set
  is checked against the ordered score kernel through the identity
  LOO_i = (pdf_sum_i - 1) / (n - 1) at the grid's two widest bandwidths, and
  the four smallest sets against the FP64 brute-force LOO at all 32;
* learned score: n = 2^22 points of the 32-component full-covariance
  mixture, MLP 16-256-256-256-16 trained on exact shifts of 262 144 rows,
  65 536 held-out queries: no fallback, log-densities within 1e-3 of exact
  SD-KDE debiasing, and closer to it than plain KDE.
"""

import json
import os

import numpy as np
import pytest
import torch

from flash_sd_kde import _ops
from flash_sd_kde.bandwidth import scott_mean_std_bandwidth
from oracle import fp64

pytestmark = pytest.mark.gpu


def record(name, **vals):
    print(name, vals, flush=True)
    path = os.environ.get("FSK_METRICS_OUT")
    if path:
        with open(path, "a") as f:
            f.write(json.dumps({"test": name, **vals}) + "\n")


def test_batched_loo_grid_full_configuration(gpu):
    from flash_sd_kde.batched import LooBandwidthGrid
    from flash_sd_kde.datasets import gmm_isotropic, uneven_sizes

    sizes = uneven_sizes(16, 2048, 262144, seed=0)
    sets = [gmm_isotropic(n, k=4, mean_scale=2.0, seed=100 + i) for i, n in enumerate(sizes)]
    factors = np.geomspace(0.05, 2.0, 32)
    grid = LooBandwidthGrid(sets, factors=factors, device=gpu, autotune=True, graph=True)
    eager = grid.run(use_graph=False).cpu().numpy()
    replay = grid.run().cpu().numpy()
    np.testing.assert_allclose(replay, eager, rtol=1e-12)
    assert np.all(np.isfinite(eager))
    # every set: the LOO grid against the score kernel's diagonal-inclusive sums
    worst_id = 0.0
    for k, x in enumerate(sets):
        xd = torch.as_tensor(x, device=gpu)
        n = x.shape[0]
        for b in (30, 31):
            h = float(grid.bandwidths[k, b])
            pdf, _ = _ops.score16(xd, h, variant="ordered")
            s = pdf.double().cpu().numpy() - 1.0
            assert s.min() > 0.05
            log_norm = -np.log(n - 1) - 16 * np.log(h) - 8 * np.log(2 * np.pi)
            ref = float(np.mean(np.log(s) + log_norm))
            e = abs(eager[k, b] - ref) / abs(ref)
            worst_id = max(worst_id, e)
            assert e <= 1e-5, (k, b, e)
    # the four smallest sets against FP64 brute force over the whole grid
    worst64 = 0.0
    for k in np.argsort(sizes)[:4]:
        ref = fp64.loo_loglik(sets[k], grid.bandwidths[k])
        e = float(np.max(np.abs(eager[k] - ref) / np.abs(ref)))
        worst64 = max(worst64, e)
        assert e <= 1e-5, (k, e)
    record("batched_full", sets=len(sets), points=int(sum(sizes)), tune=list(grid.tune),
           identity_rel=worst_id, fp64_rel_smallest4=worst64)


def test_learned_score_full_configuration(gpu):
    from flash_sd_kde.datasets import gmm_full_cov, gmm_full_cov_logpdf, gmm_full_cov_params
    from flash_sd_kde.surrogate import surrogate_fit_transform

    params = gmm_full_cov_params(32, mean_scale=4.0, seed=0)
    x = gmm_full_cov(1 << 22, params=params, seed=0)
    q = gmm_full_cov(65536, params=params, seed=1)
    h = scott_mean_std_bandwidth(x)
    xd = torch.as_tensor(x, device=gpu)
    qd = torch.as_tensor(q, device=gpu)
    deb_s, info = surrogate_fit_transform(xd, h, tol=1e-3)
    assert not info["fallback"], info["holdout_max_rel_logp_err"]
    assert info["holdout_max_rel_logp_err"] <= 0.5e-3
    pdf, w = _ops.score16(xd, h, variant="symmetric")  # exact FP32 score of all n points
    deb_e = _ops.debias(xd, pdf, w, h, 1e-12)
    lp_e = _ops.kde16_logdensity(deb_e, qd, h)
    lp_s = _ops.kde16_logdensity(deb_s, qd, h)
    lp_k = _ops.kde16_logdensity(xd, qd, h)
    r = ((lp_s - lp_e).abs() / lp_e.abs()).cpu().numpy()
    rk = ((lp_k - lp_e).abs() / lp_e.abs()).cpu().numpy()
    truth = gmm_full_cov_logpdf(q, params)
    mae = {k: float(np.mean(np.abs(v.cpu().numpy() - truth)))
           for k, v in (("exact", lp_e), ("surrogate", lp_s), ("plain", lp_k))}
    record("surrogate_full", max_rel_vs_exact=float(r.max()), median_rel_vs_exact=float(np.median(r)),
           plain_max_rel_vs_exact=float(rk.max()), train_s=info["train_s"], mae_vs_truth=mae,
           holdout=info["holdout_max_rel_logp_err"])
    assert r.max() <= 1e-3
    # what the surrogate gains over the method without it (plain KDE): much
    # closer to exact SD-KDE, and closer to the true log-density
    assert np.median(r) < 0.2 * np.median(rk)
    assert mae["surrogate"] < mae["plain"]
