"""GPU: held-out-likelihood bandwidth selection (reference
benchmarks/mnist_fashion_pca16_ood.py:63-92) and KDE / SD-KDE for d <= 16 by
zero padding, against the FP64 oracle."""

import math

import numpy as np
import pytest
import torch

from flash_sd_kde.bandwidth import silverman_bandwidth_nd
from flash_sd_kde.datasets import gmm_isotropic_split
from flash_sd_kde.ndim import kde_eval_nd, kde_logdensity_nd, sd_kde_fit_transform_nd
from flash_sd_kde.selection import select_bandwidth, select_bandwidth_lse, validation_loglik_grid
from oracle import fp64

pytestmark = pytest.mark.gpu

MULTS = [0.5, 0.7, 1.0, 1.4, 2.0]


def test_select_bandwidth_matches_oracle(gpu):
    train, val, _ = gmm_isotropic_split(4096, 1024, k=8, mean_scale=3.0, seed=0)
    base = silverman_bandwidth_nd(train)
    best_h, best_ll, scores = select_bandwidth(train, val, base, MULTS, device=gpu,
                                               return_scores=True)
    ref = []
    for mult in MULTS:
        dens = fp64.kde_density(val, train, base * mult)
        ref.append(float(np.mean(np.log(dens + 1e-12))))
    got = [ll for _, ll in scores]
    assert [h for h, _ in scores] == [base * m for m in MULTS]
    # float32 host log/mean of float32 densities vs FP64 throughout
    np.testing.assert_allclose(got, ref, rtol=1e-5)
    k = int(np.argmax(ref))
    assert best_h == base * MULTS[k] and best_ll == got[k]


def test_validation_loglik_grid_matches_fp64(gpu):
    train, val, _ = gmm_isotropic_split(4096, 1024, k=8, mean_scale=3.0, seed=2)
    base = silverman_bandwidth_nd(train)
    hs = [base * m for m in MULTS]
    got = validation_loglik_grid(train, val, hs, device=gpu)
    ref = np.array([fp64.log_density(val, train, h).mean() for h in hs])
    np.testing.assert_allclose(got, ref, rtol=2e-6)
    h, ll = select_bandwidth_lse(train, val, base, MULTS, device=gpu)
    assert h == hs[int(np.argmax(ref))] and ll == pytest.approx(ref.max(), rel=2e-6)


@pytest.mark.parametrize("d", [1, 2, 5, 11, 16])
def test_kde_nd_by_padding_matches_fp64(gpu, d):
    rng = np.random.default_rng(d)
    x = rng.standard_normal((3000, d)).astype(np.float32)
    q = rng.standard_normal((700, d)).astype(np.float32)
    h = 0.9 * float(np.std(x)) * 3000 ** (-1.0 / (d + 4))
    dens = kde_eval_nd(x, q, h, device=gpu).cpu().numpy()
    ref = fp64.kde_density(q, x, h)
    assert np.max(np.abs(dens - ref) / ref) < 5e-5
    # log-density: the FP32 cross term rounds at the scale of |q|^2 / (2 h^2)
    # (about 140 for d = 1, h = 0.18, |q| = 3), an absolute error in log p
    # that a relative bound cannot express where |log p| is near 0
    lp = kde_logdensity_nd(x, q, h, device=gpu).cpu().numpy()
    lref = fp64.log_density(q, x, h)
    assert np.all(np.abs(lp - lref) < 3e-5 + 2e-6 * np.abs(lref))


@pytest.mark.parametrize("d", [2, 7])
def test_sd_kde_nd_by_padding_matches_fp64(gpu, d):
    rng = np.random.default_rng(10 + d)
    x = rng.standard_normal((2500, d)).astype(np.float32)
    h = 0.9 * float(np.std(x)) * 2500 ** (-1.0 / (d + 4))
    deb = sd_kde_fit_transform_nd(x, h, device=gpu).cpu().numpy()
    assert deb.shape == (2500, d)
    pdf, w = fp64.score_sums(x, h)
    ref, _ = fp64.debias(x, pdf, w, h)
    assert np.max(np.abs(deb - ref)) < 2e-5 * max(1.0, float(np.abs(ref).max()))


@pytest.mark.parametrize("mode", ["kde", "sd_kde"])
def test_estimator_other_dimensions(gpu, mode):
    """FlashSDKDE with 5 features (the reference accepts only 16 for SD-KDE
    and 1 or 16 for KDE): log(density + eps) of the padded kernels against the
    FP64 pipeline; scoring twice reuses the device copy of the evaluation set."""
    from flash_sd_kde import FlashSDKDE

    rng = np.random.default_rng(31)
    x = rng.standard_normal((2000, 5)).astype(np.float32)
    q = rng.standard_normal((300, 5)).astype(np.float32)
    est = FlashSDKDE(mode=mode, bandwidth=0.6, device=gpu).fit(x)
    ls = est.score_samples(q)
    assert est.n_features_in_ == 5 and est.X_fit_.shape == (2000, 5)
    data = x
    if mode == "sd_kde":
        pdf, w = fp64.score_sums(x, 0.6)
        data, _ = fp64.debias(x, pdf, w, 0.6)
        data = data.astype(np.float32)
    ref = np.log(fp64.kde_density(q, data, 0.6) + 1e-12)
    assert np.max(np.abs(ls - ref)) < 5e-5
    np.testing.assert_array_equal(est.score_samples(q), ls)


WIDE = [17, 32, 40, 64]


def _wide_case(d, n=3001, m=700, seed=0):
    rng = np.random.default_rng(100 + d + seed)
    x = (rng.standard_normal((n, d)) + 2.0).astype(np.float32)  # off-centre: centring matters
    q = (rng.standard_normal((m, d)) + 2.0).astype(np.float32)
    return x, q, 0.35 * math.sqrt(d)


@pytest.mark.parametrize("d", WIDE)
def test_kde_wide_dimensions_match_fp64(gpu, d):
    """d in 17..64: the library pads to 16-wide slices inside (one MFMA chain
    over all slices); sums against FP64 direct differences, with a ragged n,
    and the Laplacian form with its 1 + d/2 factor."""
    from flash_sd_kde.ndim import kde_sums_nd

    # unnormalised sums: the normalised densities (2 pi)^(-d/2) h^(-d) / n sum
    # fall below FP32's range at d = 64 (1e-58 here), so the comparison is on
    # the sums and on the log-densities
    x, q, h = _wide_case(d)
    sums, n, dd = kde_sums_nd(x, q, h, device=gpu)
    assert (n, dd) == (x.shape[0], d)
    ref = fp64.kde_sums(q, x, h)
    e = float(np.max(np.abs(sums.cpu().numpy() - ref) / ref))
    print(f"kde d={d} rel err vs FP64 {e:.2e}", flush=True)
    assert e < 1e-5
    lap = kde_sums_nd(x, q, h, device=gpu, laplacian=True)[0].cpu().numpy()
    lref = fp64.kde_sums(q, x, h, laplacian=True)
    assert np.max(np.abs(lap - lref)) <= 1e-5 * np.max(np.abs(lref))
    lp = kde_logdensity_nd(x, q, h, device=gpu).cpu().numpy()
    assert np.max(np.abs(lp - fp64.log_density(q, x, h))) < 2e-5


def test_kde_wide_strided_rows(gpu):
    """Rows read through a stride of d + 7 floats (a column slice of a wider
    array) give the contiguous result bit for bit."""
    x, q, h = _wide_case(40, n=999, m=130)
    xb = torch.zeros((999, 47), device=gpu)
    xb[:, :40] = torch.as_tensor(x, device=gpu)
    a = kde_eval_nd(xb[:, :40], q, h, device=gpu)
    b = kde_eval_nd(x, q, h, device=gpu)
    assert torch.equal(a, b)


@pytest.mark.parametrize("d", [17, 40, 64])
def test_sd_kde_wide_dimensions_match_fp64(gpu, d):
    x, _, h = _wide_case(d, n=2003)
    deb = sd_kde_fit_transform_nd(x, h, device=gpu).cpu().numpy()
    assert deb.shape == (2003, d)
    pdf, w = fp64.score_sums(x, h)
    ref, _ = fp64.debias(x, pdf, w, h)
    e = float(np.max(np.abs(deb - ref))) / h
    print(f"sd-kde d={d} debias err vs FP64 {e:.2e} h", flush=True)
    assert e < 1e-5


@pytest.mark.parametrize("mode", ["kde", "sd_kde"])
def test_estimator_wide_dimension(gpu, mode):
    from flash_sd_kde import FlashSDKDE

    x, q, h = _wide_case(40, n=1500, m=200)
    est = FlashSDKDE(mode=mode, bandwidth=h, device=gpu).fit(x)
    ls = est.score_samples(q)
    data = x
    if mode == "sd_kde":
        pdf, w = fp64.score_sums(x, h)
        data = fp64.debias(x, pdf, w, h)[0].astype(np.float32)
    ref = np.log(fp64.kde_density(q, data, h) + 1e-12)
    assert np.max(np.abs(ls - ref)) < 5e-5
