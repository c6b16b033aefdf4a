"""Pins the CPU oracle: reference golden vectors, FP64 closed forms, invariants."""

import math

import numpy as np
import pytest

from oracle import fp64, reference_np as R


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


@pytest.mark.parametrize("case", ["kde16", "kde_chunk"])
def test_kde_nd_matches_reference(golden, case):
    x, q, h = golden(case, "x"), golden(case, "q"), golden(case, "h")
    out = R.kde_eval_nd_numpy(q, x, h)
    assert rel(out, golden(case, "density")) < 2e-6


def test_kde_nd_blocking_does_not_change_rows(golden):
    x, q, h = golden("kde_chunk", "x"), golden("kde_chunk", "q"), golden("kde_chunk", "h")
    a = R.kde_eval_nd_numpy(q, x, h, block=4096)
    b = R.kde_eval_nd_numpy(q, x, h, block=100)
    assert rel(a, b) < 2e-6


@pytest.mark.parametrize("case", ["kde16_lin", "kde16_lin_b"])
def test_kde_nd_linearized_matches_reference(golden, case):
    x, q, h = golden(case, "x"), golden(case, "q"), golden(case, "h")
    assert rel(R.kde_eval_nd_linearized_numpy(q, x, h), golden(case, "density")) < 2e-6


def test_kde_1d_matches_reference(golden):
    x, q, h = golden("kde1d", "x"), golden("kde1d", "q"), golden("kde1d", "h")
    assert rel(R.kde_eval_1d_numpy(q, x, h), golden("kde1d", "density")) < 1e-6
    x, q, h = golden("kde1d_lin", "x"), golden("kde1d_lin", "q"), golden("kde1d_lin", "h")
    assert rel(R.kde_eval_1d_linearized_numpy(q, x, h), golden("kde1d_lin", "density")) < 1e-6


@pytest.mark.parametrize("case", ["score_a", "score_b", "score_large"])
def test_empirical_score_matches_reference(golden, case):
    x, h = golden(case, "x"), golden(case, "h")
    pdf, w, score = R.empirical_score_nd_numpy(x, h)
    assert rel(pdf, golden(case, "pdf")) < 2e-6
    np.testing.assert_allclose(w, golden(case, "weighted"), rtol=2e-5, atol=2e-5)
    np.testing.assert_allclose(score, golden(case, "score"), rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("case", ["score_a", "score_large"])
def test_score_rows_matches_reference_rows(golden, case):
    """A subset of rows against all points = those rows of the reference's
    full empirical-score accumulators (reference.py:276-282)."""
    x, h = golden(case, "x"), golden(case, "h")
    idx = np.random.default_rng(3).choice(x.shape[0], size=min(37, x.shape[0]), replace=False)
    pdf, w = R.score_rows(x[idx], x, h)
    assert rel(pdf, golden(case, "pdf")[idx]) < 2e-6
    np.testing.assert_allclose(w, golden(case, "weighted")[idx], rtol=2e-5, atol=2e-5)
    pdf64, w64 = fp64.score_rows(x[idx], x, h)
    assert rel(pdf64, golden(case, "pdf")[idx]) < 1e-4
    np.testing.assert_allclose(w64, golden(case, "weighted")[idx], rtol=1e-4, atol=2e-5)


def test_transform_matches_reference(golden):
    x, h = golden("transform", "x"), golden("transform", "h")
    np.testing.assert_allclose(R.empirical_sd_kde_transform_nd_numpy(x, h),
                               golden("transform", "debiased"), rtol=1e-5, atol=1e-5)


def test_gmm_small_pipeline_matches_reference(golden):
    c = "cfg_gmm_small"
    x, q, h = golden(c, "x"), golden(c, "q"), golden(c, "h")
    pdf, w, _ = R.empirical_score_nd_numpy(x, h)
    assert rel(pdf, golden(c, "pdf")) < 5e-6
    deb = R.empirical_sd_kde_transform_nd_numpy(x, h)
    np.testing.assert_allclose(deb, golden(c, "debiased"), rtol=1e-5, atol=1e-5)
    assert rel(R.kde_eval_nd_numpy(q, deb, h), golden(c, "density")) < 1e-5
    assert abs(R.silverman_bandwidth_nd(x) - golden(c, "silverman")) < 1e-9


def test_bandwidths_match_reference(golden):
    x = golden("cfg_plain", "x")
    assert abs(R.silverman_bandwidth_nd(x) - golden("cfg_plain", "h_silverman")) < 1e-9
    x1 = golden("one_d", "x")
    assert abs(R.silverman_bandwidth(x1) - golden("one_d", "h_silverman")) < 1e-9


def test_one_d_score_matches_reference(golden):
    x1, h1 = golden("one_d", "x"), golden("one_d", "h_silverman")
    np.testing.assert_allclose(R.kde_score_eval_1d(x1, x1, h1), golden("one_d", "score"),
                               rtol=1e-5, atol=1e-5)


# ---- FP64 oracle: closed forms and agreement with the FP32 reference ----

def test_fp64_single_point_closed_form():
    x = np.random.default_rng(0).normal(size=(1, 16))
    h = 0.7
    # one data point, query on top of it: density = (2 pi)^-8 h^-16
    dens = fp64.kde_density(x, x, h)
    assert abs(dens[0] / ((2 * math.pi) ** -8 * h ** -16) - 1) < 1e-13
    assert abs(fp64.log_density(x, x, h)[0] - math.log((2 * math.pi) ** -8 * h ** -16)) < 1e-12
    pdf, w = fp64.score_sums(x, h)
    assert pdf[0] == 1.0 and np.allclose(w, x)
    deb, score = fp64.debias(x, pdf, w, h)
    assert np.allclose(score, 0) and np.allclose(deb, x)


def test_fp64_two_points_closed_form():
    x = np.zeros((2, 16))
    x[1, 3] = 1.7
    h = 0.9
    phi = math.exp(-(1.7 ** 2) / (2 * h * h))
    pdf, w = fp64.score_sums(x, h)
    np.testing.assert_allclose(pdf, [1 + phi, 1 + phi], rtol=1e-14)
    np.testing.assert_allclose(w[0, 3], phi * 1.7, rtol=1e-14)
    np.testing.assert_allclose(w[1, 3], 1.7, rtol=1e-14)
    # Laplacian factor at distance r: 1 + d/2 - r^2/(2h^2)
    s = fp64.kde_sums(x[:1], x[1:], h, laplacian=True)
    np.testing.assert_allclose(s, [phi * (9 - 1.7 ** 2 / (2 * h * h))], rtol=1e-14)


def test_fp64_identical_points():
    x = np.tile(np.random.default_rng(1).normal(size=(1, 16)), (50, 1))
    pdf, w = fp64.score_sums(x, 1.3)
    np.testing.assert_allclose(pdf, 50.0, rtol=1e-14)
    _, score = fp64.debias(x, pdf, w, 1.3)
    assert np.max(np.abs(score)) < 1e-10


def test_fp64_agrees_with_reference_fp32(golden):
    x, q, h = golden("kde16", "x"), golden("kde16", "q"), golden("kde16", "h")
    assert rel(fp64.kde_density(q, x, h), golden("kde16", "density")) < 2e-5
    c = "cfg_plain"
    x, q, h = golden(c, "x"), golden(c, "q"), golden(c, "h_baseline")
    lp = fp64.log_density(q, x, h)
    lp2 = fp64.log_density_expansion(q, x, h, sample_chunk=1000)
    assert np.max(np.abs(lp - lp2)) < 1e-9
    assert rel(np.exp(lp), golden(c, "density_baseline")) < 2e-5
    pdf, w = fp64.score_sums(golden("score_b", "x"), golden("score_b", "h"))
    assert rel(pdf, golden("score_b", "pdf")) < 2e-5


def test_fp64_loo_matches_bruteforce_loop():
    x = np.random.default_rng(3).normal(size=(9, 16))
    hs = [0.5, 1.0, 2.0]
    got = fp64.loo_loglik(x, hs)
    for k, h in enumerate(hs):
        acc = 0.0
        for i in range(9):
            s = 0.0
            for j in range(9):
                if j != i:
                    s += math.exp(-float(np.sum((x[i] - x[j]) ** 2)) / (2 * h * h))
            acc += math.log(s / (8 * (2 * math.pi) ** 8 * h ** 16))
        assert abs(got[k] - acc / 9) < 1e-10


def test_planted_mistake_is_caught(golden):
    """A dropped clamp or wrong normalisation exponent must fail the golden check."""
    x, q, h = golden("kde16", "x"), golden("kde16", "q"), golden("kde16", "h")
    wrong = R.kde_eval_nd_numpy(q, x, h) * h  # h^-15 instead of h^-16
    assert rel(wrong, golden("kde16", "density")) > 1e-3


def test_fp64_loo_small_bandwidth_is_finite():
    """For h far below the nearest-neighbour distance the direct sum underflows;
    the log-sum-exp form must still give the nearest-neighbour term."""
    x = np.random.default_rng(4).normal(size=(6, 16)) * 3
    h = 0.05
    got = fp64.loo_loglik(x, [h])[0]
    d2 = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d2, np.inf)
    nn = d2.min(axis=1)
    # only the nearest neighbour matters at this h (others are e^-hundreds smaller)
    expect = np.mean(-nn / (2 * h * h) - math.log(5) - 8 * math.log(2 * math.pi) - 16 * math.log(h))
    assert np.isfinite(got) and abs(got - expect) / abs(expect) < 1e-12
