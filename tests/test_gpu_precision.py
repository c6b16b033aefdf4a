"""GPU: FP64 reference kernels against the CPU FP64 oracle, and the
reduced-precision forms (score accumulators and full SD-KDE log-densities)
against FP64. Every case records its error (FSK_METRICS_OUT)."""

import json
import os

import numpy as np
import pytest
import torch

from flash_sd_kde import _ops, precision
from oracle import fp64

pytestmark = pytest.mark.gpu


def record(name, **vals):
    path = os.environ.get("FSK_METRICS_OUT")
    print(name, vals, flush=True)
    if path:
        with open(path, "a") as f:
            f.write(json.dumps({"test": name, **vals}) + "\n")


def test_ref64_kernels_match_cpu_fp64(gpu, golden):
    c = "cfg_gmm_small"
    x, q, h = golden(c, "x"), golden(c, "q"), golden(c, "h")
    xd = torch.as_tensor(x, device=gpu, dtype=torch.float64)
    qd = torch.as_tensor(q[:256], device=gpu, dtype=torch.float64)
    lp = _ops.ref64_logdensity(xd, qd, h).cpu().numpy()
    np.testing.assert_allclose(lp, fp64.log_density(q[:256], x, h), rtol=1e-12)
    pdf, w = _ops.ref64_score(xd[:1500], h)
    pe, we = fp64.score_sums(x[:1500], h)
    np.testing.assert_allclose(pdf.cpu().numpy(), pe, rtol=1e-12)
    np.testing.assert_allclose(w.cpu().numpy(), we, rtol=1e-11, atol=1e-11)


# The cross term q.x enters exp2 unscaled by any normalisation: a relative
# error eps of the 16-bit operands becomes ~eps * |2c q.x| in the exponent
# (|2c q.x| ~ 100 on this mixture), hence the per-form tolerances.
@pytest.mark.parametrize("prec,tol_pdf,tol_w", [("fp32", 2e-5, 2e-5), ("bf16x6", 2e-5, 5e-5),
                                                 ("fp16x3", 2e-5, 5e-5), ("fp16x3s", 5e-5, 1e-4),
                                                 ("bf16x3s", 1e-3, 5e-4),
                                                 ("bf16x3", 1e-3, 5e-4),
                                                 ("fp16", 5e-2, 2e-2), ("bf16", 0.5, 0.2)])
def test_score_precision_forms(gpu, golden, prec, tol_pdf, tol_w):
    c = "cfg_gmm_small"
    x, h = golden(c, "x"), golden(c, "h")
    pdf, w = precision.emp_score(x, h, precision=prec, device=gpu)
    pe, we = fp64.score_sums(x, h)
    e_pdf = float(np.max(np.abs(pdf.cpu().numpy() - pe) / pe))
    e_w = float(np.max(np.abs(w.cpu().numpy() - we)) / np.max(np.abs(we)))
    # score = (w/pdf - x)/h^2 is what the debias uses: error of w/pdf in units of h
    ratio = w.cpu().numpy().astype(np.float64) / pdf.cpu().numpy()[:, None]
    e_mean = float(np.max(np.abs(ratio - we / pe[:, None])) / h)
    record(f"score_{prec}", pdf_max_rel=e_pdf, w_max_rel=e_w, mean_shift_err_over_h=e_mean)
    assert e_pdf <= tol_pdf and e_w <= tol_w


@pytest.mark.parametrize("prec,tol", [("bf16x3", 1e-3), ("bf16x6", 5e-5), ("fp16x3", 5e-5),
                                      ("fp16x3s", 1e-3), ("bf16x3s", 1e-3), ("fp16", 5e-2),
                                      ("bf16", 0.5)])
def test_score_lp_ragged_and_transposed_layout(gpu, prec, tol):
    """The 16-bit kernels take data tiles in pairs: n = 1 and 17 end inside
    the first pair, 300 and 1025 inside a later one, and 5000 runs three data
    splits whose last one ends inside a pair."""
    for n in (1, 17, 300, 1025, 5000):
        x = np.random.default_rng(n).normal(size=(n, 16)).astype(np.float32) + 3.0
        pdf, w = _ops.score16(torch.as_tensor(x, device=gpu), 1.1, variant="ordered",
                              precision=prec, transposed_w=True)
        assert w.stride() == (1, n)
        pe, we = fp64.score_sums(x, 1.1)
        np.testing.assert_allclose(pdf.cpu().numpy(), pe, rtol=tol)
        np.testing.assert_allclose(w.cpu().numpy(), we, rtol=tol, atol=tol * np.abs(we).max())


@pytest.mark.parametrize("prec,tol", [("bf16x3", 1e-3), ("bf16x6", 5e-5), ("fp16x3", 5e-5)])
@pytest.mark.parametrize("m,n", [(1, 5), (37, 1000), (1000, 4097), (700, 1)])
def test_score_rows_lp_forms_match_fp64(gpu, prec, tol, m, n):
    """fsk_score16_rows in the 16-bit forms: rows that are not data points,
    m != n, ragged sizes on both sides."""
    rng = np.random.default_rng(m * 11 + n)
    rows = (rng.normal(size=(m, 16)) + 2.0).astype(np.float32)
    x = (rng.normal(size=(n, 16)) + 2.0).astype(np.float32)
    h = 1.05
    pdf, w = _ops.score16_rows(torch.as_tensor(rows, device=gpu), torch.as_tensor(x, device=gpu),
                               h, precision=prec)
    pe, we = fp64.score_rows(rows, x, h)
    keep = pe > 1e-20  # rows far from every data point have pdf ~ 0 in every form
    np.testing.assert_allclose(pdf.cpu().numpy()[keep], pe[keep], rtol=tol)
    np.testing.assert_allclose(w.cpu().numpy(), we, rtol=tol, atol=tol * np.abs(we).max())


@pytest.mark.parametrize("prec", ["fp16x3", "fp16x3s", "bf16x3s", "bf16x3", "fp16"])
@pytest.mark.parametrize("unit", [1e-6, 1.0, 1e5])
def test_lowp_forms_are_unit_free(gpu, prec, unit):
    """The 16-bit forms prescale the centred points by a power of two chosen
    from the bandwidth, so data in any units (here 1e-6 and 1e5 times the
    unit-scale case, bandwidth scaled alike) reach the same accuracy against
    FP64: without it FP16 pieces of 1e-6-scale data fall into subnormals."""
    rng = np.random.default_rng(17)
    x = (rng.normal(size=(1500, 16)) * unit).astype(np.float32)
    h = 1.1 * unit
    pdf, w = _ops.score16(torch.as_tensor(x, device=gpu), h, variant="ordered", precision=prec)
    pe, we = fp64.score_sums(x.astype(np.float64), h)
    tol = 5e-2 if prec == "fp16" else 1e-3
    np.testing.assert_allclose(pdf.cpu().numpy(), pe, rtol=tol)
    np.testing.assert_allclose(w.cpu().numpy(), we, rtol=tol, atol=tol * np.abs(we).max())
    q = torch.as_tensor(x[:300] + np.float32(0.3 * unit), device=gpu)
    lp = _ops.kde16_logdensity(torch.as_tensor(x, device=gpu), q, h, precision=prec)
    lref = fp64.log_density(x[:300].astype(np.float64) + 0.3 * unit, x.astype(np.float64), h)
    rel = np.abs(lp.cpu().numpy() - lref) / np.abs(lref)
    assert rel.max() < (1e-2 if prec == "fp16" else 1e-4), rel.max()


@pytest.mark.parametrize("prec", ["bf16x3", "fp16x3", "fp16x3s", "bf16x3s", "bf16x6"])
def test_score_rows_on_the_data_equal_ordered(gpu, prec):
    """rows = data runs the same plan and kernel as the ordered score: the
    results are bitwise equal."""
    x = torch.as_tensor(np.random.default_rng(4).normal(size=(3000, 16)).astype(np.float32),
                        device=gpu)
    p1, w1 = _ops.score16(x, 0.9, variant="ordered", precision=prec)
    p2, w2 = _ops.score16_rows(x, x, 0.9, precision=prec)
    assert torch.equal(p1, p2) and torch.equal(w1, w2)


@pytest.mark.parametrize("prec,tol", [("fp32", 1e-5), ("bf16x6", 1e-5), ("fp16x3", 1e-5),
                                      ("fp16x3s", 1e-5), ("bf16x3", 1e-4), ("bf16x3s", 1e-4),
                                      ("fp16", 1e-3),
                                      ("bf16", 1.0)])
def test_sdkde_logdensity_precision_forms(gpu, golden, prec, tol):
    c = "cfg_gmm_small"
    x, q, h = golden(c, "x"), golden(c, "q"), golden(c, "h")
    ref = precision.ref64_sdkde_logdensity(x, q, h, device=gpu)
    lp = precision.sdkde_logdensity(x, q, h, precision=prec, device=gpu)
    mx, med = precision.rel_log_error(lp, ref)
    record(f"sdkde_logdensity_{prec}", max_rel=mx, median_rel=med)
    assert mx <= tol


@pytest.mark.parametrize("score_prec,prec,tol", [("fp16", "fp16x3", 1e-3), ("fp16x3s", "fp32", 1e-5)])
def test_sdkde_logdensity_mixed_forms(gpu, golden, score_prec, prec, tol):
    """Score and KDE in different forms: the single-FP16 score under the
    FP16-split KDE stays inside the 1e-3 log-density bound."""
    c = "cfg_gmm_small"
    x, q, h = golden(c, "x"), golden(c, "q"), golden(c, "h")
    ref = precision.ref64_sdkde_logdensity(x, q, h, device=gpu)
    lp = precision.sdkde_logdensity(x, q, h, precision=prec, score_precision=score_prec, device=gpu)
    mx, med = precision.rel_log_error(lp, ref)
    record(f"sdkde_logdensity_{score_prec}_score_{prec}", max_rel=mx, median_rel=med)
    assert mx <= tol


@pytest.mark.parametrize("n", [1, 17, 1000, 2500])
def test_valu_score_matches_fp64(gpu, n):
    """FP32-VALU score form (cross term and phi.X as FMAs, one thread per row):
    the data set's own accumulators and separate rows, against FP64, to the
    FP32 MFMA form's bounds."""
    rng = np.random.default_rng(40 + n)
    x = (rng.normal(size=(n, 16)) + 3.0).astype(np.float32)
    h = 0.95
    xd = torch.as_tensor(x, device=gpu)
    pdf, w = _ops.score16(xd, h, variant="ordered", precision="fp32_valu")
    pe, we = fp64.score_sums(x, h)
    np.testing.assert_allclose(pdf.cpu().numpy(), pe, rtol=1e-5)
    scale = np.maximum(np.abs(we).max(1, keepdims=True), 1e-30)
    assert np.max(np.abs(w.cpu().numpy() - we) / scale) < 1e-5
    rows = (rng.normal(size=(333, 16)) + 3.0).astype(np.float32)
    pr, wr = _ops.score16_rows(torch.as_tensor(rows, device=gpu), xd, h, precision="fp32_valu")
    pre, wre = fp64.score_rows(rows, x, h)
    np.testing.assert_allclose(pr.cpu().numpy(), pre, rtol=1e-5, atol=1e-30)
    scale = np.maximum(np.abs(wre).max(1, keepdims=True), 1e-30)
    assert np.max(np.abs(wr.cpu().numpy() - wre) / scale) < 1e-5


# ---------------- negligible tiles of the FP16-phi score forms ----------------

def _spread_mixture(gpu, n, seed):
    """The precision-form workload at a smaller n: 8 isotropic clusters in
    16-D (means ~ N(0, 9 I), sigma 1), put in spatial order, with its
    Silverman bandwidth: most 16 x 32 tile pairs lie in different clusters."""
    from flash_sd_kde.bandwidth import silverman_bandwidth_nd
    from flash_sd_kde.datasets import gmm_isotropic
    from flash_sd_kde.spatial import spatial_order

    x = gmm_isotropic(n, k=8, mean_scale=3.0, seed=seed)
    h = silverman_bandwidth_nd(x)
    xd = torch.as_tensor(x, device=gpu)
    return xd[spatial_order(xd)].contiguous(), h


def _negligible_tile_frac(x, h, t_max):
    """Share of (16-row tile, 32-point tile pair) blocks of x whose every
    exponent -|x_i - x_j|^2 log2(e) / (2 h^2) is at or below t_max (FP64)."""
    xs = x.double()
    n = xs.shape[0]
    c = 1.4426950408889634 / (2 * h * h)
    far = total = 0
    for r0 in range(0, n - 15, 16 * 7):  # every 7th row tile
        rows = xs[r0:r0 + 16]
        d2 = torch.cdist(rows, xs) ** 2  # (16, n)
        npair = n // 32
        blk = (-c * d2[:, :npair * 32]).reshape(16, npair, 32).amax(dim=(0, 2))
        far += int((blk <= t_max).sum())
        total += npair
    return far / total


def test_fp16_score_skipping_is_bitwise_dense(gpu):
    """F16 form: a tile pair whose every phi <= 2^-26 adds exact zeros (those
    phi are 0 in FP16, and both the row sums and phi.X take the FP16 phi), so
    leaving such tiles out gives bitwise the dense results. The data are
    chosen so that most tiles are left out; the same holds for rows that are
    not the data (fsk_score16_rows)."""
    x, h = _spread_mixture(gpu, 20000, 5)
    frac = _negligible_tile_frac(x, h, -27.0)
    record("fp16_skip_negligible_tile_frac", frac=frac)
    assert frac > 0.5
    p1, w1 = _ops.score16(x, h, variant="ordered", precision="fp16")
    p2, w2 = _ops.score16(x, h, variant="ordered", precision="fp16", dense=True)
    assert torch.equal(p1, p2) and torch.equal(w1, w2)
    rows = x[::3].contiguous()
    p1, w1 = _ops.score16_rows(rows, x, h, precision="fp16")
    p2, w2 = _ops.score16_rows(rows, x, h, precision="fp16", dense=True)
    assert torch.equal(p1, p2) and torch.equal(w1, w2)


@pytest.mark.parametrize("gap", [5.6, 6.0, 6.4, 7.0, 9.0])
def test_fp16_far_blocks_at_the_threshold_are_bitwise_dense(gpu, gap):
    """Two tight clusters whose closest pairs sit around the F16 threshold
    (c |q - x|^2 = 26 at |q - x| ~ 6 h): blocks the box bound proves far and
    blocks it cannot prove far both end bitwise as in the dense kernel, for
    the data set's own score and for rows that are not the data."""
    rng = np.random.default_rng(int(gap * 10))
    u = rng.normal(size=16)
    u /= np.linalg.norm(u)
    a = rng.normal(size=(3000, 16)) * 0.1 + 0.5 * gap * u
    b = rng.normal(size=(3000, 16)) * 0.1 - 0.5 * gap * u + 40.0 * (gap > 8.0)
    x = torch.as_tensor(np.concatenate([a, b]).astype(np.float32), device=gpu)
    p1, w1 = _ops.score16(x, 1.0, variant="ordered", precision="fp16")
    p2, w2 = _ops.score16(x, 1.0, variant="ordered", precision="fp16", dense=True)
    assert torch.equal(p1, p2) and torch.equal(w1, w2)
    rows = (x[::7] + 0.05).contiguous()
    p1, w1 = _ops.score16_rows(rows, x, 1.0, precision="fp16")
    p2, w2 = _ops.score16_rows(rows, x, 1.0, precision="fp16", dense=True)
    assert torch.equal(p1, p2) and torch.equal(w1, w2)


@pytest.mark.parametrize("shape", ["identical", "line", "plane"])
def test_skipping_on_rank_deficient_data(gpu, shape):
    """Data of rank 0, 1 and 2 leave directions of the far-block frame without
    variance (they come out zero, or as any unit vector orthogonal to the
    others; the bound holds for every orthonormal frame): the FP16 score
    still ends bitwise as the dense kernel, and the 16-bit log-density within
    the dense kernel's rounding."""
    rng = np.random.default_rng(21)
    n = 20000
    base = rng.normal(size=16).astype(np.float32)
    if shape == "identical":
        x = np.tile(base, (n, 1))
    else:
        k = 1 if shape == "line" else 2
        dirs = rng.normal(size=(k, 16))
        x = (base + (rng.normal(size=(n, k)) * 2.0) @ dirs).astype(np.float32)
    x = torch.as_tensor(x, device=gpu)
    h = 1.0
    p1, w1 = _ops.score16(x, h, variant="ordered", precision="fp16")
    p2, w2 = _ops.score16(x, h, variant="ordered", precision="fp16", dense=True)
    assert torch.equal(p1, p2) and torch.equal(w1, w2)
    q = (x[::5] + 0.01).contiguous()
    ld = _ops.kde16_logdensity(x, q, h, precision="fp16x3", dense=True)
    ls = _ops.kde16_logdensity(x, q, h, precision="fp16x3")
    assert bool(torch.isfinite(ls).all())
    diff = float((ls - ld).abs().max())
    record(f"rank_deficient_skip_{shape}", logdensity_vs_dense=diff)
    assert diff <= 2e-6


@pytest.mark.parametrize("prec", ["fp16x3s", "bf16x3s"])
def test_one_phi_score_skipping_bounds(gpu, prec):
    """F16X3S / BF16X3S: phi.X takes the FP16 phi (left-out tiles add exact
    zeros to the centred sums sum phi (x - mu)), the row sums add the
    unrounded phi, of which a left-out tile pair holds at most
    2^-(ceil(log2 n) + 28) per term: pdf within 2^-28 of the dense one plus
    FP32 rounding of the other sums, and w = sum phi (x - mu) + mu pdf
    differs from the dense one by mu times pdf's difference."""
    x, h = _spread_mixture(gpu, 20000, 6)
    t = -(np.ceil(np.log2(x.shape[0])) + 28)
    frac = _negligible_tile_frac(x, h, t - 1.0)
    record(f"{prec}_skip_negligible_tile_frac", frac=frac)
    assert frac > 0.3
    p1, w1 = _ops.score16(x, h, variant="ordered", precision=prec)
    p2, w2 = _ops.score16(x, h, variant="ordered", precision=prec, dense=True)
    rel = float(((p1.double() - p2.double()).abs() / p2.double()).max())
    mu = x.double().mean(0)
    dw = (w1.double() - w2.double()) - mu[None, :] * (p1.double() - p2.double())[:, None]
    rel_w = float((dw.abs().amax(1) / w2.double().abs().amax(1)).max())
    record(f"{prec}_skip_rel_diff", pdf=rel, w_centred=rel_w)
    assert rel <= 1e-6 and rel_w <= 1e-6
    # rows that are not the data run dense: bitwise the dense results
    rows = x[1::2].contiguous()
    p1, w1 = _ops.score16_rows(rows, x, h, precision=prec)
    p2, w2 = _ops.score16_rows(rows, x, h, precision=prec, dense=True)
    assert torch.equal(p1, p2) and torch.equal(w1, w2)
    # unless the caller states that every row is a data point: the same bound
    p3, w3 = _ops.score16_rows(rows, x, h, precision=prec, rows_in_data=True)
    rel = float(((p3.double() - p2.double()).abs() / p2.double()).max())
    dw = (w3.double() - w2.double()) - mu[None, :] * (p3.double() - p2.double())[:, None]
    rel_w = float((dw.abs().amax(1) / w2.double().abs().amax(1)).max())
    record(f"{prec}_rows_in_data_skip_rel_diff", pdf=rel, w_centred=rel_w)
    assert rel <= 1e-6 and rel_w <= 1e-6


@pytest.mark.parametrize("prec,tol_pdf,tol_w", [("fp16x3s", 5e-5, 1e-4), ("bf16x3s", 1e-3, 5e-4),
                                                 ("fp16", 5e-2, 2e-2)])
def test_spatially_ordered_score_matches_fp64(gpu, prec, tol_pdf, tol_w):
    """emp_score puts the data of the FP16-phi forms in spatial order and
    returns the sums in the caller's order: against the FP64 GPU score on
    data in random order, to the forms' bounds, and against the same form in
    the caller's order to rounding."""
    x, h = _spread_mixture(gpu, 12000, 7)
    x = x[torch.randperm(x.shape[0], generator=torch.Generator().manual_seed(0)).to(gpu)]
    x = x.contiguous()
    pdf, w = precision.emp_score(x, h, precision=prec, device=gpu)
    pe, we = _ops.ref64_score(x.double(), h)
    e_pdf = float(((pdf.double() - pe).abs() / pe).max())
    e_w = float((w.double() - we).abs().max() / we.abs().max())
    record(f"spatial_score_{prec}", pdf_max_rel=e_pdf, w_max_rel=e_w)
    assert e_pdf <= tol_pdf and e_w <= tol_w
    p0, w0 = precision.emp_score(x, h, precision=prec, spatial=False, device=gpu)
    assert float(((pdf - p0).abs() / p0).max()) <= tol_pdf


# ---------------- negligible terms of the 16-bit log-densities ----------------

@pytest.mark.parametrize("prec", ["fp16x3", "bf16x6", "fp16"])
def test_logdensity_skipping_bounds(gpu, prec):
    """From n = 16384 on, the 16-bit log-densities bound each query's sum from
    below by a pass over every k-th data point and leave out terms below
    2^-(ceil(log2 n) + 28) of that bound (far blocks by the box bitmap, tiles
    by the kernel's own test): the log-density moves by at most 2^-28 plus
    FP32 rounding of the other terms, against the dense kernel; data and
    queries in spatial order (the Python layer's default) and in random order
    (fewer terms left out, same bound)."""
    from flash_sd_kde.datasets import gmm_isotropic

    x, h = _spread_mixture(gpu, 20000, 8)
    q = torch.as_tensor(gmm_isotropic(4000, k=8, mean_scale=3.0, seed=9), device=gpu)
    ld = _ops.kde16_logdensity(x, q, h, precision=prec, dense=True)
    lp = precision.kde_logdensity(x, q, h, precision=prec, device=gpu)
    lr = _ops.kde16_logdensity(x, q, h, precision=prec)
    d_sp = float((lp - ld).abs().max())
    d_rand = float((lr - ld).abs().max())
    record(f"logdensity_skip_{prec}", spatial=d_sp, caller_order=d_rand)
    # FP32 partial sums run over 512 terms between FP64 flushes: leaving terms
    # out shifts their rounding by ~1e-6 relative
    assert d_sp <= 2e-6 and d_rand <= 2e-6


@pytest.mark.parametrize("prec,tol", [("fp16x3", 1e-5), ("bf16x6", 1e-5)])
def test_spatial_logdensity_matches_fp64(gpu, prec, tol):
    """kde_logdensity of the FP32-class 16-bit forms with terms left out,
    against the FP64 GPU log-density (relative)."""
    from flash_sd_kde.datasets import gmm_isotropic

    x, h = _spread_mixture(gpu, 20000, 10)
    q = torch.as_tensor(gmm_isotropic(3000, k=8, mean_scale=3.0, seed=11), device=gpu)
    lp = precision.kde_logdensity(x, q, h, precision=prec, device=gpu)
    ref = _ops.ref64_logdensity(x.double(), q.double(), h)
    rel = float(((lp - ref).abs() / ref.abs()).max())
    record(f"spatial_logdensity_{prec}", max_rel=rel)
    assert rel <= tol


@pytest.mark.parametrize("n,m", [(16385, 1), (16400, 257), (20001, 3)])
def test_logdensity_skipping_ragged(gpu, n, m):
    """The skipping log-density at ragged sizes: n just past the threshold
    and not a multiple of the 16-point tile or the 32-point pair, a single
    query and partial row blocks; against the dense kernel (the 2^-28 bound
    plus FP32 rounding of the partial sums: a lane's partial runs over up to
    128 terms, at most 128 * 2^-24 = 7.6e-6 relative, and leaving terms out
    reorders them; far queries whose sums hold many comparable terms come
    closest)."""
    from flash_sd_kde.datasets import gmm_isotropic

    x = torch.as_tensor(gmm_isotropic(n, k=8, mean_scale=3.0, seed=n), device=gpu)
    q = torch.as_tensor(gmm_isotropic(m, k=8, mean_scale=3.0, seed=m + 7), device=gpu)
    h = 1.4
    for prec in ("fp16x3", "bf16"):
        lp = precision.kde_logdensity(x, q, h, precision=prec, device=gpu)
        ld = _ops.kde16_logdensity(x, q, h, precision=prec, dense=True)
        assert torch.isfinite(lp).all()
        assert float((lp - ld).abs().max()) <= 1e-5
