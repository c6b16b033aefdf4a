"""Precision forms of the SD-KDE pipeline (no reference counterpart).

The plain FP32 form (the drop-in, flash_sd_kde.kde) evaluates the cross term
q.x and phi.X with FP32 MFMA. The forms here evaluate them with 16-bit MFMA
operands and FP32 accumulation:
  "bf16x6"  BF16 three-piece split of both operands, seven cross-term
            products (every term above FP32 rounding): an FP32-class cross
            term at BF16 MFMA rate; phi.X as in bf16x3
  "fp16x3"  FP16 hi/lo split cross term, three products (11-bit pieces, the
            dropped lo.lo term ~2^-22 relative); phi.X as in bf16x3
  "fp16x3s" the fp16x3 cross term; phi.X with phi rounded once to FP16
            against X in FP16 hi/lo pieces (one MFMA instead of two and no
            phi split; phi to 2^-12 relative). Score forms; its
            log-densities are fp16x3's
  "bf16x3s" the bf16x3 cross term (BF16 range) with fp16x3s's phi.X (the FP16
            pieces there are x - mu itself: |x - mu| < 6e4); log-densities
            are bf16x3's
  "bf16x3"  BF16 hi/lo split, three products (hi.hi + lo.hi + hi.lo)
  "fp16"    single FP16 operands
  "bf16"    single BF16 operands
  "fp32", "fp32_valu"  FP32 MFMA / FP32 VALU (16 FMAs per pair) cross term;
            the VALU score also forms phi.X with FMAs (one thread per row)
with points centred on the data mean, log-densities by log-sum-exp
(per-split running maximum, FP64 combine) and FP64 debias. The score and the
log-density may use different forms (`sdkde_logdensity(score_precision=...)`).

In the forms whose phi.X takes phi rounded to FP16 ("fp16", "fp16x3s",
"bf16x3s") the score kernel leaves out each row tile x tile pair whose every
term is negligible (exactly 0 in FP16 for "fp16": bitwise the dense result;
below n 2^-28 of a row's self term in the FP32 row sums of the other two).
`emp_score` therefore runs these forms on the data in spatial order
(`spatial.spatial_order`) and returns the results in the caller's order; the
order changes the sums only by rounding. The 16-bit log-densities leave out
terms below 2^-(ceil(log2 n) + 28) of a query's sum (bounded by a first pass
over a sample of the data), so `kde_logdensity` and `sdkde_logdensity` put
data and queries in spatial order for them as well.

`ref64_*` are FP64 direct-difference GPU kernels used to measure the error of
every form at sizes the CPU oracle cannot reach.
"""

from __future__ import annotations

import torch

from flash_sd_kde import _ops
from flash_sd_kde.spatial import spatial_order

# forms whose score kernel leaves out negligible tiles, forms whose
# log-density does, and the smallest data set that is put in spatial order
# for them (below it the kernels have few tiles to leave out)
SKIP_PRECISIONS = ("fp16", "fp16x3s", "bf16x3s")
LSE_SKIP_PRECISIONS = ("fp16", "bf16", "fp16x3", "bf16x3", "bf16x6", "fp16x3s", "bf16x3s")
SPATIAL_MIN_N = 8192

SCORE_PRECISIONS = ("fp32", "fp32_valu", "fp16x3", "fp16x3s", "bf16x6", "bf16x3", "bf16x3s",
                    "fp16", "bf16")
LOGDENSITY_PRECISIONS = ("fp32", "fp32_valu", "fp16x3", "fp16x3s", "bf16x6", "bf16x3", "bf16x3s",
                         "fp16", "bf16")


def _logdensity(train: torch.Tensor, query: torch.Tensor, bandwidth: float, precision: str,
                spatial: bool, dense: bool, train_ordered: bool = False) -> torch.Tensor:
    if (spatial and not dense and precision in LSE_SKIP_PRECISIONS
            and train.shape[0] >= SPATIAL_MIN_N):
        if not train_ordered:
            train = train[spatial_order(train)]
        qperm = spatial_order(query)
        out_s = _ops.kde16_logdensity(train, query[qperm], bandwidth, precision=precision)
        out = torch.empty_like(out_s)
        out[qperm] = out_s
        return out
    return _ops.kde16_logdensity(train, query, bandwidth, precision=precision, dense=dense)


def kde_logdensity(data, queries, bandwidth: float, *, precision: str = "fp32",
                   spatial: bool = True, dense: bool = False, device="cuda") -> torch.Tensor:
    """log p(q) = log(1/(n (2 pi)^8 h^16) sum_j exp(-|q - x_j|^2 / 2h^2)), float64.
    The 16-bit forms run on data and queries in spatial order (spatial=False:
    the caller's order) and leave out negligible terms (dense=True: none)."""
    device = torch.device(device)
    return _logdensity(_ops.as_points16(data, device), _ops.as_points16(queries, device),
                       bandwidth, precision, spatial, dense)


def _ordered_score(train: torch.Tensor, bandwidth: float, precision: str, spatial: bool,
                   dense: bool):
    if spatial and not dense and precision in SKIP_PRECISIONS and train.shape[0] >= SPATIAL_MIN_N:
        perm = spatial_order(train)
        pdf_s, w_s = _ops.score16(train[perm], bandwidth, variant="ordered", precision=precision)
        pdf = torch.empty_like(pdf_s)
        pdf[perm] = pdf_s
        w = torch.empty_like(w_s)
        w[perm] = w_s
        return pdf, w
    return _ops.score16(train, bandwidth, variant="ordered", precision=precision, dense=dense)


def emp_score(data, bandwidth: float, *, precision: str = "fp32", spatial: bool = True,
              dense: bool = False, device="cuda"):
    """(pdf_sum (n,), weighted (n, 16)) by the deterministic ordered kernel.
    spatial=False keeps the caller's order for the forms that leave out
    negligible tiles; dense=True evaluates every tile (both only change
    rounding)."""
    device = torch.device(device)
    return _ordered_score(_ops.as_points16(data, device), bandwidth, precision, spatial, dense)


def sdkde_fit_transform(data, bandwidth: float, *, precision: str = "fp32", eps: float = 1e-12,
                        spatial: bool = True, dense: bool = False,
                        device="cuda") -> torch.Tensor:
    """One-step empirical SD-KDE debiasing with the chosen score precision."""
    device = torch.device(device)
    train = _ops.as_points16(data, device)
    pdf, w = _ordered_score(train, bandwidth, precision, spatial, dense)
    return _ops.debias(train, pdf, w, bandwidth, eps)


def sdkde_logdensity(data, queries, bandwidth: float, *, precision: str = "fp32",
                     score_precision: str | None = None, spatial: bool = True,
                     dense: bool = False, device="cuda") -> torch.Tensor:
    """SD-KDE log-density of `queries`: debias `data` (score in `score_precision`,
    default `precision`), then the log-sum-exp KDE (cross term in `precision`).

    The log-density error is set by the KDE's cross term far more than by the
    score's: the score only enters through the debias shift
    (h^2/2) score = (sum_j phi_ij x_j / sum_j phi_ij - x_i) / 2, a weighted mean
    of neighbours in which independent rounding errors of the phi_ij average
    out. So a single-FP16 score under an FP16-split KDE ("fp16" score,
    "fp16x3" KDE) stays within 1e-3 relative log-density error of FP64 at
    h in {0.5, 1, 2} x Silverman on the 8-component 16-D mixture at
    n = 2^20, where the single-FP16 KDE does not (measured errors by bandwidth:
    3.7e-4, 4.6e-6 and 1.7e-7 against 6.0e-3, 1.1e-3 and 1.6e-4). The score's
    share grows as h shrinks; a single-BF16 score exceeds 1e-3 at 0.5 x h.
    """
    device = torch.device(device)
    sp = precision if score_precision is None else score_precision
    if sp not in SCORE_PRECISIONS:
        raise ValueError(f"score_precision must be one of {SCORE_PRECISIONS}, got {sp!r}")
    train = _ops.as_points16(data, device)
    query = _ops.as_points16(queries, device)
    if (spatial and not dense and train.shape[0] >= SPATIAL_MIN_N
            and (sp in SKIP_PRECISIONS or precision in LSE_SKIP_PRECISIONS)):
        # score, debias and log-density all on the data in spatial order
        xs = train[spatial_order(train)]
        pdf, w = _ops.score16(xs, bandwidth, variant="ordered", precision=sp)
        deb = _ops.debias(xs, pdf, w, bandwidth, 1e-12)
        return _logdensity(deb, query, bandwidth, precision, True, False, train_ordered=True)
    deb = sdkde_fit_transform(train, bandwidth, precision=sp, spatial=spatial, dense=dense,
                              device=device)
    return _logdensity(deb, query, bandwidth, precision, spatial, dense)


def ref64_sdkde_logdensity(data, queries, bandwidth: float, *, eps: float = 1e-12,
                           device="cuda") -> torch.Tensor:
    """FP64 reference of sdkde_logdensity (FP64 score, debias and log-sum-exp)."""
    device = torch.device(device)
    x = torch.as_tensor(data, device=device).to(torch.float64)
    q = torch.as_tensor(queries, device=device).to(torch.float64)
    pdf, w = _ops.ref64_score(x, bandwidth)
    deb = x + 0.5 * (w / (pdf[:, None] + eps) - x)
    return _ops.ref64_logdensity(deb, q, bandwidth)


def rel_log_error(logp: torch.Tensor, ref: torch.Tensor) -> tuple[float, float]:
    """(max, median) of |logp - ref| / |ref|."""
    r = (logp.double() - ref.double()).abs() / ref.double().abs().clamp_min(1e-300)
    return float(r.max()), float(r.median())
