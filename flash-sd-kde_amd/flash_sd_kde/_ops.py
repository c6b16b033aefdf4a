"""Tensor-level operations: validate, allocate outputs/workspace with torch,
and enqueue the HIP entry points on torch's current stream.

These are the only callers of the C ABI. Higher layers (kernels/*,
flash_sd_kde.kde, the estimator) mirror the reference's Python signatures on
top of them.
"""

from __future__ import annotations

import math
from typing import Sequence

import torch

from . import _native as nat

ND = 16

PRECISION_FLAGS = {
    # drop-in precision modes of the reference (globals.py:2-3). gfx950 has no
    # TF32 MFMA (the xf32 MFMA of gfx942 is gone), so both modes run the cross
    # term as FP32 MFMA: "fast_tf32" is at least as accurate as the reference's.
    "fp32_ieee": nat.PREC_F32,
    "fast_tf32": nat.PREC_F32,
    # log-density precision forms
    "fp32": nat.PREC_F32,
    "fp32_valu": nat.PREC_F32_VALU,
    "bf16": nat.PREC_BF16,
    "fp16": nat.PREC_F16,
    "bf16x3": nat.PREC_BF16X3,
    "bf16x6": nat.PREC_BF16X6,
    "fp16x3": nat.PREC_F16X3,
    "fp16x3s": nat.PREC_F16X3S,
    "bf16x3s": nat.PREC_BF16X3S,
}


def precision_flag(precision_mode: str) -> int:
    try:
        return PRECISION_FLAGS[precision_mode]
    except KeyError:
        raise ValueError(
            f"precision_mode must be one of {sorted(PRECISION_FLAGS)}, got {precision_mode}."
        ) from None


def check_finite(t: torch.Tensor, what: str = "points") -> torch.Tensor:
    """Refuse NaN / inf coordinates. The kernels clamp exponents (exp2 of
    min(t, 0)) with NaN-ignoring min/max, so a NaN point would otherwise turn
    into a finite kernel value and silently change every density (the
    reference returns NaN there). Skipped under HIP graph capture, where the
    check's device-to-host read cannot run; the caller vouches for the data."""
    if t.is_cuda and torch.cuda.is_current_stream_capturing():
        return t
    if t.numel() and not bool(torch.isfinite(t).all()):
        raise ValueError(f"{what} contain NaN or infinite values.")
    return t


def as_points16(array_like, device: torch.device) -> torch.Tensor:
    """(n, 16) float32 contiguous tensor on `device` (reference _to_matrix_tensor,
    kernels/flash_sd_kde/common.py:28-43)."""
    if isinstance(array_like, torch.Tensor):
        t = array_like.to(device=device, dtype=torch.float32, copy=False)
    else:
        t = torch.as_tensor(array_like, dtype=torch.float32, device=device)
    if t.ndim != 2 or t.shape[1] != ND:
        raise ValueError(f"expected tensor with shape (n, {ND}), got {tuple(t.shape)}")
    return check_finite(t.contiguous())


def as_points1(array_like, device: torch.device) -> torch.Tensor:
    if isinstance(array_like, torch.Tensor):
        t = array_like.to(device=device, dtype=torch.float32, copy=False)
    else:
        t = torch.as_tensor(array_like, dtype=torch.float32, device=device)
    if t.ndim != 1:
        raise ValueError("expected 1D tensor for KDE eval.")
    return check_finite(t.contiguous())


def _check_h(bandwidth: float) -> float:
    h = float(bandwidth)
    if not (h > 0.0):
        raise ValueError("bandwidth must be positive.")
    return h


def kde16_sums(train: torch.Tensor, query: torch.Tensor, bandwidth: float, *,
               precision_mode: str = "fp32_ieee", laplacian: bool = False,
               atomic: bool = False, moment: bool = False) -> torch.Tensor:
    """sum_j exp(-|q_i - x_j|^2 / 2h^2) for each query (unnormalised); with
    `moment`, sum_j phi |q_i - x_j|^2 / h^2 instead."""
    h = _check_h(bandwidth)
    dev = train.device
    nat.require_gpu(dev, "kde16_sums")
    n, m = train.shape[0], query.shape[0]
    if n == 0 or m == 0:
        raise ValueError("data and queries must be non-empty.")
    flags = precision_flag(precision_mode) | (nat.FLAG_LAPLACIAN if laplacian else 0) \
        | (nat.FLAG_PHI_S if moment else 0)
    if (flags & 0xFF) != nat.PREC_F32:
        raise ValueError(f"precision_mode {precision_mode} is only available for log-densities")
    out = torch.empty(m, dtype=torch.float32, device=dev)
    op = nat.OP_KDE16_ATOMIC if atomic else nat.OP_KDE16_SPLITK
    ws, nbytes = nat.workspace(op, n, m, flags, dev)
    fn = nat.lib().fsk_kde16_atomic if atomic else nat.lib().fsk_kde16_splitk
    st = fn(train.data_ptr(), train.stride(0), n, query.data_ptr(), query.stride(0), m, h,
            flags, ws.data_ptr(), nbytes, out.data_ptr(), nat.stream_handle(dev))
    nat.check(st, "fsk_kde16_atomic" if atomic else "fsk_kde16_splitk")
    return out


FP16_SAFE = 3.0e4  # largest magnitude an FP16 record value may take (FP16 max 65504)


def check_fp16_range(train: torch.Tensor, query: torch.Tensor, h: float, flags: int) -> None:
    """The FP16 forms store the centred coordinates, prescaled by s =
    2^round(-log2 h), as FP16 pieces: x' = (x - mu) s on the data side,
    -2c' (q - mu) s on the query side and the norm terms c |x - mu|^2
    (c = log2(e) / (2 h^2) in magnitude, c' = c / s^2); refuse inputs whose
    values would leave FP16's range instead of returning inf / NaN. Skipped
    while a HIP graph is being captured (it needs a device-to-host read)."""
    if flags not in (nat.PREC_F16, nat.PREC_F16X3, nat.PREC_F16X3S, nat.PREC_BF16X3S):
        return
    if train.is_cuda and torch.cuda.is_current_stream_capturing():
        return  # the check reads a device value; under graph capture the caller vouches
    mu = train.mean(0)
    s = 2.0 ** round(-math.log2(h))
    if flags == nat.PREC_BF16X3S:
        # only the phi.X operand is FP16 there: the pieces of (x - mu) s
        amax = float((train - mu).abs().max()) * s
        if amax > FP16_SAFE:
            raise ValueError(f"bf16x3s holds (x - mu) s as FP16 pieces: {amax:.3g} exceeds "
                             f"{FP16_SAFE:g} (points more than {FP16_SAFE:g} bandwidths from their "
                             "mean); use bf16x3 / fp32.")
        return
    r2 = float(((train - mu) ** 2).sum(1).max())
    if query is not train:
        r2 = max(r2, float(((query - mu) ** 2).sum(1).max()))
    c = 1.4426950408889634 / (2.0 * h * h)
    if c * r2 > FP16_SAFE or 2.0 * c * r2 ** 0.5 / s > FP16_SAFE:
        raise ValueError(
            f"FP16 forms need c*max|x-mu|^2 <= {FP16_SAFE:g} (got {c * r2:.3g}): the points are too "
            "far from their mean for this bandwidth; use bf16x3 / fp32.")


def kde16_logdensity(train: torch.Tensor, query: torch.Tensor, bandwidth: float, *,
                     precision: str = "fp32", dense: bool = False) -> torch.Tensor:
    """Normalised log-density by log-sum-exp, float64 (m,). The 16-bit forms
    leave out terms below 2^-(ceil(log2 n) + 28) of a query's sum from
    n = 16384 on; dense=True evaluates every term."""
    h = _check_h(bandwidth)
    dev = train.device
    nat.require_gpu(dev, "kde16_logdensity")
    n, m = train.shape[0], query.shape[0]
    if n == 0 or m == 0:
        raise ValueError("data and queries must be non-empty.")
    flags = precision_flag(precision)
    check_fp16_range(train, query, h, flags)
    if dense:
        flags |= nat.FLAG_DENSE
    out = torch.empty(m, dtype=torch.float64, device=dev)
    ws, nbytes = nat.workspace(nat.OP_KDE16_LOGDENSITY, n, m, flags, dev)
    st = nat.lib().fsk_kde16_logdensity(train.data_ptr(), train.stride(0), n, query.data_ptr(),
                                        query.stride(0), m, h, flags, ws.data_ptr(), nbytes,
                                        out.data_ptr(), nat.stream_handle(dev))
    nat.check(st, "fsk_kde16_logdensity")
    return out


def kde1_sums(train: torch.Tensor, query: torch.Tensor, bandwidth: float, *,
              laplacian: bool = False, atomic: bool = False,
              moment: bool = False) -> torch.Tensor:
    h = _check_h(bandwidth)
    dev = train.device
    nat.require_gpu(dev, "kde1_sums")
    n, m = train.numel(), query.numel()
    if n == 0 or m == 0:
        raise ValueError("data and queries must be non-empty.")
    flags = (nat.FLAG_LAPLACIAN if laplacian else 0) | (nat.FLAG_PHI_S if moment else 0)
    out = torch.empty(m, dtype=torch.float32, device=dev)
    op = nat.OP_KDE1_ATOMIC if atomic else nat.OP_KDE1_SPLITK
    ws, nbytes = nat.workspace(op, n, m, flags, dev)
    fn = nat.lib().fsk_kde1_atomic if atomic else nat.lib().fsk_kde1_splitk
    st = fn(train.data_ptr(), n, query.data_ptr(), m, h, flags, ws.data_ptr(), nbytes,
            out.data_ptr(), nat.stream_handle(dev))
    nat.check(st, "fsk_kde1_atomic" if atomic else "fsk_kde1_splitk")
    return out


def kde1_score_sums(train: torch.Tensor, query: torch.Tensor, bandwidth: float):
    """(sum_j phi(z_ij), sum_j -z_ij phi(z_ij)) with z = (q - x)/h, phi = exp(-z^2/2)."""
    h = _check_h(bandwidth)
    dev = train.device
    nat.require_gpu(dev, "kde1_score_sums")
    n, m = train.numel(), query.numel()
    if n == 0 or m == 0:
        raise ValueError("data and queries must be non-empty.")
    pdf = torch.empty(m, dtype=torch.float32, device=dev)
    der = torch.empty(m, dtype=torch.float32, device=dev)
    ws, nbytes = nat.workspace(nat.OP_KDE1_SCORE, n, m, 0, dev)
    st = nat.lib().fsk_kde1_score(train.data_ptr(), n, query.data_ptr(), m, h, ws.data_ptr(),
                                  nbytes, pdf.data_ptr(), der.data_ptr(),
                                  nat.stream_handle(dev))
    nat.check(st, "fsk_kde1_score")
    return pdf, der


def score16(train: torch.Tensor, bandwidth: float, *, variant: str = "symmetric",
            transposed_w: bool = False, precision: str = "fp32", dense: bool = False):
    """Empirical-score accumulators (pdf (n,), weighted (n,16)).

    transposed_w=True stores the weighted sums in a (16, n) buffer and returns
    its transpose, the layout of the reference's flash backend
    (kernels/flash_sd_kde/empirical.py:180,223). dense=True evaluates every
    tile pair in the forms that otherwise leave out negligible ones (fp16,
    fp16x3s, bf16x3s)."""
    h = _check_h(bandwidth)
    dev = train.device
    nat.require_gpu(dev, "score16")
    n = train.shape[0]
    if n == 0:
        raise ValueError("data must be non-empty.")
    pdf = torch.empty(n, dtype=torch.float32, device=dev)
    if transposed_w:
        wbuf = torch.empty((ND, n), dtype=torch.float32, device=dev)
        ld_row, ld_col = 1, n
    else:
        wbuf = torch.empty((n, ND), dtype=torch.float32, device=dev)
        ld_row, ld_col = ND, 1
    if variant == "symmetric":
        op, fn, name = nat.OP_SCORE16_SYM, nat.lib().fsk_score16_sym_atomic, "fsk_score16_sym_atomic"
    elif variant == "ordered":
        op, fn, name = nat.OP_SCORE16_ORDERED, nat.lib().fsk_score16_ordered, "fsk_score16_ordered"
    else:
        raise ValueError(f"unknown score variant {variant}")
    flags = precision_flag(precision)
    if flags not in (nat.PREC_F32, nat.PREC_F32_VALU, nat.PREC_BF16, nat.PREC_F16,
                     nat.PREC_BF16X3, nat.PREC_BF16X6, nat.PREC_F16X3, nat.PREC_F16X3S,
                     nat.PREC_BF16X3S):
        raise ValueError(f"precision {precision} is not available for the score pass")
    if flags != nat.PREC_F32 and variant != "ordered":
        raise ValueError("reduced-precision score accumulators use the ordered kernel")
    check_fp16_range(train, train, h, flags)
    if dense and variant == "ordered":
        flags |= nat.FLAG_DENSE
    ws, nbytes = nat.workspace(op, n, n, flags, dev)
    st = fn(train.data_ptr(), train.stride(0), n, h, flags, ws.data_ptr(), nbytes, pdf.data_ptr(),
            wbuf.data_ptr(), ld_row, ld_col, nat.stream_handle(dev))
    nat.check(st, name)
    return pdf, (wbuf.t() if transposed_w else wbuf)


def score16_rows(rows: torch.Tensor, train: torch.Tensor, bandwidth: float, *,
                 precision: str = "fp32", dense: bool = False, rows_in_data: bool = False):
    """(pdf (m,), weighted (m,16)) of `rows` under the KDE of `train`, in the
    FP32 form or one of the 16-bit MFMA forms of the ordered score
    (dense=True: every tile pair evaluated, as in score16). rows_in_data=True
    states that every row is one of the data points (FSK_FLAG_ROWS_IN_DATA)."""
    h = _check_h(bandwidth)
    dev = train.device
    nat.require_gpu(dev, "score16_rows")
    n, m = train.shape[0], rows.shape[0]
    if n == 0 or m == 0:
        raise ValueError("rows and data must be non-empty.")
    flags = precision_flag(precision)
    if flags not in (nat.PREC_F32, nat.PREC_F32_VALU, nat.PREC_BF16, nat.PREC_F16,
                     nat.PREC_BF16X3, nat.PREC_BF16X6, nat.PREC_F16X3, nat.PREC_F16X3S,
                     nat.PREC_BF16X3S):
        raise ValueError(f"precision {precision} is not available for the score pass")
    check_fp16_range(train, rows, h, flags)
    if dense:
        flags |= nat.FLAG_DENSE
    if rows_in_data and flags != nat.PREC_F32 and flags != nat.PREC_F32_VALU:
        flags |= nat.FLAG_ROWS_IN_DATA
    pdf = torch.empty(m, dtype=torch.float32, device=dev)
    w = torch.empty((m, ND), dtype=torch.float32, device=dev)
    ws, nbytes = nat.workspace(nat.OP_SCORE16_ROWS, n, m, flags, dev)
    st = nat.lib().fsk_score16_rows(rows.data_ptr(), rows.stride(0), m, train.data_ptr(),
                                    train.stride(0), n, h, flags, ws.data_ptr(), nbytes,
                                    pdf.data_ptr(), w.data_ptr(), ND, 1, nat.stream_handle(dev))
    nat.check(st, "fsk_score16_rows")
    return pdf, w


def spatial_codes16(x: torch.Tensor, bits: int = 10) -> torch.Tensor:
    """Morton codes (int32, n) of the rows of a float32 (n, 16) CUDA tensor
    along their three leading principal axes (fsk_spatial_codes16): no host
    synchronization."""
    n = x.shape[0]
    if x.dim() != 2 or x.shape[1] != ND or x.dtype != torch.float32 or not x.is_cuda:
        raise ValueError("spatial_codes16 takes a float32 (n, 16) CUDA tensor")
    if x.stride(1) != 1 or x.stride(0) < ND:
        x = x.contiguous()
    dev = x.device
    codes = torch.empty(n, dtype=torch.int32, device=dev)
    ws, nbytes = nat.workspace(nat.OP_SPATIAL_CODES16, n, 0, 0, dev)
    st = nat.lib().fsk_spatial_codes16(x.data_ptr(), x.stride(0), n, bits, codes.data_ptr(),
                                       ws.data_ptr(), nbytes, nat.stream_handle(dev))
    nat.check(st, "fsk_spatial_codes16")
    return codes


def debias(train: torch.Tensor, pdf: torch.Tensor, weighted: torch.Tensor, bandwidth: float,
           eps: float, *, return_score: bool = False):
    """x + (h^2/2) * ((w/(pdf+eps) - x)/h^2), evaluated in FP64 on the GPU."""
    h = _check_h(bandwidth)
    dev = train.device
    n = train.shape[0]
    if pdf.dtype != torch.float32 or weighted.dtype != torch.float32:
        raise ValueError("pdf and weighted sums must be float32")
    pdf = pdf.contiguous()
    out = torch.empty((n, ND), dtype=torch.float32, device=dev)
    score = torch.empty((n, ND), dtype=torch.float32, device=dev) if return_score else None
    st = nat.lib().fsk_sdkde_debias_f64(
        train.data_ptr(), train.stride(0), n, pdf.data_ptr(), weighted.data_ptr(),
        weighted.stride(0), weighted.stride(1), h, float(eps), out.data_ptr(), ND,
        nat.ptr(score), nat.stream_handle(dev))
    nat.check(st, "fsk_sdkde_debias_f64")
    return (out, score) if return_score else out


# ---- any dimension d <= 64 (zero-padded 16-wide slices inside the library) ----

def as_points_nd(array_like, device: torch.device) -> torch.Tensor:
    """(rows, d) float32 tensor on `device` with unit feature stride."""
    if isinstance(array_like, torch.Tensor):
        t = array_like.to(device=device, dtype=torch.float32, copy=False)
    else:
        t = torch.as_tensor(array_like, dtype=torch.float32, device=device)
    if t.ndim != 2 or t.shape[0] == 0 or not 1 <= t.shape[1] <= nat.ND_MAX_DIM:
        raise ValueError(f"expected a non-empty (rows, d) array with 1 <= d <= {nat.ND_MAX_DIM}, "
                         f"got shape {tuple(t.shape)}")
    return check_finite(t if t.stride(1) == 1 else t.contiguous())


def kde_nd_sums(train: torch.Tensor, query: torch.Tensor, bandwidth: float, *,
                laplacian: bool = False) -> torch.Tensor:
    """sum_j exp(-|q_i - x_j|^2 / 2h^2) in d dimensions (unnormalised)."""
    h = _check_h(bandwidth)
    dev = train.device
    nat.require_gpu(dev, "kde_nd_sums")
    n, d = train.shape
    m = query.shape[0]
    if query.shape[1] != d:
        raise ValueError(f"data has {d} features but queries have {query.shape[1]}.")
    flags = nat.FLAG_LAPLACIAN if laplacian else 0
    out = torch.empty(m, dtype=torch.float32, device=dev)
    nbytes = int(nat.lib().fsk_nd_workspace_bytes(nat.OP_KDE_ND_SPLITK, n, m, d))
    ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=dev)
    st = nat.lib().fsk_kde_nd_splitk(train.data_ptr(), train.stride(0), n, query.data_ptr(),
                                     query.stride(0), m, d, h, flags, ws.data_ptr(), nbytes,
                                     out.data_ptr(), nat.stream_handle(dev))
    nat.check(st, "fsk_kde_nd_splitk")
    return out


def score_nd(train: torch.Tensor, bandwidth: float):
    """(pdf (n,), weighted (n, d)) empirical-score accumulators in d dimensions."""
    h = _check_h(bandwidth)
    dev = train.device
    nat.require_gpu(dev, "score_nd")
    n, d = train.shape
    pdf = torch.empty(n, dtype=torch.float32, device=dev)
    w = torch.empty((n, d), dtype=torch.float32, device=dev)
    nbytes = int(nat.lib().fsk_nd_workspace_bytes(nat.OP_SCORE_ND_ORDERED, n, n, d))
    ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=dev)
    st = nat.lib().fsk_score_nd_ordered(train.data_ptr(), train.stride(0), n, d, h, 0,
                                        ws.data_ptr(), nbytes, pdf.data_ptr(), w.data_ptr(), d, 1,
                                        nat.stream_handle(dev))
    nat.check(st, "fsk_score_nd_ordered")
    return pdf, w


def debias_nd(train: torch.Tensor, pdf: torch.Tensor, weighted: torch.Tensor, bandwidth: float,
              eps: float) -> torch.Tensor:
    """x + (h^2/2) ((w/(pdf+eps) - x)/h^2) in d dimensions, FP64 on the GPU."""
    h = _check_h(bandwidth)
    dev = train.device
    n, d = train.shape
    out = torch.empty((n, d), dtype=torch.float32, device=dev)
    st = nat.lib().fsk_sdkde_debias_nd_f64(train.data_ptr(), train.stride(0), n, d,
                                           pdf.contiguous().data_ptr(), weighted.data_ptr(),
                                           weighted.stride(0), weighted.stride(1), h, float(eps),
                                           out.data_ptr(), d, nat.stream_handle(dev))
    nat.check(st, "fsk_sdkde_debias_nd_f64")
    return out


def reduce_partials(partials: torch.Tensor) -> torch.Tensor:
    if partials.ndim != 2:
        raise ValueError("partials must be 2D (n_splits, n_query).")
    nat.require_gpu(partials.device, "reduce_pdf_partials")
    p = partials.to(torch.float32)
    if p.stride(1) != 1:
        p = p.contiguous()
    S, m = p.shape
    out = torch.empty(m, dtype=torch.float32, device=p.device)
    st = nat.lib().fsk_reduce_partials_f64(p.data_ptr(), S, m, p.stride(0), out.data_ptr(),
                                           nat.stream_handle(p.device))
    nat.check(st, "fsk_reduce_partials_f64")
    return out


def reduce_weighted(partials: torch.Tensor) -> torch.Tensor:
    if partials.ndim != 3:
        raise ValueError("partials must be 3D (n_splits, n_query, features).")
    nat.require_gpu(partials.device, "reduce_weighted_partials")
    if partials.shape[2] != ND:
        raise ValueError(f"weighted partials must have {ND} features")
    p = partials.to(torch.float32)
    if p.stride(2) != 1 or p.stride(1) != ND:
        p = p.contiguous()
    S, m, _ = p.shape
    out = torch.empty((m, ND), dtype=torch.float32, device=p.device)
    st = nat.lib().fsk_reduce_weighted_partials_f64(p.data_ptr(), S, m, p.stride(0),
                                                    out.data_ptr(), ND, 1,
                                                    nat.stream_handle(p.device))
    nat.check(st, "fsk_reduce_weighted_partials_f64")
    return out


def infer_ndim(data: Sequence | torch.Tensor) -> int:
    """Reference _infer_ndim, flash_sd_kde/kde.py:36-46."""
    if isinstance(data, torch.Tensor):
        return data.ndim
    import numpy as np

    if isinstance(data, np.ndarray):
        return data.ndim
    if len(data) == 0:
        return 1
    first = data[0]
    if isinstance(first, (list, tuple, np.ndarray)):
        return 2
    return 1


def ref64_score(train64: torch.Tensor, bandwidth: float):
    """FP64 reference accumulators (pdf (n,), w (n,16)), float64 on the GPU."""
    h = _check_h(bandwidth)
    dev = train64.device
    nat.require_gpu(dev, "ref64_score")
    x = train64.to(torch.float64).contiguous()
    n = x.shape[0]
    pdf = torch.empty(n, dtype=torch.float64, device=dev)
    w = torch.empty((n, ND), dtype=torch.float64, device=dev)
    st = nat.lib().fsk_ref64_score16(x.data_ptr(), n, h, pdf.data_ptr(), w.data_ptr(),
                                     nat.stream_handle(dev))
    nat.check(st, "fsk_ref64_score16")
    return pdf, w


def ref64_logdensity(train64: torch.Tensor, query64: torch.Tensor, bandwidth: float):
    """FP64 reference log-densities (m,), float64 on the GPU."""
    h = _check_h(bandwidth)
    dev = train64.device
    nat.require_gpu(dev, "ref64_logdensity")
    x = train64.to(torch.float64).contiguous()
    q = query64.to(torch.float64).contiguous()
    n, m = x.shape[0], q.shape[0]
    out = torch.empty(m, dtype=torch.float64, device=dev)
    nbytes = int(nat.lib().fsk_ref64_workspace_bytes(n, m))
    ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=dev)
    st = nat.lib().fsk_ref64_logdensity16(x.data_ptr(), n, q.data_ptr(), m, h, ws.data_ptr(),
                                          nbytes, out.data_ptr(), nat.stream_handle(dev))
    nat.check(st, "fsk_ref64_logdensity16")
    return out
