"""Host-side API behaviour that needs no GPU: validation, error paths,
bandwidth rules, backend selection (the reference's error behaviour in
flash_sd_kde/kde.py:64-71, estimator.py:93-157, config.py:46-75)."""

import numpy as np
import pytest
import torch

import flash_sd_kde
from flash_sd_kde import bandwidth as bw
from flash_sd_kde.config import (
    select_emp_score_backend,
    validate_emp_score_backend,
    validate_kde_backend,
    validate_precision_mode,
)
from globals import EMP_SCORE_BACKEND_FLASH_SD_KDE


def test_validators():
    validate_precision_mode("fast_tf32")
    validate_precision_mode("fp32_ieee")
    with pytest.raises(ValueError):
        validate_precision_mode("tf16")
    validate_kde_backend("atomic")
    with pytest.raises(ValueError):
        validate_kde_backend("dense")
    validate_emp_score_backend("ordered_splitk")
    with pytest.raises(ValueError):
        validate_emp_score_backend("nope")
    assert select_emp_score_backend(10) == EMP_SCORE_BACKEND_FLASH_SD_KDE


def test_bandwidth_rules_match_reference(golden):
    x = golden("cfg_plain", "x")
    assert abs(bw.silverman_bandwidth_nd(x) - golden("cfg_plain", "h_silverman")) < 1e-12
    assert abs(bw.scott_mean_std_bandwidth(x) - golden("cfg_plain", "h_baseline")) < 1e-12
    x1 = golden("one_d", "x")
    assert abs(bw.silverman_bandwidth(x1) - golden("one_d", "h_silverman")) < 1e-12
    with pytest.raises(ValueError):
        bw.silverman_bandwidth_nd(np.zeros((0, 16), np.float32))


def test_kde_eval_argument_errors():
    x = np.zeros((4, 16), np.float32)
    with pytest.raises(ValueError, match="bandwidth"):
        flash_sd_kde.kde_eval(x, x, 0.0)
    with pytest.raises(ValueError, match="precision_mode"):
        flash_sd_kde.kde_eval(x, x, 1.0, precision_mode="tf16")
    with pytest.raises(ValueError, match="kde_backend"):
        flash_sd_kde.kde_eval(x, x, 1.0, kde_backend="dense")
    with pytest.raises(ValueError, match="CUDA"):
        flash_sd_kde.kde_eval(x, x, 1.0, device="cpu")
    if not torch.cuda.is_available():
        with pytest.raises(RuntimeError, match="CUDA is not available"):
            flash_sd_kde.kde_eval(x, x, 1.0)
    with pytest.raises(ValueError, match="bandwidth"):
        flash_sd_kde.emp_sd_kde_fit_transform(x, -1.0)
    with pytest.raises(ValueError, match="emp_score_backend"):
        flash_sd_kde.emp_sd_kde_fit_transform(x, 1.0, emp_score_backend="x")


def test_estimator_errors():
    est = flash_sd_kde.FlashSDKDE(mode="kde")
    with pytest.raises(ValueError, match="not fitted"):
        est.score_samples(np.zeros((2, 16)))
    with pytest.raises(ValueError, match="shape"):
        est.fit(np.zeros(5))
    with pytest.raises(ValueError, match="at least one sample"):
        est.fit(np.zeros((0, 16)))
    with pytest.raises(ValueError, match="mode"):
        flash_sd_kde.FlashSDKDE(mode="bogus", device="cpu").fit(np.zeros((3, 16)))
    est = flash_sd_kde.FlashSDKDE(mode="kde", bandwidth=-1.0)
    with pytest.raises(ValueError, match="bandwidth"):
        est.fit(np.zeros((3, 16)))
    assert est.bandwidth_ is None and est.X_fit_ is None  # a failed fit sets nothing
    with pytest.raises(ValueError, match="bandwidth"):
        flash_sd_kde.FlashSDKDE(bandwidth=-1.0).fit(np.zeros((3, 16)))
    with pytest.raises(ValueError, match="Unsupported bandwidth"):
        flash_sd_kde.FlashSDKDE(bandwidth="scott").fit(np.zeros((3, 16)))


def test_cpu_tensors_are_rejected_not_computed():
    from kernels.reduce_partials import reduce_pdf_partials

    with pytest.raises((ValueError, RuntimeError)):
        reduce_pdf_partials(torch.ones(3, 5))


def test_ndim_rejects_bad_dimensions():
    from flash_sd_kde.ndim import pad16

    with pytest.raises(ValueError, match="dimension"):
        pad16(np.zeros((4, 17), np.float32), "cpu")
    with pytest.raises(ValueError, match="2D"):
        pad16(np.zeros((2, 2, 2), np.float32), "cpu")
    with pytest.raises(ValueError, match="non-empty"):
        pad16(np.zeros((0, 3), np.float32), "cpu")
    with pytest.raises(ValueError, match="positive"):
        from flash_sd_kde.ndim import kde_eval_nd

        kde_eval_nd(np.zeros((4, 3)), np.zeros((2, 3)), 0.0)


def test_wide_dimensions_are_bounded():
    from flash_sd_kde import _native as nat
    from flash_sd_kde.ndim import kde_eval_nd

    with pytest.raises(ValueError, match="dimension"):
        kde_eval_nd(np.zeros((4, 65), np.float32), np.zeros((2, 65), np.float32), 1.0)
    with pytest.raises(ValueError, match="features"):
        kde_eval_nd(np.zeros((4, 20), np.float32), np.zeros((2, 21), np.float32), 1.0)
    lib = nat.load_library()
    assert lib.fsk_nd_workspace_bytes(nat.OP_KDE_ND_SPLITK, 1000, 10, 65) == 0
    assert lib.fsk_nd_workspace_bytes(nat.OP_KDE_ND_SPLITK, 1000, 10, 40) > 1000 * 48 * 4
    # the C entry rejects a stride below the dimension before any GPU work
    st = lib.fsk_kde_nd_splitk(ctypes_ptr(), 30, 10, ctypes_ptr(), 40, 10, 40, 1.0, 0,
                               ctypes_ptr(), 1 << 20, ctypes_ptr(), None)
    assert st == nat.FSK_EINVAL and "stride" in nat.last_error()


def ctypes_ptr():
    import ctypes

    return ctypes.c_void_p(0x1000)


def test_fp16_forms_refuse_out_of_range_inputs():
    from flash_sd_kde import _native as nat
    from flash_sd_kde import _ops

    x = torch.zeros((4, 16))
    x[0, 0] = 100.0  # c |x - mu|^2 ~ 0.72 * 5600 / h^2
    _ops.check_fp16_range(x, x, 1.0, nat.PREC_BF16X3)  # BF16 forms have no such limit
    with pytest.raises(ValueError, match="FP16 forms"):
        _ops.check_fp16_range(x, x, 0.1, nat.PREC_F16X3)
    _ops.check_fp16_range(x, x, 10.0, nat.PREC_F16)  # in range
    _ops.check_fp16_range(x, x, 0.1, nat.PREC_BF16X3S)  # only x - mu is FP16 there
    with pytest.raises(ValueError, match="bf16x3s"):
        y = torch.zeros((4, 16))
        y[0, 0] = 1e5
        _ops.check_fp16_range(y, y, 1.0, nat.PREC_BF16X3S)


def test_mixed_precision_form_rejects_unknown_score_precision():
    from flash_sd_kde import precision

    x = np.zeros((4, 16), np.float32)
    with pytest.raises(ValueError, match="score_precision"):
        precision.sdkde_logdensity(x, x, 1.0, precision="fp16x3", score_precision="fp8",
                                   device="cpu")


def test_bench_precision_forms_name_valid_forms():
    """Every (score, log-density) pair the benchmark times is a form the
    precision module accepts, and the metric's bound is checked at the
    workload's bandwidth and at half and twice it."""
    import bench
    from flash_sd_kde import precision

    names = [f[0] for f in bench.PRECISION_FORMS]
    assert len(names) == len(set(names))
    for _, sp, ep in bench.PRECISION_FORMS:
        assert sp in precision.SCORE_PRECISIONS and ep in precision.LOGDENSITY_PRECISIONS
    assert bench.ERROR_BANDWIDTHS == (0.5, 1.0, 2.0)
    assert bench.score_variant("fp32") == "symmetric" and bench.score_variant("fp16") == "ordered"


def test_non_finite_points_are_refused():
    """NaN / inf coordinates raise ValueError in the input conversion (the
    kernels would otherwise clamp them into finite, meaningless sums)."""
    from flash_sd_kde import _ops

    x = np.zeros((5, 16), np.float32)
    x[2, 7] = np.nan
    with pytest.raises(ValueError, match="NaN or infinite"):
        _ops.as_points16(x, torch.device("cpu"))
    with pytest.raises(ValueError, match="NaN or infinite"):
        _ops.as_points1(np.array([0.0, np.inf], np.float32), torch.device("cpu"))
    with pytest.raises(ValueError, match="NaN or infinite"):
        _ops.as_points_nd(np.full((3, 20), -np.inf, np.float32), torch.device("cpu"))
    _ops.as_points16(np.zeros((5, 16), np.float32), torch.device("cpu"))  # finite passes
