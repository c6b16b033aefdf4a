"""Option names, their validation, and the empirical-score backend choice.

Interface of the reference's flash_sd_kde/config.py (:46-75): the validators
raise ValueError naming the option, the allowed values and the value given;
KernelHeuristics carries the same fields, and select_emp_score_backend makes
the same choice (the flash backend whenever it is the configured default,
else the symmetric kernel up to n_sym_max points and the ordered kernel
above). The reference's Triton tile records and split-count helpers have no
counterpart: the HIP library plans its own tiles and splits from (n, m)
(flash-sd-kde_amd/csrc/fsk_plan.h), which is what keeps the split-K paths
bitwise reproducible.
"""

from __future__ import annotations

from dataclasses import dataclass

from globals import (
    DEFAULT_EMP_SCORE_BACKEND,
    DEFAULT_EPS,
    DEFAULT_KDE_BACKEND,
    DEFAULT_PRECISION_MODE,
    EMP_SCORE_BACKEND_FLASH_SD_KDE,
    EMP_SCORE_BACKEND_ORDERED_SPLITK,
    EMP_SCORE_BACKEND_SYMMETRIC_ATOMIC,
    KDE_BACKEND_ATOMIC,
    KDE_BACKEND_SPLITK_STREAM,
    PRECISION_FAST_TF32,
    PRECISION_FP32_IEEE,
)

# Accepted values of each string option. Both precision modes run the FP32
# MFMA cross term on gfx950 (it has no TF32 MFMA); the names stay so that
# callers passing either keep working.
CHOICES: dict[str, frozenset[str]] = {
    "precision_mode": frozenset({PRECISION_FAST_TF32, PRECISION_FP32_IEEE}),
    "kde_backend": frozenset({KDE_BACKEND_SPLITK_STREAM, KDE_BACKEND_ATOMIC}),
    "emp_score_backend": frozenset({EMP_SCORE_BACKEND_ORDERED_SPLITK,
                                    EMP_SCORE_BACKEND_SYMMETRIC_ATOMIC,
                                    EMP_SCORE_BACKEND_FLASH_SD_KDE}),
}


def check_choice(option: str, value: str) -> None:
    """Raise ValueError unless `value` is one of the accepted values of `option`."""
    allowed = CHOICES[option]
    if value not in allowed:
        raise ValueError(f"{option} must be one of {sorted(allowed)}, got {value}.")


def validate_precision_mode(precision_mode: str) -> None:
    check_choice("precision_mode", precision_mode)


def validate_kde_backend(kde_backend: str) -> None:
    check_choice("kde_backend", kde_backend)


def validate_emp_score_backend(emp_score_backend: str) -> None:
    check_choice("emp_score_backend", emp_score_backend)


@dataclass(frozen=True)
class KernelHeuristics:
    """Defaults a caller may override and pass to the selection rule.

    n_sym_max: largest n for which the symmetric kernel is chosen when the
        configured default is not the flash backend.
    max_partial_weighted_bytes: accepted for callers that set it; the
        library caps its own split-K partial workspace (2 GiB).
    """

    n_sym_max: int = 50_000
    max_partial_weighted_bytes: int = 256 * 1024 * 1024
    eps: float = DEFAULT_EPS
    default_precision_mode: str = DEFAULT_PRECISION_MODE
    default_kde_backend: str = DEFAULT_KDE_BACKEND
    default_emp_score_backend: str = DEFAULT_EMP_SCORE_BACKEND


DEFAULT_HEURISTICS = KernelHeuristics()


def select_emp_score_backend(n: int, *, heuristics: KernelHeuristics = DEFAULT_HEURISTICS) -> str:
    """Backend used when the caller names none."""
    configured = heuristics.default_emp_score_backend
    if configured == EMP_SCORE_BACKEND_FLASH_SD_KDE:
        return configured
    by_size = {True: EMP_SCORE_BACKEND_SYMMETRIC_ATOMIC, False: EMP_SCORE_BACKEND_ORDERED_SPLITK}
    return by_size[n <= heuristics.n_sym_max]
