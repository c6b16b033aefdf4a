"""ctypes binding of libflashsdkde.so, the gfx950 HIP library (C ABI in
include/flashsdkde.h).

Every public function in this package ends in one of these entry points; there
is no other compute path. If the library is missing, or no GPU is visible,
calls raise instead of falling back to anything else.
"""

from __future__ import annotations

import ctypes
import os
import threading
from typing import Optional

import torch

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_ROOT, "lib", "libflashsdkde.so")

FSK_OK = 0
FSK_EINVAL = -1
FSK_EUNSUPPORTED = -2
FSK_ELAUNCH = -3
FSK_EWORKSPACE = -4

PREC_F32 = 0
PREC_F32_VALU = 1
PREC_BF16 = 2
PREC_F16 = 3
PREC_BF16X3 = 4
PREC_BF16X6 = 5
PREC_F16X3 = 6
PREC_F16X3S = 7
PREC_BF16X3S = 8
FLAG_LAPLACIAN = 1 << 8
FLAG_PHI_S = 1 << 9
FLAG_DENSE = 1 << 10
FLAG_ROWS_IN_DATA = 1 << 11

OP_KDE16_SPLITK = 1
OP_KDE16_ATOMIC = 2
OP_KDE1_SPLITK = 3
OP_KDE1_ATOMIC = 4
OP_SCORE16_SYM = 5
OP_SCORE16_ORDERED = 6
OP_KDE16_LOGDENSITY = 7
OP_KDE1_SCORE = 8
OP_LOO_GRID = 9
OP_SCORE16_ROWS = 10
OP_KDE_ND_SPLITK = 11
OP_SCORE_ND_ORDERED = 12
OP_SPATIAL_CODES16 = 13
ND_MAX_DIM = 64

_p = ctypes.c_void_p
_i64 = ctypes.c_int64
_d = ctypes.c_double
_u = ctypes.c_uint
_sz = ctypes.c_size_t

# name -> argument types (all return int status unless listed in _RESTYPES)
SIGNATURES = {
    "fsk_last_error": [],
    "fsk_version": [],
    "fsk_workspace_bytes": [ctypes.c_int, _i64, _i64, _u],
    "fsk_profile_enable": [ctypes.c_int],
    "fsk_profile_last_ms": [ctypes.c_int],
    "fsk_microbench": [ctypes.c_int, ctypes.c_int, ctypes.c_int, _p, _p, _p],
    "fsk_kde16_splitk": [_p, _i64, _i64, _p, _i64, _i64, _d, _u, _p, _sz, _p, _p],
    "fsk_kde16_atomic": [_p, _i64, _i64, _p, _i64, _i64, _d, _u, _p, _sz, _p, _p],
    "fsk_kde16_splitk_noalloc": [_p, _i64, _i64, _p, _i64, _i64, _d, _u, _p, _p],
    "fsk_kde16_logdensity": [_p, _i64, _i64, _p, _i64, _i64, _d, _u, _p, _sz, _p, _p],
    "fsk_kde1_splitk": [_p, _i64, _p, _i64, _d, _u, _p, _sz, _p, _p],
    "fsk_kde1_atomic": [_p, _i64, _p, _i64, _d, _u, _p, _sz, _p, _p],
    "fsk_kde1_score": [_p, _i64, _p, _i64, _d, _p, _sz, _p, _p, _p],
    "fsk_score16_sym_atomic": [_p, _i64, _i64, _d, _u, _p, _sz, _p, _p, _i64, _i64, _p],
    "fsk_score16_ordered": [_p, _i64, _i64, _d, _u, _p, _sz, _p, _p, _i64, _i64, _p],
    "fsk_score16_rows": [_p, _i64, _i64, _p, _i64, _i64, _d, _u, _p, _sz, _p, _p, _i64, _i64, _p],
    "fsk_reduce_partials_f64": [_p, _i64, _i64, _i64, _p, _p],
    "fsk_spatial_codes16": [_p, _i64, _i64, ctypes.c_int, _p, _p, _sz, _p],
    "fsk_reduce_weighted_partials_f64": [_p, _i64, _i64, _i64, _p, _i64, _i64, _p],
    "fsk_sdkde_debias_f64": [_p, _i64, _i64, _p, _p, _i64, _i64, _d, _d, _p, _i64, _p, _p],
    "fsk_ref64_workspace_bytes": [_i64, _i64],
    "fsk_mlp_linear_fwd": [_p, _i64, _i64, ctypes.c_int, _p, _p, ctypes.c_int, ctypes.c_int, _p,
                           _p, _i64, _p],
    "fsk_mlp_linear_fwd_mse": [_p, _i64, _i64, ctypes.c_int, _p, _p, ctypes.c_int, _p, _d, _p,
                               _p, _p],
    "fsk_mlp_linear_bwd": [_p, _i64, ctypes.c_int, _p, _i64, ctypes.c_int, ctypes.c_int, _p, _p,
                           _p, _p, _p, _p],
    "fsk_mlp_rows_step": [_p, _p],
    "fsk_mlp_wgrad_multi": [_p, ctypes.c_int, _i64, _p],
    "fsk_mlp_linear_bwd_weight": [_p, _i64, ctypes.c_int, _p, _i64, ctypes.c_int, ctypes.c_int,
                                  _p, _p, _p],
    "fsk_mlp_adam": [_p, _p, _p, _p, ctypes.c_int, _i64, ctypes.c_int, ctypes.c_int, _d, _d, _d,
                     _d, ctypes.c_int, _p, _p, _p, _p],
    "fsk_mlp_adam_multi": [_p, ctypes.c_int, _d, _d, _d, _d, ctypes.c_int, _p, _p],
    "fsk_mlp_adam_multi_step": [_p, ctypes.c_int, _d, _d, _d, _d, ctypes.c_int, _p, _p, _p],
    "fsk_mlp_gather_batch": [_p, _p, _i64, ctypes.c_int, ctypes.c_int, _p, _i64, _p, _i64, _p,
                             _p, _p, _p, _p],
    "fsk_mlp_step_advance": [_p, _p],
    "fsk_mlp_cast_bf16": [_p, ctypes.c_int, ctypes.c_int, _p, _p, _p],
    "fsk_mlp_mse": [_p, _p, _i64, _d, _p, _p, _p],
    "fsk_loo_workspace_bytes": [_p, ctypes.c_int, ctypes.c_int],
    "fsk_loo_bandwidth_grid": [_p, _i64, _p, ctypes.c_int, _p, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, ctypes.c_int, _p, _sz, _p, _p, _p],
    "fsk_ref64_logdensity16": [_p, _i64, _p, _i64, _d, _p, _sz, _p, _p],
    "fsk_ref64_score16": [_p, _i64, _d, _p, _p, _p],
    "fsk_nd_workspace_bytes": [ctypes.c_int, _i64, _i64, ctypes.c_int],
    "fsk_kde_nd_splitk": [_p, _i64, _i64, _p, _i64, _i64, ctypes.c_int, _d, _u, _p, _sz, _p, _p],
    "fsk_score_nd_ordered": [_p, _i64, _i64, ctypes.c_int, _d, _u, _p, _sz, _p, _p, _i64, _i64,
                             _p],
    "fsk_sdkde_debias_nd_f64": [_p, _i64, _i64, ctypes.c_int, _p, _p, _i64, _i64, _d, _d, _p,
                                _i64, _p],
}
_RESTYPES = {
    "fsk_last_error": ctypes.c_char_p,
    "fsk_version": ctypes.c_int,
    "fsk_workspace_bytes": _sz,
    "fsk_profile_last_ms": _d,
    "fsk_ref64_workspace_bytes": _sz,
    "fsk_loo_workspace_bytes": _sz,
    "fsk_nd_workspace_bytes": _sz,
}

class AdamTensor(ctypes.Structure):
    """fsk_adam_tensor of include/flashsdkde.h."""
    _fields_ = [("w", ctypes.c_void_p), ("m", ctypes.c_void_p), ("v", ctypes.c_void_p),
                ("grad_part", ctypes.c_void_p), ("w_bf16", ctypes.c_void_p),
                ("wt_bf16", ctypes.c_void_p), ("count", ctypes.c_int64), ("rows", ctypes.c_int),
                ("cols", ctypes.c_int), ("splits", ctypes.c_int), ("reserved", ctypes.c_int)]


ADAM_MAX_TENSORS = 16


class WgradLayer(ctypes.Structure):
    """fsk_mlp_wgrad_layer of include/flashsdkde.h."""
    _fields_ = [("dz", ctypes.c_void_p), ("x", ctypes.c_void_p), ("ldx", ctypes.c_int64),
                ("N", ctypes.c_int), ("K", ctypes.c_int), ("splits", ctypes.c_int),
                ("x_bf16", ctypes.c_int), ("dw_part", ctypes.c_void_p),
                ("db_part", ctypes.c_void_p)]


class RowsStepArgs(ctypes.Structure):
    """fsk_mlp_rows_step_args of include/flashsdkde.h."""
    _fields_ = [("x", ctypes.c_void_p), ("t", ctypes.c_void_p), ("n", ctypes.c_int64),
                ("idx", ctypes.c_void_p), ("batch", ctypes.c_int64), ("steps", ctypes.c_int64),
                ("table", ctypes.c_void_p), ("counter", ctypes.c_void_p),
                ("sched", ctypes.c_void_p), ("in_dim", ctypes.c_int), ("hidden", ctypes.c_int),
                ("out_dim", ctypes.c_int), ("reserved", ctypes.c_int),
                ("w_bf16", ctypes.c_void_p * 4), ("wt_bf16", ctypes.c_void_p * 4),
                ("bias", ctypes.c_void_p * 4), ("xb", ctypes.c_void_p),
                ("h", ctypes.c_void_p * 3), ("dz", ctypes.c_void_p * 4),
                ("mse_scale", ctypes.c_double), ("loss", ctypes.c_void_p)]

_lib: Optional[ctypes.CDLL] = None
_lock = threading.Lock()


class NativeLibraryError(RuntimeError):
    """The HIP library could not be loaded."""


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load (once) and type the shared library. Raises if it is absent."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise NativeLibraryError(
                f"libflashsdkde.so not found at {path}; build it with "
                "`make -C flash-sd-kde_amd/csrc` (hipcc, gfx950)."
            )
        lib = ctypes.CDLL(path)
        for name, argtypes in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = argtypes
            fn.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = lib
        return lib


def lib() -> ctypes.CDLL:
    return load_library()


def last_error() -> str:
    msg = lib().fsk_last_error()
    return msg.decode() if msg else ""


def check(status: int, what: str) -> None:
    if status == FSK_OK:
        return
    msg = last_error()
    if status == FSK_EINVAL:
        raise ValueError(f"{what}: {msg}")
    raise RuntimeError(f"{what} failed with status {status}: {msg}")


def require_gpu(device: torch.device, what: str) -> None:
    if device.type != "cuda":
        raise ValueError(f"{what} requires a CUDA (ROCm/HIP) device, got {device}.")
    if not torch.cuda.is_available():
        raise RuntimeError(f"CUDA is not available but was requested for {what}.")
    load_library()


def stream_handle(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def workspace(op: int, n: int, m: int, flags: int, device: torch.device) -> tuple[torch.Tensor, int]:
    nbytes = int(lib().fsk_workspace_bytes(op, n, m, flags))
    buf = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=device)
    return buf, nbytes


def ptr(t: Optional[torch.Tensor]) -> Optional[int]:
    return None if t is None else t.data_ptr()
