"""C ABI: the library loads without a GPU, exports every declared symbol, and
rejects bad arguments before touching the device."""

import ctypes
import os
import re

import pytest

from flash_sd_kde import _native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    inc = os.path.join(ROOT, "include")
    for name in os.listdir(inc):
        if name.endswith(".h"):
            text = open(os.path.join(inc, name)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            syms |= set(re.findall(r"\b(fsk_[a-z0-9_]+)\s*\(", text))
    return syms


def test_library_loads_and_exports_every_declared_symbol():
    lib = nat.load_library()
    syms = declared_symbols()
    assert len(syms) >= 14
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # every declared symbol is typed in the Python binding
    assert syms <= set(nat.SIGNATURES), syms - set(nat.SIGNATURES)


def test_version_and_workspace_sizes():
    lib = nat.load_library()
    assert lib.fsk_version() >= 10000
    n, m = 1 << 20, 65536
    ws = lib.fsk_workspace_bytes(nat.OP_KDE16_SPLITK, n, m, 0)
    assert ws >= (n + m) * 4  # norms
    assert lib.fsk_workspace_bytes(nat.OP_KDE16_ATOMIC, n, m, 0) < ws
    # ordered score partials are capped (the reference's grow as n^2)
    assert lib.fsk_workspace_bytes(nat.OP_SCORE16_ORDERED, n, n, 0) <= (2 << 30) + 64 * n
    assert lib.fsk_workspace_bytes(nat.OP_SCORE16_SYM, n, n, 0) >= n * 17 * 8
    assert lib.fsk_workspace_bytes(999, n, m, 0) == 0
    assert lib.fsk_workspace_bytes(nat.OP_KDE16_SPLITK, 0, m, 0) == 0


def _fake(addr=0x1000):
    return ctypes.c_void_p(addr)


@pytest.mark.parametrize("h", [0.0, -1.0, float("nan"), float("inf")])
def test_bad_bandwidth_is_rejected(h):
    lib = nat.load_library()
    st = lib.fsk_kde16_splitk(_fake(), 16, 10, _fake(), 16, 10, h, 0, _fake(), 1 << 20,
                              _fake(), None)
    assert st == nat.FSK_EINVAL
    assert "bandwidth" in nat.last_error()


def test_bad_shapes_and_strides_are_rejected():
    lib = nat.load_library()
    # ld < 16
    assert lib.fsk_kde16_splitk(_fake(), 8, 10, _fake(), 16, 10, 1.0, 0, _fake(), 1 << 20,
                                _fake(), None) == nat.FSK_EINVAL
    # empty data
    assert lib.fsk_kde16_atomic(_fake(), 16, 0, _fake(), 16, 10, 1.0, 0, _fake(), 1 << 20,
                                _fake(), None) == nat.FSK_EINVAL
    # misaligned pointer
    assert lib.fsk_score16_ordered(_fake(0x1004), 16, 10, 1.0, 0, _fake(), 1 << 20, _fake(),
                                   _fake(), 16, 1, None) == nat.FSK_EINVAL
    # null output
    assert lib.fsk_score16_sym_atomic(_fake(), 16, 10, 1.0, 0, _fake(), 1 << 20, None, _fake(),
                                      16, 1, None) == nat.FSK_EINVAL
    assert lib.fsk_kde1_splitk(None, 10, _fake(), 10, 1.0, 0, _fake(), 1 << 20, _fake(),
                               None) == nat.FSK_EINVAL


def test_unsupported_flags_and_small_workspace():
    lib = nat.load_library()
    assert lib.fsk_kde16_splitk(_fake(), 16, 10, _fake(), 16, 10, 1.0, nat.PREC_BF16, _fake(),
                                1 << 20, _fake(), None) == nat.FSK_EUNSUPPORTED
    assert lib.fsk_kde16_logdensity(_fake(), 16, 10, _fake(), 16, 10, 1.0, nat.FLAG_LAPLACIAN,
                                    _fake(), 1 << 20, _fake(), None) == nat.FSK_EUNSUPPORTED
    assert lib.fsk_kde1_splitk(_fake(), 10, _fake(), 10, 1.0, nat.PREC_BF16, _fake(), 1 << 20,
                               _fake(), None) == nat.FSK_EUNSUPPORTED
    # precision values past the last defined form are rejected
    assert lib.fsk_kde16_logdensity(_fake(), 16, 10, _fake(), 16, 10, 1.0, nat.PREC_BF16X3S + 1,
                                    _fake(), 1 << 20, _fake(), None) == nat.FSK_EUNSUPPORTED
    assert lib.fsk_score16_ordered(_fake(), 16, 10, 1.0, nat.PREC_BF16X3S + 1, _fake(), 1 << 20,
                                   _fake(), _fake(), 16, 1, None) == nat.FSK_EUNSUPPORTED
    assert lib.fsk_kde16_splitk(_fake(), 16, 1000, _fake(), 16, 1000, 1.0, 0, _fake(), 16,
                                _fake(), None) == nat.FSK_EWORKSPACE
    assert lib.fsk_score16_ordered(_fake(), 16, 1000, 1.0, 0, None, 0, _fake(), _fake(), 16, 1,
                                   None) == nat.FSK_EWORKSPACE
    assert "workspace" in nat.last_error()


def test_check_maps_status_to_python_errors():
    lib = nat.load_library()
    lib.fsk_kde16_splitk(_fake(), 16, 10, _fake(), 16, 10, -1.0, 0, _fake(), 1 << 20, _fake(),
                         None)
    with pytest.raises(ValueError, match="bandwidth"):
        nat.check(nat.FSK_EINVAL, "fsk_kde16_splitk")
    with pytest.raises(RuntimeError):
        nat.check(nat.FSK_EWORKSPACE, "x")


def test_row_counts_above_the_elementwise_limit_are_rejected():
    """Score, debias and weighted-combine paths run one thread per (row,
    feature); above 2^27 rows they refuse before any launch."""
    lib = nat.load_library()
    big = (1 << 27) + 1
    assert lib.fsk_score16_sym_atomic(_fake(), 16, big, 1.0, 0, _fake(), 1 << 20, _fake(),
                                      _fake(), 16, 1, None) == nat.FSK_EUNSUPPORTED
    assert "rows" in nat.last_error()
    assert lib.fsk_score16_ordered(_fake(), 16, big, 1.0, 0, _fake(), 1 << 20, _fake(),
                                   _fake(), 16, 1, None) == nat.FSK_EUNSUPPORTED
    assert lib.fsk_sdkde_debias_f64(_fake(), 16, big, _fake(), _fake(), 16, 1, 1.0, 1e-12,
                                    _fake(), 16, None, None) == nat.FSK_EUNSUPPORTED
    assert lib.fsk_reduce_weighted_partials_f64(_fake(), 2, big, 16 * big, _fake(), 16, 1,
                                                None) == nat.FSK_EUNSUPPORTED
