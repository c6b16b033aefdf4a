"""GPU: the library's issue-rate and copy microbenchmarks (the measured side
of the roofline) run, return positive rates, and the copy moves the bytes it
claims to."""

import ctypes

import pytest
import torch

from flash_sd_kde import _native

pytestmark = pytest.mark.gpu


def test_microbench_rates_positive(gpu):
    lib = _native.lib()
    out = torch.empty(64 * 256, device=gpu)
    r = ctypes.c_double()
    for which in (0, 1, 2, 3, 4, 6, 7, 8, 9):
        assert lib.fsk_microbench(which, 64, 10, out.data_ptr(), ctypes.addressof(r), None) == 0
        assert r.value > 0


def test_microbench_copy_moves_first_half(gpu):
    lib = _native.lib()
    blocks = 1000  # 256,000 floats; halves of 128,000
    buf = torch.zeros(blocks * 256, device=gpu)
    half = blocks * 128
    buf[:half] = torch.arange(half, device=gpu, dtype=torch.float32)
    r = ctypes.c_double()
    assert lib.fsk_microbench(5, blocks, 2, buf.data_ptr(), ctypes.addressof(r), None) == 0
    assert r.value > 0
    assert torch.equal(buf[half:], buf[:half])
    # modes past the last one are refused before anything is launched
    assert lib.fsk_microbench(10, blocks, 2, buf.data_ptr(), ctypes.addressof(r), None) != 0
