"""fsk_kde16_splitk_noalloc (the allocation-free variant of the split-K KDE,
scratch in a static device array) against fsk_kde16_splitk with a caller
workspace, through the C ABI: two calls in a row, the second on smaller
inputs, must give the same sums bit for bit (same kernels, same plan), and
both agree with the FP64 oracle; sizes beyond the static buffer are refused
before any GPU work."""

import ctypes

import numpy as np
import pytest
import torch

from flash_sd_kde import _native as nat
from oracle import fp64

pytestmark = pytest.mark.gpu


def _sums(lib, fn, x, q, h, flags, dev):
    n, m = x.shape[0], q.shape[0]
    out = torch.full((m,), float("nan"), device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    if fn == "noalloc":
        st = lib.fsk_kde16_splitk_noalloc(x.data_ptr(), 16, n, q.data_ptr(), 16, m, h, flags,
                                          out.data_ptr(), stream)
    else:
        nbytes = int(lib.fsk_workspace_bytes(nat.OP_KDE16_SPLITK, n, m, flags))
        ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=dev)
        st = lib.fsk_kde16_splitk(x.data_ptr(), 16, n, q.data_ptr(), 16, m, h, flags,
                                  ws.data_ptr(), nbytes, out.data_ptr(), stream)
    nat.check(st, fn)
    torch.cuda.synchronize(dev)
    return out


def test_noalloc_matches_workspace_entry_point_on_two_calls(gpu):
    lib = nat.load_library()
    # second call smaller than the first: the scratch keeps the first call's
    # contents, and nothing of it may leak into the second result
    for n, m, h, flags, seed in [(6000, 900, 1.1, 0, 1), (1500, 200, 0.9, nat.FLAG_LAPLACIAN, 2)]:
        rng = np.random.default_rng(seed)
        x = rng.normal(size=(n, 16)).astype(np.float32)
        q = rng.normal(size=(m, 16)).astype(np.float32)
        xd, qd = torch.as_tensor(x, device=gpu), torch.as_tensor(q, device=gpu)
        print(f"noalloc case n={n} m={m} flags={flags}", flush=True)
        a = _sums(lib, "noalloc", xd, qd, h, flags, gpu)
        b = _sums(lib, "workspace", xd, qd, h, flags, gpu)
        assert torch.equal(a, b)
        exact = fp64.kde_sums(q, x, h, laplacian=bool(flags))
        got = a.cpu().numpy().astype(np.float64)
        if flags:  # the Laplacian factor changes sign: compare on the result's scale
            assert np.max(np.abs(got - exact)) <= 1e-5 * np.max(np.abs(exact))
        else:
            assert np.max(np.abs(got - exact) / exact) <= 1e-5


def test_noalloc_refuses_sizes_beyond_the_static_buffer(gpu):
    lib = nat.load_library()
    x = torch.zeros((16, 16), device=gpu)
    out = torch.zeros(16, device=gpu)
    n, m = 1 << 20, 65536  # needs ~90 MB of scratch; refused before any launch
    st = lib.fsk_kde16_splitk_noalloc(x.data_ptr(), 16, n, x.data_ptr(), 16, m, 1.0, 0,
                                      out.data_ptr(), None)
    assert st == nat.FSK_EUNSUPPORTED
    assert "static buffer" in nat.last_error()
    assert bool((out == 0).all())
