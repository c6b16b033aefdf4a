"""GPU: the pair kernels are capturable in a HIP graph (no allocation, no
host synchronisation, fills by kernel instead of memset) and every replay,
with unrelated eager work and synchronisations between replays, reproduces
the eager result bit for bit."""

import numpy as np
import pytest
import torch

from flash_sd_kde import _ops

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp32", "fp16x3"])
def test_sdkde_step_replays_bitwise(gpu, precision):
    rng = np.random.default_rng(7)
    x = torch.as_tensor(rng.normal(size=(3000, 16)).astype(np.float32), device=gpu)
    q = torch.as_tensor(rng.normal(size=(700, 16)).astype(np.float32), device=gpu)
    h = 1.1

    def step():
        pdf, w = _ops.score16(x, h, variant="ordered", precision=precision)
        deb = _ops.debias(x, pdf, w, h, 1e-12)
        if precision == "fp32":
            return _ops.kde16_sums(deb, q, h)
        return _ops.kde16_logdensity(deb, q, h, precision=precision)

    eager = step()
    torch.cuda.synchronize(gpu)
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream(device=gpu)
    side.wait_stream(torch.cuda.current_stream(gpu))
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            out = step()
    torch.cuda.current_stream(gpu).wait_stream(side)
    torch.cuda.synchronize(gpu)
    for k in range(3):
        torch.randn(1000, device=gpu).sum().item()  # unrelated eager work between replays
        g.replay()
        torch.cuda.synchronize(gpu)
        assert torch.equal(out, eager), f"replay {k} differs from the eager result"


def test_symmetric_sdkde_step_replays(gpu):
    """The default FP32 step (symmetric score: FP64 atomic scatter, so replays
    agree to rounding rather than bitwise) is capturable: its accumulators are
    cleared by fill kernels inside the graph, and every replay, with unrelated
    eager work between, matches the eager result."""
    rng = np.random.default_rng(11)
    x = torch.as_tensor(rng.normal(size=(2600, 16)).astype(np.float32), device=gpu)
    q = torch.as_tensor(rng.normal(size=(700, 16)).astype(np.float32), device=gpu)
    h = 1.1

    def step():
        pdf, w = _ops.score16(x, h, variant="symmetric")
        deb = _ops.debias(x, pdf, w, h, 1e-12)
        return _ops.kde16_sums(deb, q, h)

    eager = step()
    torch.cuda.synchronize(gpu)
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream(device=gpu)
    side.wait_stream(torch.cuda.current_stream(gpu))
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            out = step()
    torch.cuda.current_stream(gpu).wait_stream(side)
    torch.cuda.synchronize(gpu)
    for k in range(3):
        torch.randn(1000, device=gpu).sum().item()  # unrelated eager work between replays
        g.replay()
        torch.cuda.synchronize(gpu)
        torch.testing.assert_close(out, eager, rtol=1e-6, atol=0.0,
                                   msg=lambda m: f"replay {k}: {m}")


def test_one_d_path_refuses_graph_capture(gpu):
    """The 1-D kernels sort with library kernels that record memset nodes,
    which do not replay reliably from a HIP graph: under capture the entry
    point refuses with a clear error and launches nothing; outside capture
    the same call works."""
    x = torch.linspace(-3.0, 3.0, 5000, device=gpu)
    q = torch.linspace(-2.0, 2.0, 300, device=gpu)
    eager = _ops.kde1_sums(x, q, 0.3)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream(device=gpu)
    side.wait_stream(torch.cuda.current_stream(gpu))
    with pytest.raises(RuntimeError, match="capture"):
        with torch.cuda.stream(side):
            with torch.cuda.graph(g, stream=side):
                _ops.kde1_sums(x, q, 0.3)
    torch.cuda.current_stream(gpu).wait_stream(side)
    torch.cuda.synchronize()
    again = _ops.kde1_sums(x, q, 0.3)
    assert torch.equal(again, eager)
