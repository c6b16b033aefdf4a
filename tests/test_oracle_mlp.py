"""Pins the MLP oracle: manual backward = autograd (FP64, no rounding), and
autograd = central finite differences of the MSE loss."""

import torch

from oracle import mlp


def _net(seed=0, dims=(16, 24, 24, 16)):
    g = torch.Generator().manual_seed(seed)
    Ws = [torch.randn(n, k, generator=g, dtype=torch.float64) / k ** 0.5
          for k, n in zip(dims[:-1], dims[1:])]
    bs = [torch.randn(n, generator=g, dtype=torch.float64) * 0.1 for n in dims[1:]]
    x = torch.randn(7, dims[0], generator=g, dtype=torch.float64)
    t = torch.randn(7, dims[-1], generator=g, dtype=torch.float64)
    return Ws, bs, x, t


def test_manual_backward_equals_autograd():
    Ws, bs, x, t = _net()
    y, cache = mlp.forward(Ws, bs, x, round_bf16=False)
    dy, _ = mlp.mse_grad(y, t)
    grads = mlp.backward(Ws, cache, dy, round_bf16=False)
    Wr = [W.clone().requires_grad_(True) for W in Ws]
    br = [b.clone().requires_grad_(True) for b in bs]
    y2, _ = mlp.forward(Wr, br, x, round_bf16=False)
    ((y2 - t) ** 2).mean().backward()
    for (dW, db), W, b in zip(grads, Wr, br):
        assert torch.allclose(dW, W.grad, rtol=1e-12, atol=1e-14)
        assert torch.allclose(db, b.grad, rtol=1e-12, atol=1e-14)


def test_autograd_equals_finite_differences():
    Ws, bs, x, t = _net(1)
    Wr = [W.clone().requires_grad_(True) for W in Ws]
    y, _ = mlp.forward(Wr, bs, x, round_bf16=False)
    ((y - t) ** 2).mean().backward()
    eps = 1e-6
    for l in range(len(Ws)):
        for (i, j) in ((0, 0), (3, 5), (Ws[l].shape[0] - 1, Ws[l].shape[1] - 1)):
            Wp = [W.clone() for W in Ws]
            Wm = [W.clone() for W in Ws]
            Wp[l][i, j] += eps
            Wm[l][i, j] -= eps
            lp = ((mlp.forward(Wp, bs, x, round_bf16=False)[0] - t) ** 2).mean()
            lm = ((mlp.forward(Wm, bs, x, round_bf16=False)[0] - t) ** 2).mean()
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - Wr[l].grad[i, j]) < 1e-7 * max(1.0, abs(fd))


def test_bf16_rounding_is_nearest_even():
    v = torch.tensor([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9], dtype=torch.float64)
    assert mlp.bf16(v).tolist() == [1.0, 1.0 + 2 ** -6, 1.0]
