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


def _adam_oracle_run(w0, grads, lrs, *, fixed_step=None):
    w, m, v = w0.clone(), torch.zeros_like(w0), torch.zeros_like(w0)
    out = []
    for k, (g, lr) in enumerate(zip(grads, lrs), start=1):
        w, m, v = mlp.adam(w, m, v, g, fixed_step or k, lr)
        out.append(w.clone())
    return out


def test_adam_matches_torch_optim_over_steps_with_a_schedule():
    """oracle.mlp.adam against torch.optim.Adam (bias-corrected first and
    second moments) over 6 steps whose learning rate changes every step, as
    the trainer's cosine schedule does; steps >= 2 are where the bias
    corrections 1 - beta^t differ from their first-step values."""
    g = torch.Generator().manual_seed(11)
    w0 = torch.randn(5, 7, generator=g, dtype=torch.float64)
    grads = [torch.randn(5, 7, generator=g, dtype=torch.float64) * (0.1 + k) for k in range(6)]
    lrs = [1e-2 * (0.5 + 0.5 * torch.cos(torch.tensor(k / 6 * 3.14159)).item()) for k in range(6)]
    p = torch.nn.Parameter(w0.clone())
    opt = torch.optim.Adam([p], lr=lrs[0], betas=(0.9, 0.999), eps=1e-8)
    ours = _adam_oracle_run(w0, grads, lrs)
    for k in range(6):
        for group in opt.param_groups:
            group["lr"] = lrs[k]
        p.grad = grads[k].clone()
        opt.step()
        assert torch.allclose(ours[k], p.detach(), rtol=1e-13, atol=1e-15), k
    # a planted mistake in the oracle (bias correction frozen at step 1)
    # separates from torch from the second step on
    wrong = _adam_oracle_run(w0, grads, lrs, fixed_step=1)
    assert torch.allclose(wrong[0], ours[0], rtol=1e-13, atol=1e-15)
    assert not torch.allclose(wrong[2], ours[2], rtol=1e-6, atol=1e-9)


def test_adam_first_step_closed_form():
    """At step 1 the corrected moments are g and g^2, so the update is
    lr * g / (|g| + eps): every weight moves by ~lr against its gradient."""
    w = torch.tensor([1.0, -2.0, 0.5], dtype=torch.float64)
    gr = torch.tensor([3.0, -1e-3, 0.0], dtype=torch.float64)
    w1, m1, v1 = mlp.adam(w, torch.zeros(3, dtype=torch.float64), torch.zeros(3, dtype=torch.float64),
                          gr, 1, 0.1)
    expect = w - 0.1 * gr / (gr.abs() + 1e-8)
    assert torch.allclose(w1, expect, rtol=0, atol=1e-15)
    assert torch.allclose(m1, 0.1 * gr) and torch.allclose(v1, 0.001 * gr * gr)
