"""GPU: MLP kernels against the CPU oracle (forward, gradients, Adam),
training, checkpoint round trip, and the surrogate SD-KDE pipeline including
what it gains over plain KDE (no debiasing)."""

import json
import os

import numpy as np
import pytest
import torch

from flash_sd_kde import _ops
from flash_sd_kde.bandwidth import scott_mean_std_bandwidth
from flash_sd_kde.datasets import gmm_full_cov, gmm_full_cov_logpdf, gmm_full_cov_params
from flash_sd_kde.surrogate import ScoreMLP, exact_shift, surrogate_fit_transform
from oracle import mlp

pytestmark = pytest.mark.gpu


def record(name, **vals):
    path = os.environ.get("FSK_METRICS_OUT")
    print(name, vals, flush=True)
    if path:
        with open(path, "a") as f:
            f.write(json.dumps({"test": name, **vals}) + "\n")


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _oracle_weights(net):
    Ws = [L["W"].cpu().double() for L in net.layers]
    bs = [L["b"].cpu().double() for L in net.layers]
    return Ws, bs


@pytest.mark.parametrize("M,splits", [(1, 1), (37, 3), (1000, 1), (1000, 4), (4096, 16)])
def test_mlp_forward_and_backward_match_oracle(gpu, M, splits):
    """Split counts include empty trailing splits (37 rows over 3 splits of 64)."""
    net = ScoreMLP(seed=3, device=gpu)
    g = torch.Generator().manual_seed(M)
    x = torch.randn(M, 16, generator=g)
    t = torch.randn(M, 16, generator=g)
    cache = []
    y = net.forward(x.to(gpu), cache)
    Ws, bs = _oracle_weights(net)
    y_ref, cache_ref = mlp.forward(Ws, bs, x)
    e_fwd = rel(y.cpu(), y_ref)
    dy = net.mse(y, t.to(gpu), None)
    dy_ref, _ = mlp.mse_grad(y_ref, t)
    grads = net.backward(dy, cache, splits)
    grads_ref = mlp.backward(Ws, cache_ref, dy_ref)
    e_w = max(rel(gw.sum(0).cpu(), gr[0]) for (gw, _), gr in zip(grads, grads_ref))
    e_b = max(rel(gb.sum(0).cpu(), gr[1]) for (_, gb), gr in zip(grads, grads_ref))
    # GPU (FP32) and oracle (FP64) round the same activations to BF16; an
    # activation within rounding noise of a BF16 tie can land on the other
    # neighbour, so single elements may differ by ~2^-8 of one input term:
    # bound the maximum loosely and the typical element tightly.
    med_fwd = float(np.median(np.abs(y.cpu().numpy() - y_ref.numpy())) / np.abs(y_ref.numpy()).max())
    record(f"mlp_M{M}_s{splits}", fwd_rel=e_fwd, fwd_median_rel=med_fwd, dW_rel=e_w, db_rel=e_b)
    assert e_fwd < 3e-3 and med_fwd < 1e-5 and e_w < 3e-3 and e_b < 3e-3


def test_mlp_adam_matches_oracle(gpu):
    """Three Adam steps of every tensor against the oracle (FP64, pinned to
    torch.optim.Adam on the CPU), once with the step count passed from the
    host and once with the device schedule [lr, 1 - b1^t, 1 - b2^t] the
    captured training step reads; the learning rate changes every step.
    Steps 2 and 3 exercise the bias corrections beyond their step-1 values."""
    g = torch.Generator(device=gpu).manual_seed(5)
    steps = [[(torch.randn(2, L["N"], L["K"], device=gpu, generator=g),
               torch.randn(2, L["N"], device=gpu, generator=g))
              for L in ScoreMLP(seed=4, device=gpu).layers] for _ in range(3)]
    lrs = [1e-2, 7e-3, 4e-3]
    for mode in ("host_step", "device_schedule"):
        net = ScoreMLP(seed=4, device=gpu)
        ref = [[L["W"].cpu().double(), L["b"].cpu().double()] for L in net.layers]
        mom = [[torch.zeros_like(t) for t in pair] for pair in ref]
        vel = [[torch.zeros_like(t) for t in pair] for pair in ref]
        sched = torch.empty(3, dtype=torch.float32, device=gpu)
        for k, (grads, lr) in enumerate(zip(steps, lrs), start=1):
            if mode == "host_step":
                net.adam_step(grads, lr)
            else:
                sched.copy_(torch.tensor([lr, 1 - 0.9 ** k, 1 - 0.999 ** k]))
                net.adam_step(grads, 123.0, sched=sched)  # host lr ignored
            for l, (dW, db) in enumerate(grads):
                for j, part in enumerate((dW, db)):
                    ref[l][j], mom[l][j], vel[l][j] = mlp.adam(
                        ref[l][j], mom[l][j], vel[l][j], part.sum(0).cpu().double(), k, lr)
            errs = [rel(L["W"].cpu(), r[0]) for L, r in zip(net.layers, ref)]
            errs += [rel(L["b"].cpu(), r[1]) for L, r in zip(net.layers, ref)]
            record(f"mlp_adam_{mode}_step{k}", max_rel=max(errs))
            assert max(errs) < 5e-7, (mode, k, errs)
        # BF16 copies follow the master weights (and the transposed copy its transpose)
        for L in net.layers:
            Wb = L["Wb"].cpu().view(torch.bfloat16).double()
            Wt = L["Wt"].cpu().view(torch.bfloat16).double()
            assert torch.equal(Wb, mlp.bf16(L["W"].cpu().double()))
            assert torch.equal(Wt, Wb.T)


def test_adam_multi_equals_per_tensor_updates(gpu):
    """The one-launch optimizer step updates every tensor exactly as the
    per-tensor entry point does (same arithmetic, bitwise)."""
    from flash_sd_kde import _native as nat

    a = ScoreMLP(seed=6, device=gpu)
    b = ScoreMLP(seed=6, device=gpu)
    g = torch.Generator(device=gpu).manual_seed(3)
    grads = [(torch.randn(3, L["N"], L["K"], device=gpu, generator=g),
              torch.randn(3, L["N"], device=gpu, generator=g)) for L in a.layers]
    for step in (1, 2):
        a.adam_step(grads, 1e-2)
        for L, (dW, db) in zip(b.layers, grads):
            for t, mm, vv, part, wb, wt, rows, cols in (
                    (L["W"], L["mW"], L["vW"], dW, L["Wb"], L["Wt"], L["N"], L["K"]),
                    (L["b"], L["mb"], L["vb"], db, None, None, 1, L["N"])):
                st = nat.lib().fsk_mlp_adam(t.data_ptr(), mm.data_ptr(), vv.data_ptr(),
                                            part.data_ptr(), part.shape[0], t.numel(), rows, cols,
                                            1e-2, 0.9, 0.999, 1e-8, step, nat.ptr(wb), nat.ptr(wt),
                                            None, None)
                nat.check(st, "fsk_mlp_adam")
        b.step_count += 1
    for La, Lb in zip(a.layers, b.layers):
        for key in ("W", "b", "mW", "vW", "Wb", "Wt"):
            assert torch.equal(La[key], Lb[key]), key


def test_mlp_trains_and_checkpoints(gpu, tmp_path):
    g = torch.Generator(device=gpu).manual_seed(0)
    x = torch.randn(20000, 16, device=gpu, generator=g)
    t = torch.stack([torch.sin(x[:, k]) * x[:, (k + 1) % 16] for k in range(16)], dim=1)
    net = ScoreMLP(seed=5, device=gpu)
    hist = net.fit(x, t, steps=600, batch=4096, lr=3e-3, log_every=100)
    record("mlp_train", first=hist[0], last=hist[-1])
    assert hist[-1] < 0.1 * hist[0]
    path = str(tmp_path / "net.npz")
    net.save(path)
    net2 = ScoreMLP.load(path, device=gpu)
    assert torch.equal(net.predict(x[:100]), net2.predict(x[:100]))


def test_mlp_graph_training_equals_eager(gpu):
    """Replays of the captured training step give bitwise the same weights as
    eager steps (same batches, same schedule), checked after several replays."""
    g = torch.Generator(device=gpu).manual_seed(1)
    x = torch.randn(5000, 16, device=gpu, generator=g)
    t = torch.tanh(x) * 2.0
    a = ScoreMLP(seed=9, device=gpu)
    b = ScoreMLP(seed=9, device=gpu)
    for steps in (4, 9):  # two fits in a row: the second continues from the first
        a.fit(x, t, steps=steps, batch=1024, lr=1e-3, seed=steps, graph=True, log_every=3)
        b.fit(x, t, steps=steps, batch=1024, lr=1e-3, seed=steps, graph=False, log_every=3)
        torch.ones(10, device=gpu).sum().item()  # unrelated eager work in between
        for La, Lb in zip(a.layers, b.layers):
            assert torch.equal(La["W"], Lb["W"]) and torch.equal(La["b"], Lb["b"])
            assert torch.equal(La["Wb"], Lb["Wb"]) and torch.equal(La["Wt"], Lb["Wt"])


def test_surrogate_sdkde_gain_over_plain_kde(gpu):
    params = gmm_full_cov_params(8, mean_scale=4.0, seed=7)
    x = gmm_full_cov(32768, params=params, seed=7)
    q = gmm_full_cov(2048, params=params, seed=8)
    h = scott_mean_std_bandwidth(x)
    xd = torch.as_tensor(x, device=gpu)
    qd = torch.as_tensor(q, device=gpu)
    deb_s, info = surrogate_fit_transform(xd, h, n_label=16384, n_holdout=4096, n_probe=1024,
                                          steps=1500, batch=4096, device=gpu, tol=1e-3)
    deb_e = xd + 0.5 * exact_shift(xd, xd, h)
    lp_e = _ops.kde16_logdensity(deb_e, qd, h).cpu().numpy()
    lp_s = _ops.kde16_logdensity(deb_s, qd, h).cpu().numpy()
    lp_k = _ops.kde16_logdensity(xd, qd, h).cpu().numpy()
    truth = gmm_full_cov_logpdf(q, params)
    err_s = float(np.max(np.abs(lp_s - lp_e) / np.abs(lp_e)))
    err_k = float(np.max(np.abs(lp_k - lp_e) / np.abs(lp_e)))
    mae = lambda a: float(np.mean(np.abs(a - truth)))
    record("surrogate_small", fallback=info["fallback"],
           holdout_err=info["holdout_max_rel_logp_err"], surrogate_vs_exact=err_s,
           plain_kde_vs_exact=err_k, mae_truth_exact=mae(lp_e), mae_truth_surrogate=mae(lp_s),
           mae_truth_plain=mae(lp_k), loss_last=info["loss"][-1])
    # the learned shift recovers most of what SD-KDE changes, and the KDE
    # built from it is closer to the true density than plain KDE
    assert err_s < 0.25 * err_k
    assert mae(lp_s) < mae(lp_k)


@pytest.mark.parametrize("M", [1, 300, 8192])
def test_fused_output_loss_equals_separate_launches(gpu, M):
    """fsk_mlp_linear_fwd_mse (output GEMM with the MSE gradient in its
    epilogue) against the output GEMM followed by fsk_mlp_mse: the same dy bit
    for bit and the same loss up to FP64 summation order."""
    from flash_sd_kde import _native as nat

    net = ScoreMLP(seed=8, device=gpu)
    L = net.layers[-1]
    g = torch.Generator(device=gpu).manual_seed(M)
    h = torch.randn(M, L["K"], device=gpu, generator=g)
    t = torch.randn(M, L["N"], device=gpu, generator=g)
    y = torch.empty(M, L["N"], device=gpu)
    st = nat.lib().fsk_mlp_linear_fwd(h.data_ptr(), L["K"], M, L["K"], L["Wb"].data_ptr(),
                                      L["b"].data_ptr(), L["N"], 0, None, y.data_ptr(), L["N"],
                                      None)
    nat.check(st, "fwd")
    loss_a = torch.zeros(1, dtype=torch.float64, device=gpu)
    dy_a = net.mse(y, t, loss_a)
    loss_b = torch.zeros(1, dtype=torch.float64, device=gpu)
    dy_b = torch.empty_like(dy_a)
    st = nat.lib().fsk_mlp_linear_fwd_mse(h.data_ptr(), L["K"], M, L["K"], L["Wb"].data_ptr(),
                                          L["b"].data_ptr(), L["N"], t.data_ptr(),
                                          2.0 / y.numel(), dy_b.data_ptr(), loss_b.data_ptr(), None)
    nat.check(st, "fwd_mse")
    assert torch.equal(dy_a, dy_b)
    assert abs(float(loss_a) - float(loss_b)) <= 1e-12 * float(loss_a)


@pytest.mark.parametrize("layer,M,splits", [(1, 8192, 8), (3, 300, 3), (2, 1, 1)])
def test_fused_backward_equals_separate_launches(gpu, layer, M, splits):
    """fsk_mlp_linear_bwd (weight gradient and data gradient x silu' in one
    launch) gives the partials and dZ of the two separate entry points bit for
    bit."""
    from flash_sd_kde import _native as nat

    net = ScoreMLP(seed=9, device=gpu)
    L = net.layers[layer]
    g = torch.Generator(device=gpu).manual_seed(layer * 100 + M)
    dz = torch.randn(M, L["N"], device=gpu, generator=g)
    h = torch.randn(M, L["K"], device=gpu, generator=g)
    zp = torch.randn(M, L["K"], device=gpu, generator=g)
    outs = []
    for fused in (True, False):
        dW = torch.full((splits, L["N"], L["K"]), float("nan"), device=gpu)
        db = torch.full((splits, L["N"]), float("nan"), device=gpu)
        dx = torch.full((M, L["K"]), float("nan"), device=gpu)
        if fused:
            st = nat.lib().fsk_mlp_linear_bwd(dz.data_ptr(), M, L["N"], h.data_ptr(), L["K"],
                                              L["K"], splits, dW.data_ptr(), db.data_ptr(),
                                              L["Wt"].data_ptr(), zp.data_ptr(), dx.data_ptr(),
                                              None)
            nat.check(st, "bwd")
        else:
            nat.check(nat.lib().fsk_mlp_linear_bwd_weight(dz.data_ptr(), M, L["N"], h.data_ptr(),
                                                          L["K"], L["K"], splits, dW.data_ptr(),
                                                          db.data_ptr(), None), "wgrad")
            nat.check(nat.lib().fsk_mlp_linear_fwd(dz.data_ptr(), L["N"], M, L["N"],
                                                   L["Wt"].data_ptr(), None, L["K"], 2,
                                                   zp.data_ptr(), dx.data_ptr(), L["K"], None),
                      "dgrad")
        outs.append((dW, db, dx))
    for a, b in zip(*outs):
        assert torch.equal(a, b)


@pytest.mark.parametrize("M", [8192, 300, 1])
def test_wgrad_multi_bf16_inputs_equal_fp32_inputs(gpu, M):
    """fsk_mlp_wgrad_multi with BF16 layer inputs (x_bf16 = 1, the step
    kernel's activations) gives the partials of the same inputs passed as FP32
    (rounded to BF16 inside) bit for bit, for every layer shape at once,
    including ragged batches; a misaligned BF16 row stride is refused."""
    import ctypes

    from flash_sd_kde import _native as nat

    net = ScoreMLP(seed=5, device=gpu)
    g = torch.Generator(device=gpu).manual_seed(M)
    dzs = [torch.randn(M, L["N"], device=gpu, generator=g) for L in net.layers]
    xs = [torch.randn(M, L["K"], device=gpu, generator=g) for L in net.layers]
    outs = []
    for bf in (0, 1):
        wl = (nat.WgradLayer * 4)()
        keep = []
        for l, L in enumerate(net.layers):
            S = 3 if M > 1 else 1
            dW = torch.full((S, L["N"], L["K"]), float("nan"), device=gpu)
            db = torch.full((S, L["N"]), float("nan"), device=gpu)
            xin = xs[l].to(torch.bfloat16) if bf else xs[l]
            keep.append((dW, db, xin))
            wl[l] = nat.WgradLayer(dzs[l].data_ptr(), xin.data_ptr(), L["K"], L["N"], L["K"], S, bf,
                                   dW.data_ptr(), db.data_ptr())
        nat.check(nat.lib().fsk_mlp_wgrad_multi(ctypes.byref(wl), 4, M, None), "wgrad_multi")
        torch.cuda.synchronize(gpu)
        outs.append([(dW, db) for dW, db, _ in keep])
    for (a, b), (c, d) in zip(*outs):
        assert torch.equal(a, c) and torch.equal(b, d)
    bad = (nat.WgradLayer * 1)()
    xin = xs[1].to(torch.bfloat16)
    bad[0] = nat.WgradLayer(dzs[1].data_ptr(), xin.data_ptr() + 2, 256, 256, 256, 1, 1,
                            keep[1][0].data_ptr(), keep[1][1].data_ptr())
    assert nat.lib().fsk_mlp_wgrad_multi(ctypes.byref(bad), 1, M, None) == nat.FSK_EINVAL


@pytest.mark.parametrize("batch,graph", [(300, False), (8192, True)])
def test_fused_training_step_equals_layered(gpu, batch, graph):
    """Training with the three-launch step (row-local kernel, multi-layer
    weight gradients, Adam with the counter) gives the weights of the
    layer-by-layer launches bit for bit, eager and graph-replayed; a ragged
    batch (300 rows: a partial last row block) included."""
    g = torch.Generator(device=gpu).manual_seed(batch)
    x = torch.randn(5000, 16, device=gpu, generator=g)
    t = torch.randn(5000, 16, device=gpu, generator=g)
    nets = []
    for fused in (True, False):
        net = ScoreMLP(seed=11, device=gpu)
        hist = net.fit(x, t, steps=6, batch=batch, lr=1e-2, graph=graph, fused=fused,
                       log_every=3)
        nets.append((net, hist))
    (a, ha), (b, hb) = nets
    for La, Lb in zip(a.layers, b.layers):
        for k in ("W", "b", "mW", "vW", "Wb", "Wt"):
            assert torch.equal(La[k], Lb[k]), k
    np.testing.assert_allclose(ha, hb, rtol=1e-12)  # loss sums differ in FP64 order only
