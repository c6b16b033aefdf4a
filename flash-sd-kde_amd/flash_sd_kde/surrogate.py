"""Learned score surrogate for SD-KDE debiasing (no reference counterpart).

The exact one-step SD-KDE shifts every point by half its mean shift,
    x_i -> x_i + (h^2 / 2) score_i = x_i + delta_i / 2,   delta_i = w_i / pdf_i - x_i,
which costs n^2 kernel evaluations. Here a small MLP, 16 -> 256 -> 256 ->
256 -> 16 with SiLU (BF16 weights, FP32 accumulation), learns x -> delta from
exact labels on a subset (the score of the subset rows under the KDE of ALL
points, fsk_score16_rows) and then shifts all n points at O(n) cost.

Training runs entirely in this library's HIP kernels: forward GEMMs with
fused bias + SiLU, SiLU backward, split-batch weight gradients, data
gradients through the transposed BF16 weights, and fused Adam.

Safety net: after training, the surrogate's shifts and the exact shifts of a
held-out point set are each used to build a KDE, and the maximum relative
log-density difference at probe queries is measured. This is a sample
estimate, on a smaller KDE and fewer queries, of the error the surrogate
causes on the full problem, so the check keeps a margin: if it exceeds
`check_margin * tol` (0.5 x 1e-3 by default), surrogate_fit_transform falls
back to the exact score kernel for all points.
"""

from __future__ import annotations

import ctypes
import math
import time

import numpy as np
import torch

from flash_sd_kde import _native as nat
from flash_sd_kde import _ops
from flash_sd_kde.spatial import spatial_order

DIMS = (16, 256, 256, 256, 16)


def _bf16_buffer(shape, device):
    return torch.empty(shape, dtype=torch.int16, device=device)


class ScoreMLP:
    def __init__(self, dims=DIMS, seed: int = 0, device="cuda"):
        self.device = torch.device(device)
        nat.require_gpu(self.device, "ScoreMLP")
        self.dims = tuple(int(d) for d in dims)
        for d in self.dims:
            if d % 8:
                raise ValueError("layer widths must be multiples of 8")
        rng = np.random.default_rng(seed)
        self.layers = []
        for K, N in zip(self.dims[:-1], self.dims[1:]):
            bound = 1.0 / math.sqrt(K)
            W = torch.as_tensor(rng.uniform(-bound, bound, (N, K)).astype(np.float32),
                                device=self.device)
            b = torch.as_tensor(rng.uniform(-bound, bound, N).astype(np.float32),
                                device=self.device)
            L = {"W": W, "b": b, "mW": torch.zeros_like(W), "vW": torch.zeros_like(W),
                 "mb": torch.zeros_like(b), "vb": torch.zeros_like(b),
                 "Wb": _bf16_buffer((N, K), self.device), "Wt": _bf16_buffer((K, N), self.device),
                 "K": K, "N": N}
            self.layers.append(L)
        self._cast()
        self.step_count = 0
        # input / output normalisation (set by fit)
        self.x_mean = torch.zeros(16, device=self.device)
        self.x_scale = torch.ones(16, device=self.device)
        self.y_scale = 1.0

    def _stream(self):
        return nat.stream_handle(self.device)

    def _cast(self):
        for L in self.layers:
            st = nat.lib().fsk_mlp_cast_bf16(L["W"].data_ptr(), L["N"], L["K"], L["Wb"].data_ptr(),
                                             L["Wt"].data_ptr(), self._stream())
            nat.check(st, "fsk_mlp_cast_bf16")

    # ---- raw network (normalised units) ----
    def forward(self, x: torch.Tensor, cache: list | None = None) -> torch.Tensor:
        h = x.contiguous()
        M = h.shape[0]
        last = len(self.layers) - 1
        for l, L in enumerate(self.layers):
            act = 0 if l == last else 1
            y = torch.empty((M, L["N"]), dtype=torch.float32, device=self.device)
            z = torch.empty_like(y) if (cache is not None and act) else None
            st = nat.lib().fsk_mlp_linear_fwd(h.data_ptr(), h.stride(0), M, L["K"],
                                              L["Wb"].data_ptr(), L["b"].data_ptr(), L["N"], act,
                                              nat.ptr(z), y.data_ptr(), L["N"], self._stream())
            nat.check(st, "fsk_mlp_linear_fwd")
            if cache is not None:
                cache.append((h, z))
            h = y
        return h

    def forward_loss(self, x: torch.Tensor, t: torch.Tensor, loss_acc: torch.Tensor | None,
                     cache: list) -> torch.Tensor:
        """Training forward pass: the hidden layers as forward(), the output
        layer with the MSE gradient fused into its GEMM; returns dL/dy for
        L = mean (y - t)^2 and adds sum (y - t)^2 to loss_acc."""
        h = x.contiguous()
        M = h.shape[0]
        for L in self.layers[:-1]:
            y = torch.empty((M, L["N"]), dtype=torch.float32, device=self.device)
            z = torch.empty_like(y)
            st = nat.lib().fsk_mlp_linear_fwd(h.data_ptr(), h.stride(0), M, L["K"],
                                              L["Wb"].data_ptr(), L["b"].data_ptr(), L["N"], 1,
                                              z.data_ptr(), y.data_ptr(), L["N"], self._stream())
            nat.check(st, "fsk_mlp_linear_fwd")
            cache.append((h, z))
            h = y
        L = self.layers[-1]
        dy = torch.empty((M, L["N"]), dtype=torch.float32, device=self.device)
        st = nat.lib().fsk_mlp_linear_fwd_mse(h.data_ptr(), h.stride(0), M, L["K"],
                                              L["Wb"].data_ptr(), L["b"].data_ptr(), L["N"],
                                              t.data_ptr(), 2.0 / (M * L["N"]), dy.data_ptr(),
                                              nat.ptr(loss_acc), self._stream())
        nat.check(st, "fsk_mlp_linear_fwd_mse")
        cache.append((h, None))
        return dy

    @staticmethod
    def auto_splits(M: int, N: int, K: int, target_groups: int = 512,
                    max_splits: int = 32) -> int:
        """Batch splits of a weight-gradient launch: enough (64 x 64 tile,
        split) workgroups to fill the GPU (~512), at most one split per 64
        batch rows and at most `max_splits` (the optimizer sums the split
        partials of each weight one after another: the 16-wide layers, with
        4 tiles, would otherwise take 128)."""
        tiles = -(-N // 64) * -(-K // 64)
        return max(1, min(-(-target_groups // tiles), max(1, M // 64), max_splits))

    def backward(self, dy: torch.Tensor, cache: list, splits: int | None = None):
        """Gradient partials per layer: list of (dW_part, db_part). The data
        gradient GEMM of layer l also applies silu'(z_{l-1}), so it hands the
        previous layer its dZ directly. `splits` fixes the batch split count of
        every layer; None picks it per layer (auto_splits)."""
        M = dy.shape[0]
        grads = [None] * len(self.layers)
        dz = dy.contiguous()
        for l in range(len(self.layers) - 1, -1, -1):
            L = self.layers[l]
            h, _ = cache[l]
            S = splits or self.auto_splits(M, L["N"], L["K"])
            dWp = torch.empty((S, L["N"], L["K"]), dtype=torch.float32, device=self.device)
            dbp = torch.empty((S, L["N"]), dtype=torch.float32, device=self.device)
            grads[l] = (dWp, dbp)
            if l == 0:
                st = nat.lib().fsk_mlp_linear_bwd_weight(dz.data_ptr(), M, L["N"], h.data_ptr(),
                                                         h.stride(0), L["K"], S, dWp.data_ptr(),
                                                         dbp.data_ptr(), self._stream())
                nat.check(st, "fsk_mlp_linear_bwd_weight")
                break
            # weight gradient and the previous layer's dZ (data gradient times
            # silu'(z_prev)) in one launch
            z_prev = cache[l - 1][1]
            dz_prev = torch.empty((M, L["K"]), dtype=torch.float32, device=self.device)
            st = nat.lib().fsk_mlp_linear_bwd(dz.data_ptr(), M, L["N"], h.data_ptr(), h.stride(0),
                                              L["K"], S, dWp.data_ptr(), dbp.data_ptr(),
                                              L["Wt"].data_ptr(), z_prev.data_ptr(),
                                              dz_prev.data_ptr(), self._stream())
            nat.check(st, "fsk_mlp_linear_bwd")
            dz = dz_prev
        return grads

    def adam_step(self, grads, lr: float, beta1=0.9, beta2=0.999, eps=1e-8,
                  sched: torch.Tensor | None = None, counter: torch.Tensor | None = None):
        """One Adam update of every weight and bias in a single launch. With
        `sched` (device [lr, 1-b1^t, 1-b2^t]) the learning rate and bias
        corrections are read on the device, so the call can live inside a
        captured graph."""
        self.step_count += 1
        step = 0 if sched is not None else self.step_count
        arr = (nat.AdamTensor * (2 * len(self.layers)))()
        k = 0
        for L, (dWp, dbp) in zip(self.layers, grads):
            arr[k] = nat.AdamTensor(L["W"].data_ptr(), L["mW"].data_ptr(), L["vW"].data_ptr(),
                                    dWp.data_ptr(), L["Wb"].data_ptr(), L["Wt"].data_ptr(),
                                    L["W"].numel(), L["N"], L["K"], dWp.shape[0], 0)
            arr[k + 1] = nat.AdamTensor(L["b"].data_ptr(), L["mb"].data_ptr(), L["vb"].data_ptr(),
                                        dbp.data_ptr(), None, None, L["b"].numel(), 1, L["N"],
                                        dbp.shape[0], 0)
            k += 2
        for a in range(0, k, nat.ADAM_MAX_TENSORS):
            cnt = min(nat.ADAM_MAX_TENSORS, k - a)
            last = a + cnt >= k  # the step counter advances once, with the last group
            st = nat.lib().fsk_mlp_adam_multi_step(ctypes.addressof(arr[a]), cnt, lr, beta1,
                                                   beta2, eps, step, nat.ptr(sched),
                                                   nat.ptr(counter) if last else None,
                                                   self._stream())
            nat.check(st, "fsk_mlp_adam_multi_step")

    def mse(self, y: torch.Tensor, t: torch.Tensor, loss_acc: torch.Tensor | None):
        dy = torch.empty_like(y)
        st = nat.lib().fsk_mlp_mse(y.data_ptr(), t.data_ptr(), y.numel(), 2.0 / y.numel(),
                                   dy.data_ptr(), nat.ptr(loss_acc), self._stream())
        nat.check(st, "fsk_mlp_mse")
        return dy

    # ---- training / inference in data units ----
    def fit(self, x: torch.Tensor, target: torch.Tensor, *, steps: int = 3000, batch: int = 8192,
            lr: float = 2e-3, seed: int = 0, log_every: int = 250, graph: bool = True,
            beta1: float = 0.9, beta2: float = 0.999, splits: int | None = None,
            warmup: int = 0, fused: bool | None = None):
        """Fit x (n,16) -> target (n,16); inputs standardised per feature and
        targets by their overall scale; cosine learning-rate decay; minibatches
        drawn uniformly with replacement; `warmup` steps of linear learning-rate
        ramp before the decay. With graph=True one training step
        (gather, forward, loss, backward, Adam, step counter) is captured once
        as a HIP graph and replayed: the step's batch indices and schedule
        entry are selected on the device by a step counter, so nothing is
        launched from the host between replays.

        fused (default: whenever the network is 16-256-256-256-16) runs each
        step as three launches - the row-local step (gather, forward, loss,
        data gradients; fsk_mlp_rows_step), all weight gradients
        (fsk_mlp_wgrad_multi) and Adam with the step counter - instead of the
        layer-by-layer launches; the values are the same bit for bit."""
        x = x.float().contiguous()
        target = target.float().contiguous()
        self.x_mean = x.mean(0)
        self.x_scale = 1.0 / x.std(0).clamp_min(1e-12)
        self.y_scale = float(target.std().clamp_min(1e-12))
        xn = ((x - self.x_mean) * self.x_scale).contiguous()
        tn = (target / self.y_scale).contiguous()
        n = xn.shape[0]
        batch = min(batch, n)
        # splits: None = per-layer choice (auto_splits)
        dev = self.device
        gen = torch.Generator(device=dev)
        gen.manual_seed(seed)
        t0 = self.step_count
        k = torch.arange(1, steps + 1, dtype=torch.float64)
        ramp = torch.clamp(k / max(warmup, 1), max=1.0) if warmup > 0 else torch.ones_like(k)
        table = torch.stack([lr * ramp * 0.5 * (1.0 + torch.cos(math.pi * (k - 1) / steps)),
                             1.0 - beta1 ** (k + t0), 1.0 - beta2 ** (k + t0)], dim=1)
        table = table.to(torch.float32).to(dev)
        sched = torch.empty(3, dtype=torch.float32, device=dev)
        # every step's batch indices up front (int32, steps x batch) and a device
        # step counter: the gather reads both, so one captured step replays as is
        idx = torch.randint(0, n, (steps, batch), device=dev, generator=gen, dtype=torch.int32)
        counter = torch.zeros(1, dtype=torch.int64, device=dev)
        xb = torch.empty((batch, xn.shape[1]), dtype=torch.float32, device=dev)
        tb = torch.empty((batch, tn.shape[1]), dtype=torch.float32, device=dev)
        loss_acc = torch.zeros(1, dtype=torch.float64, device=dev)
        lib = nat.lib()

        if fused is None:
            fused = self.dims == (16, 256, 256, 256, 16)
        if fused:
            train_step = self._fused_step(xn, tn, idx, batch, steps, table, counter, sched,
                                          loss_acc, splits, lr, beta1, beta2)
        else:
            train_step = None

        def layered_step():
            st = lib.fsk_mlp_gather_batch(xn.data_ptr(), tn.data_ptr(), n, xn.shape[1],
                                          tn.shape[1], idx.data_ptr(), batch, table.data_ptr(),
                                          steps, counter.data_ptr(), xb.data_ptr(), tb.data_ptr(),
                                          sched.data_ptr(), self._stream())
            nat.check(st, "fsk_mlp_gather_batch")
            cache = []
            dy = self.forward_loss(xb, tb, loss_acc, cache)
            grads = self.backward(dy, cache, splits)
            self.adam_step(grads, lr, beta1, beta2, sched=sched, counter=counter)

        train_step = train_step or layered_step
        # One graph holds `chunk` consecutive steps (the largest divisor of
        # log_every up to 16), so the host issues one replay per chunk: a
        # step takes less GPU time than a replay call from Python.
        chunk = max(d for d in range(1, 17) if log_every % d == 0)
        graphs = {}

        def replay(k):
            if k not in graphs:
                torch.cuda.synchronize(dev)
                g = torch.cuda.CUDAGraph()
                side = torch.cuda.Stream(device=dev)
                side.wait_stream(torch.cuda.current_stream(dev))
                with torch.cuda.stream(side):
                    with torch.cuda.graph(g, stream=side):
                        for _ in range(k):
                            train_step()
                torch.cuda.current_stream(dev).wait_stream(side)
                self.step_count -= k  # capture records, it does not run
                graphs[k] = g
            graphs[k].replay()
            self.step_count += k

        history = []
        s, since_log = 0, 0
        while s < steps:
            if not graph or s < 2:
                train_step()  # eager (also the warm-up before capture)
                k = 1
            else:
                # whole chunks from a multiple of `chunk`, single steps elsewhere
                k = chunk if (s % chunk == 0 and steps - s >= chunk) else 1
                replay(k)
            s += k
            since_log += k
            if s % log_every == 0 or s == steps:
                history.append(float(loss_acc.item()) / (since_log * batch * tn.shape[1]))
                loss_acc.zero_()
                since_log = 0
        torch.cuda.synchronize(dev)
        return history

    def _fused_step(self, xn, tn, idx, batch, steps, table, counter, sched, loss_acc, splits,
                    lr, beta1, beta2):
        """Buffers and argument records of the three-launch training step."""
        dev = self.device
        H = self.dims[1]
        f32 = torch.float32
        bufs = {"xb": torch.empty((batch, 16), dtype=f32, device=dev),
                # the activations as BF16: what the next layer and the weight
                # gradients multiply (the step kernel stores them rounded)
                "h": [torch.empty((batch, H), dtype=torch.bfloat16, device=dev) for _ in range(3)],
                "dz": [torch.empty((batch, H), dtype=f32, device=dev) for _ in range(3)]
                + [torch.empty((batch, 16), dtype=f32, device=dev)]}
        grads = []
        wl = (nat.WgradLayer * len(self.layers))()
        inputs = [bufs["xb"]] + bufs["h"]
        for l, L in enumerate(self.layers):
            S = splits or self.auto_splits(batch, L["N"], L["K"])
            dWp = torch.empty((S, L["N"], L["K"]), dtype=f32, device=dev)
            dbp = torch.empty((S, L["N"]), dtype=f32, device=dev)
            grads.append((dWp, dbp))
            wl[l] = nat.WgradLayer(bufs["dz"][l].data_ptr(), inputs[l].data_ptr(),
                                   inputs[l].stride(0), L["N"], L["K"], S,
                                   int(inputs[l].dtype == torch.bfloat16), dWp.data_ptr(),
                                   dbp.data_ptr())
        a = nat.RowsStepArgs()
        a.x, a.t, a.n = xn.data_ptr(), tn.data_ptr(), xn.shape[0]
        a.idx, a.batch, a.steps = idx.data_ptr(), batch, steps
        a.table, a.counter, a.sched = table.data_ptr(), counter.data_ptr(), sched.data_ptr()
        a.in_dim, a.hidden, a.out_dim = self.dims[0], H, self.dims[-1]
        for l, L in enumerate(self.layers):
            a.w_bf16[l], a.wt_bf16[l], a.bias[l] = L["Wb"].data_ptr(), L["Wt"].data_ptr(), \
                L["b"].data_ptr()
            a.dz[l] = bufs["dz"][l].data_ptr()
        for l in range(3):
            a.h[l] = bufs["h"][l].data_ptr()
        a.xb = bufs["xb"].data_ptr()
        a.mse_scale = 2.0 / (batch * self.dims[-1])
        a.loss = loss_acc.data_ptr()
        lib = nat.lib()

        def step():
            nat.check(lib.fsk_mlp_rows_step(ctypes.byref(a), self._stream()), "fsk_mlp_rows_step")
            nat.check(lib.fsk_mlp_wgrad_multi(ctypes.byref(wl), len(self.layers), batch,
                                              self._stream()), "fsk_mlp_wgrad_multi")
            self.adam_step(grads, lr, beta1, beta2, sched=sched, counter=counter)

        step.keep = (a, wl, bufs, grads)  # records and buffers live as long as the step
        return step

    def predict(self, x: torch.Tensor, chunk: int = 1 << 20) -> torch.Tensor:
        out = torch.empty((x.shape[0], self.dims[-1]), dtype=torch.float32, device=self.device)
        for a in range(0, x.shape[0], chunk):
            xn = ((x[a:a + chunk].float() - self.x_mean) * self.x_scale).contiguous()
            out[a:a + chunk] = self.forward(xn) * self.y_scale
        return out

    # ---- checkpoint ----
    def state_dict(self) -> dict:
        sd = {"dims": np.asarray(self.dims), "x_mean": self.x_mean.cpu().numpy(),
              "x_scale": self.x_scale.cpu().numpy(), "y_scale": np.asarray(self.y_scale)}
        for l, L in enumerate(self.layers):
            sd[f"W{l}"] = L["W"].cpu().numpy()
            sd[f"b{l}"] = L["b"].cpu().numpy()
        return sd

    def save(self, path: str) -> None:
        np.savez(path, **self.state_dict())

    @classmethod
    def load(cls, path: str, device="cuda") -> "ScoreMLP":
        with np.load(path, allow_pickle=False) as z:
            net = cls(dims=tuple(int(d) for d in z["dims"]), device=device)
            for l, L in enumerate(net.layers):
                L["W"].copy_(torch.as_tensor(z[f"W{l}"]))
                L["b"].copy_(torch.as_tensor(z[f"b{l}"]))
            net.x_mean = torch.as_tensor(z["x_mean"], device=net.device)
            net.x_scale = torch.as_tensor(z["x_scale"], device=net.device)
            net.y_scale = float(z["y_scale"])
        net._cast()
        return net


def exact_shift(rows: torch.Tensor, data: torch.Tensor, h: float, eps: float = 1e-12, *,
                precision: str = "fp32", rows_in_data: bool = False):
    """delta = w/(pdf + eps) - x of `rows` under the KDE of `data` (score kernel
    in `precision`, FP64 debias): x + delta/2 is the exact one-step SD-KDE
    point. rows_in_data: every row is one of the data points (the 16-bit
    forms may then leave out negligible tile pairs). A 16-bit form whose
    range the inputs exceed (ValueError from the FP16 range check) falls back
    to the FP32 form."""
    try:
        pdf, w = _ops.score16_rows(rows, data, h, precision=precision, rows_in_data=rows_in_data)
    except ValueError:
        if precision == "fp32":
            raise
        pdf, w = _ops.score16_rows(rows, data, h)
    deb = _ops.debias(rows, pdf, w, h, eps)
    return 2.0 * (deb - rows)


def surrogate_fit_transform(data, bandwidth: float, *, n_label: int = 262144,
                            n_holdout: int = 16384, n_probe: int = 16384, tol: float = 1e-3,
                            check_margin: float = 0.5, steps: int = 8000, batch: int = 8192,
                            lr: float = 1e-2, seed: int = 0, device="cuda",
                            net: ScoreMLP | None = None, label_precision: str = "fp16x3s"):
    """One-step SD-KDE debiasing with the learned shift field.

    The training labels come from the score kernel in `label_precision`
    (default fp16x3s: FP32-class log-density accuracy at three times the FP32
    form's rate; the label rows are data points, so each row's self pair is
    in its sums; FP32 when the inputs leave FP16's range). The held-out check
    and the fallback use the FP32 form.

    Returns (debiased (n,16) tensor, info dict). info["fallback"] is True when
    the held-out log-density error exceeded check_margin * tol and the exact
    score kernel produced the result instead."""
    device = torch.device(device)
    x = _ops.as_points16(data, device)
    n = x.shape[0]
    h = float(bandwidth)
    info = {}
    rng = np.random.default_rng(seed)
    perm = torch.as_tensor(rng.permutation(n), device=device)
    n_label = min(n_label, max(1, n // 2))
    n_holdout = min(n_holdout, max(1, (n - n_label) // 2))
    n_probe = min(n_probe, max(1, n - n_label - n_holdout))
    hold = x[perm[n_label:n_label + n_holdout]].contiguous()
    probe = x[perm[n_label + n_holdout:n_label + n_holdout + n_probe]].contiguous()

    t0 = time.perf_counter()
    # labels: a random subset of the data, both in spatial order, so the
    # label kernel leaves out the tile pairs whose terms are negligible
    order = spatial_order(x)
    xs = x[order]
    inv = torch.empty_like(order)
    inv[order] = torch.arange(n, device=device)
    lab = xs[torch.sort(inv[perm[:n_label]]).values].contiguous()
    target = exact_shift(lab, xs, h, precision=label_precision, rows_in_data=True)
    torch.cuda.synchronize(device)
    info["label_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    net = net or ScoreMLP(seed=seed, device=device)
    info["loss"] = net.fit(lab, target, steps=steps, batch=batch, lr=lr, seed=seed)
    torch.cuda.synchronize(device)
    info["train_s"] = time.perf_counter() - t0

    # held-out check: KDE of the held-out set shifted exactly vs by the surrogate
    t0 = time.perf_counter()
    d_exact = exact_shift(hold, x, h)
    d_hat = net.predict(hold)
    lp_exact = _ops.kde16_logdensity(hold + 0.5 * d_exact, probe, h, precision="fp32")
    lp_hat = _ops.kde16_logdensity(hold + 0.5 * d_hat, probe, h, precision="fp32")
    rel = ((lp_hat - lp_exact).abs() / lp_exact.abs().clamp_min(1e-300))
    info["holdout_max_rel_logp_err"] = float(rel.max())
    info["holdout_shift_rmse_over_h"] = float(((d_hat - d_exact).pow(2).sum(1).mean().sqrt()) / h)
    torch.cuda.synchronize(device)
    info["check_s"] = time.perf_counter() - t0
    info["fallback"] = info["holdout_max_rel_logp_err"] > check_margin * tol
    t0 = time.perf_counter()
    if info["fallback"]:
        pdf, w = _ops.score16(x, h, variant="symmetric")
        deb = _ops.debias(x, pdf, w, h, 1e-12)
    else:
        deb = x + 0.5 * net.predict(x)
    torch.cuda.synchronize(device)
    info["transform_s"] = time.perf_counter() - t0
    info["net"] = net
    return deb, info
