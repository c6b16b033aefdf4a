#!/usr/bin/env python3
"""Benchmark: 16-D SD-KDE Gpair-evals/s on one MI355X.

One step = the plain-FP32 SD-KDE pass at n = 1,048,576 training points drawn
from an 8-component isotropic Gaussian mixture (means ~ N(0, 9 I), sigma 1,
seed 0) and m = 65,536 queries from the same mixture (seed 1):
  1. empirical-score accumulators over all n^2 ordered pairs (symmetric
     kernel by default: each unordered pair once, column sums scattered with
     FP64 atomics; --variant ordered runs the deterministic split-K kernel),
  2. FP64 debias x + (h^2/2) score,
  3. KDE of the m queries against the n debiased points.
value = (m n + n^2) / step time, inputs resident in HBM.

Prints ONE JSON line. --gpus N with N > 1 exits: this build measures one GPU.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "flash-sd-kde_amd"))

METRIC = "SD-KDE 16D Gpair-evals/s at n=1M vs FP32 plain, at max rel err ≤1e-3 vs FP64"
PEAK_FP32_MFMA_TFLOPS = 157.3
# vector issue rates (whole chip) used for the exp/VALU roofline: 256 CU x 2.4 GHz,
# 128 FP32 FMA lanes/clk/CU (MI355X's 157.3 TFLOP/s FP32 vector peak) and 32
# lanes/clk/CU for transcendentals (v_exp_f32: the library's microbenchmark
# measures 1.9e13/s, 31 lanes/clk/CU, against CDNA3's 16).
R_VALU = 256 * 2.4e9 * 128
R_EXP = 256 * 2.4e9 * 32


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--m", type=int, default=65536)
    ap.add_argument("--variant", choices=["symmetric", "ordered"], default="symmetric")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-precision-forms", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-one-d", action="store_true")
    ap.add_argument("--error-queries", type=int, default=4096)
    return ap.parse_args()


def cpu_baseline(x, q, h, budget_s):
    """Time the CPU oracle (numpy FP32 restatement of the reference's
    reference.py) on a bounded sample of the same workload: blocks of 64 score
    rows and 64 KDE queries against all n points, spread over a thread pool
    (BLAS pinned to one thread per worker), until the time budget is spent."""
    from concurrent.futures import ThreadPoolExecutor

    from threadpoolctl import threadpool_limits

    from oracle import reference_np as R

    host_cores = len(os.sched_getaffinity(0))  # cores this process may run on
    threads = max(1, min(16, host_cores))  # at most the box's share of 16
    blk = 64
    nblocks = threads * 64
    pairs = 0
    t0 = time.perf_counter()
    with threadpool_limits(limits=1):
        with ThreadPoolExecutor(threads) as ex:
            def work(b):
                if time.perf_counter() - t0 > budget_s:
                    return 0
                a = (b * blk) % (x.shape[0] - blk)
                if b % 2 == 0:
                    R.score_rows(x[a:a + blk], x, h, block=blk)
                else:
                    qa = (b * blk) % max(1, q.shape[0] - blk)
                    R.kde_sums_nd(q[qa:qa + blk], x, h, block=blk)
                return blk * x.shape[0]
            for p in ex.map(work, range(nblocks)):
                pairs += p
    dt = time.perf_counter() - t0
    out = {
        "value": pairs / dt / 1e9,
        "unit": "Gpair-evals/s",
        "cores": threads,
        "host_cores": host_cores,
        "kind": "port",
        "sample": (f"{pairs / (blk * x.shape[0]):.0f} blocks of {blk} rows (score rows and KDE "
                   f"queries alternating) x all n={x.shape[0]} points, {pairs:.3g} pairs in "
                   f"{dt:.1f} s; numpy FP32 oracle, {threads} threads (of {host_cores} host "
                   f"cores visible), BLAS 1 thread each"),
    }
    out["sklearn"] = sklearn_baseline(x, q, h, budget_s / 2)
    return out


def sklearn_baseline(x, q, h, budget_s):
    """scikit-learn KernelDensity (gaussian, exact: rtol = atol = 0) fitted on
    all n points and scored on a bounded sample of the queries, as the
    reference's runtime benchmark times it
    (experiments/runtime/benchmark_triton_kde.py:63-74); one thread."""
    try:
        from sklearn.neighbors import KernelDensity
    except ImportError:
        return None
    t0 = time.perf_counter()
    kd = KernelDensity(kernel="gaussian", bandwidth=h, rtol=0.0, atol=0.0).fit(x)
    fit_s = time.perf_counter() - t0
    done, k = 0, 4
    t1 = time.perf_counter()
    while done < q.shape[0] and time.perf_counter() - t1 < budget_s:
        kd.score_samples(q[done:done + k])
        done += k
        k = min(2 * k, 256)
    dt = time.perf_counter() - t1
    return {"value": done * x.shape[0] / dt / 1e9, "unit": "Gpair-evals/s (KDE queries)",
            "cores": 1, "sample": f"{done} queries x n={x.shape[0]} in {dt:.1f} s (fit {fit_s:.1f} s)"}


# HBM bytes per launch of the dominant kernel, from two rocprofv3 --pmc runs of
# `bench.py --steps 1 --warmup 0 --no-cpu-baseline [--variant ordered] --no-batched
# --no-precision-forms --no-surrogate --no-one-d` (FETCH_SIZE, then WRITE_SIZE, KiB
# per dispatch), each scaled by what the same counter reads for the library's
# copy microbenchmark in the same run (512 MiB read and 512 MiB written per
# dispatch: FETCH_SIZE reads 262168 KiB -> scale 2.0, WRITE_SIZE reads
# 524288 KiB -> scale 1.0). Keyed by (score kernel, n); other sizes report null.
# The symmetric kernel's figure is dominated by its FP64 device-scope atomic
# scatters, which are performed past the L2 (8 bytes each).
PMC_TRAFFIC = {
    ("ordered", 1 << 20): {"fetch_kib": 2354213.0, "write_kib": 278528.0,
                           "fetch_scale": 524288.0 / 262168.0, "write_scale": 1.0},
    ("symmetric", 1 << 20): {"fetch_kib": 8820297.0, "write_kib": 160432128.0,
                             "fetch_scale": 524288.0 / 262168.0, "write_scale": 1.0},
}


def pmc_traffic_bytes(variant, n):
    t = PMC_TRAFFIC.get((variant, n))
    if t is None:
        return None
    return round((t["fetch_kib"] * t["fetch_scale"] + t["write_kib"] * t["write_scale"]) * 1024.0)


def measure_rates():
    """Issue rates measured on this GPU by the library's microbenchmarks."""
    import ctypes

    import torch

    from flash_sd_kde import _native

    lib = _native.lib()
    out = torch.empty(2048 * 256, device="cuda")
    r = ctypes.c_double()
    rates = {}
    for which, name, iters in ((0, "fp32_mfma_flops", 20000), (1, "bf16_mfma_flops", 20000),
                               (2, "exp_per_s", 20000), (3, "fma_per_s", 20000)):
        lib.fsk_microbench(which, 2048, 50, out.data_ptr(), ctypes.addressof(r), None)
        st = lib.fsk_microbench(which, 2048, iters, out.data_ptr(), ctypes.addressof(r), None)
        rates[name] = r.value if st == 0 else None
    # MFMA / VALU co-issue: 4 FP32 MFMAs + 16 exps (or 16 FMAs) per lane per
    # iteration against the two parts timed alone (2048 workgroups, 20000
    # iterations). A ratio near 1 means the two issue times add up; near
    # max/sum they overlap.
    # The same with 4 BF16 16x16x32 MFMAs (modes 7, 8), and 8 exps with 8 FMAs
    # (mode 9: whether transcendental and plain VALU issue overlap).
    t = {}
    for which in (0, 1, 2, 3, 4, 6, 7, 8, 9):
        lib.fsk_microbench(which, 2048, 50, out.data_ptr(), ctypes.addressof(r), None)
        if lib.fsk_microbench(which, 2048, 20000, out.data_ptr(), ctypes.addressof(r), None) == 0:
            t[which] = r.value
    if len(t) == 9:
        waves = 2048 * 4
        t_mfma = waves * 20000 * 4 * 2048 / t[0]
        t_mfma16 = waves * 20000 * 4 * 16384 / t[1]
        t_exp8 = waves * 64 * 20000 * 8 / t[2]
        t_fma8 = waves * 64 * 20000 * 8 / t[3]
        rates["mfma_exp_mix_time_over_sum"] = round(t[4] / (t_mfma + 2 * t_exp8), 4)
        rates["mfma_exp_mix_time_over_max"] = round(t[4] / max(t_mfma, 2 * t_exp8), 4)
        rates["mfma_fma_mix_time_over_sum"] = round(t[6] / (t_mfma + 2 * t_fma8), 4)
        rates["mfma16_exp_mix_time_over_sum"] = round(t[7] / (t_mfma16 + 2 * t_exp8), 4)
        rates["mfma16_exp_mix_time_over_max"] = round(t[7] / max(t_mfma16, 2 * t_exp8), 4)
        rates["mfma16_fma_mix_time_over_sum"] = round(t[8] / (t_mfma16 + 2 * t_fma8), 4)
        rates["exp_fma_mix_time_over_sum"] = round(t[9] / (t_exp8 + t_fma8), 4)
    # HBM copy: 512 MiB read + 512 MiB written per launch (2^20 blocks x 1 KiB)
    buf = torch.empty(1 << 28, device="cuda")
    lib.fsk_microbench(5, 1 << 20, 1, buf.data_ptr(), ctypes.addressof(r), None)
    st = lib.fsk_microbench(5, 1 << 20, 10, buf.data_ptr(), ctypes.addressof(r), None)
    rates["hbm_copy_bytes_per_s"] = r.value if st == 0 else None
    del buf
    return rates


def batched_loo(rates, steps):
    """Batched LOO bandwidth grid: 16 datasets, n_i log-uniform in [2048, 262144]
    (seed 0), 4-component mixtures (means ~ N(0, 4 I), sigma 1, seed 100 + i),
    32 bandwidths log-spaced in [0.05, 2] x the mean-std rule; one persistent
    launch replayed from a HIP graph, launch shape from the autotuner."""
    import numpy as np
    import torch

    from flash_sd_kde.batched import LooBandwidthGrid
    from flash_sd_kde.datasets import gmm_isotropic, uneven_sizes

    sizes = uneven_sizes(16, 2048, 262144, seed=0)
    sets = [gmm_isotropic(n, k=4, mean_scale=2.0, seed=100 + i) for i, n in enumerate(sizes)]
    t0 = time.perf_counter()
    grid = LooBandwidthGrid(sets, factors=np.geomspace(0.05, 2.0, 32), autotune=True, graph=True)
    setup_s = time.perf_counter() - t0
    grid.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        grid.run()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    pairs = grid.pair_count()
    exps = pairs * 32
    done = grid.exps_evaluated()
    exp_rate = rates.get("exp_per_s") or 17e12
    fma_rate = rates.get("fma_per_s") or 54e12
    # per evaluated exp: one multiply and one add on the VALU beside it
    t_roof = done / exp_rate + 2 * done / fma_rate
    ll = grid.run().cpu().numpy()
    return {
        "workload": f"16 datasets, sum n = {sum(sizes)}, 32 bandwidths, LOO",
        "ms_per_launch": round(ms, 3),
        "gpair_evals_per_s": round(pairs / ms / 1e6, 3),
        "gexp_per_s": round(exps / ms / 1e6, 3),
        "exps_evaluated_frac": round(done / exps, 4),
        "gexp_evaluated_per_s": round(done / ms / 1e6, 3),
        "exp_valu_frac": round(t_roof / (ms * 1e-3), 4),
        "tune": list(grid.tune),
        "tune_ms": {f"{w}w{rt}r{g}g": round(t, 3)
                    for (w, rt, g), t in getattr(grid, "tune_times", {}).items()},
        "setup_s": round(setup_s, 2),
        "best_factor_median": float(np.median(np.argmax(ll, axis=1))),
    }


def surrogate_form(fp32_pair_rate):
    """Learned-score form: n = 4,194,304 points of a 32-component
    full-covariance mixture (eigenvalues log-U[0.1, 3], random rotations,
    means ~ N(0, 16 I), seed 0), 65,536 held-out queries (seed 1). The MLP is
    trained on exact FP32 shift labels of 262,144 rows; its debiasing is
    compared with exact SD-KDE debiasing of all n points (log-densities of the
    held-out queries, FP32 log-sum-exp KDE)."""
    import numpy as np
    import torch

    from flash_sd_kde import _ops
    from flash_sd_kde.bandwidth import scott_mean_std_bandwidth
    from flash_sd_kde.datasets import gmm_full_cov, gmm_full_cov_logpdf, gmm_full_cov_params
    from flash_sd_kde.surrogate import surrogate_fit_transform

    params = gmm_full_cov_params(32, mean_scale=4.0, seed=0)
    x = gmm_full_cov(1 << 22, params=params, seed=0)
    q = gmm_full_cov(65536, params=params, seed=1)
    h = scott_mean_std_bandwidth(x)
    xd = torch.as_tensor(x, device="cuda")
    qd = torch.as_tensor(q, device="cuda")
    n = xd.shape[0]
    # one untimed run first, as for every other section: the first call in a
    # process also loads the MLP kernels and instantiates the training graphs
    # (the exact score kernel it is compared with is already warm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    surrogate_fit_transform(xd, h, tol=1e-3)
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t0
    t0 = time.perf_counter()
    deb_s, info = surrogate_fit_transform(xd, h, tol=1e-3)
    torch.cuda.synchronize()
    t_sur = time.perf_counter() - t0
    t0 = time.perf_counter()
    pdf, w = _ops.score16(xd, h, variant="symmetric")  # the fastest exact FP32 score
    deb_e = _ops.debias(xd, pdf, w, h, 1e-12)
    torch.cuda.synchronize()
    t_exact = time.perf_counter() - t0
    lp_e = _ops.kde16_logdensity(deb_e, qd, h)
    lp_s = _ops.kde16_logdensity(deb_s, qd, h)
    lp_k = _ops.kde16_logdensity(xd, qd, h)
    r = ((lp_s - lp_e).abs() / lp_e.abs()).cpu().numpy()
    rk = ((lp_k - lp_e).abs() / lp_e.abs()).cpu().numpy()
    truth = gmm_full_cov_logpdf(q, params)
    mae = lambda a: float(np.mean(np.abs(a.cpu().numpy() - truth)))
    return {
        "workload": "n=4194304 full-cov GMM32, 65536 queries, MLP 16-256-256-256-16 SiLU bf16",
        "bandwidth": h,
        "exact_debias_s": round(t_exact, 3),
        "surrogate_debias_s": round(t_sur, 3),
        "surrogate_first_call_s": round(t_first, 3),
        "speedup_vs_exact": round(t_exact / t_sur, 3),
        "label_s": round(info["label_s"], 3), "train_s": round(info["train_s"], 3),
        "check_s": round(info["check_s"], 3), "transform_s": round(info["transform_s"], 3),
        "final_train_loss": info["loss"][-1],
        "holdout_max_rel_logp_err": info["holdout_max_rel_logp_err"],
        "fallback_to_exact": info["fallback"],
        "max_rel_logp_err_vs_exact_sdkde": float(r.max()),
        "median_rel_logp_err_vs_exact_sdkde": float(np.median(r)),
        "plain_kde_max_rel_logp_err_vs_exact_sdkde": float(rk.max()),
        "mean_abs_logp_err_vs_truth": {"exact_sdkde": mae(lp_e), "surrogate_sdkde": mae(lp_s),
                                       "plain_kde": mae(lp_k)},
    }


def kde1_visited_pairs(data, queries, h):
    """Pairs the 1-D plain and score kernels evaluate: per block of 2048
    sorted queries, the 1024-point chunks of the sorted data that meet the
    union of the queries' reaches [q - r, q + r], r = min(sqrt(d_nn^2 +
    T / k^2), 12.8 / k) with d_nn the distance to the nearest data point,
    T = ceil(log2 n) + 32 and k = sqrt(log2(e) / (2 h^2)) (terms more than
    2^-T below a query's largest are left out, and beyond 12.8 / k every term
    is exactly 0 in FP32); the library's rule, in FP64."""
    import numpy as np

    xs = np.sort(np.asarray(data, dtype=np.float32)).astype(np.float64)
    qs = np.sort(np.asarray(queries, dtype=np.float32)).astype(np.float64)
    k = math.sqrt(math.log2(math.e) / (2.0 * h * h))
    n, rows, chunk = xs.size, 2048, 1024
    t = math.ceil(math.log2(n)) + 32.0
    pos = np.searchsorted(xs, qs, side="left")
    right = np.where(pos < n, xs[np.minimum(pos, n - 1)] - qs, np.inf)
    left = np.where(pos > 0, qs - xs[np.maximum(pos - 1, 0)], np.inf)
    dnn = np.minimum(right, left)
    r = np.minimum(np.sqrt(dnn * dnn + t / (k * k)), 12.8 / k)
    n_pad = -(-n // chunk) * chunk
    starts = np.arange(0, qs.size, rows)
    lo = np.minimum.reduceat(qs - r, starts)
    hi = np.maximum.reduceat(qs + r, starts)
    jlo = np.searchsorted(xs, lo, side="left")
    jhi = np.searchsorted(xs, hi, side="right")
    c0 = (jlo // chunk) * chunk
    c1 = np.minimum(-(-jhi // chunk) * chunk, n_pad)
    width = np.where(jhi > jlo, c1 - c0, 0)
    nrows = np.minimum(starts + rows, qs.size) - starts
    return float(np.sum(width * nrows))


def one_d_sdkde(rates, steps):
    """1-D SD-KDE as in the reference's Triton scaling sweep
    (experiments/configs/runtime/run_triton_scaling.sh, n = 2^21, m = n/8):
    empirical score of every point against all points, debias x + (h^2/2)
    score, KDE of m queries; data N(0, 1) (seed 0), queries seed 1, bandwidth
    0.4 min(std, IQR/1.34) n^(-1/9) (the reference's 1-D SD-KDE rule)."""
    import numpy as np
    import torch

    from flash_sd_kde import _native, _ops

    n, m = 1 << 21, 1 << 18
    rng = np.random.default_rng(0)
    x = rng.standard_normal(n).astype(np.float32)
    q = np.random.default_rng(1).standard_normal(m).astype(np.float32)
    iqr = float(np.percentile(x, 75) - np.percentile(x, 25))
    h = 0.4 * min(float(x.std()), iqr / 1.34) * n ** (-1.0 / 9.0)
    xd = torch.as_tensor(x, device="cuda")
    qd = torch.as_tensor(q, device="cuda")

    def step():
        pdf, der = _ops.kde1_score_sums(xd, xd, h)
        deb = xd + (0.5 * h) * (der / pdf)  # (h^2/2) * score, score = der / (h pdf)
        return _ops.kde1_sums(deb, qd, h, atomic=True), deb

    step()
    torch.cuda.synchronize()
    lib = _native.lib()
    lib.fsk_profile_enable(1)
    t0 = time.perf_counter()
    for _ in range(steps):
        out, deb = step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    score_ms = lib.fsk_profile_last_ms(_native.OP_KDE1_SCORE)
    kde_ms = lib.fsk_profile_last_ms(_native.OP_KDE1_ATOMIC)
    lib.fsk_profile_enable(0)
    pairs_s, pairs_k = float(n) * n, float(n) * m
    vis_s = kde1_visited_pairs(x, x, h)
    vis_k = kde1_visited_pairs(deb.cpu().numpy(), q, h)
    res = {"workload": f"1-D SD-KDE, n = {n}, m = {m}, N(0,1)", "bandwidth": h,
           "ms_per_step": round(ms, 3), "gpair_evals_per_s": round((pairs_s + pairs_k) / ms / 1e6, 3),
           "score_kernel_ms": round(score_ms, 3), "kde_kernel_ms": round(kde_ms, 3),
           "visited_pair_frac": {"score": round(vis_s / pairs_s, 4), "kde": round(vis_k / pairs_k, 4)},
           "outputs_finite": bool(torch.isfinite(out).all().item())}
    if rates.get("exp_per_s") and rates.get("fma_per_s"):
        # per visited pair: score 1 exp + 5 VALU (sub, mul by k, square, add,
        # fma); KDE 1 exp + 4 VALU. The transcendental issue overlaps the other
        # VALU work in these kernels (the additive exp + VALU time exceeds the
        # measured time), so each resource is reported as its own fraction; the
        # larger one binds.
        for name, pairs, v_ops, kms in (("score", vis_s, 5, score_ms), ("kde", vis_k, 4, kde_ms)):
            t_exp = pairs / rates["exp_per_s"]
            t_valu = v_ops * pairs / rates["fma_per_s"]
            res[f"{name}_roofline"] = {
                "exp_frac": round(t_exp / (kms * 1e-3), 4),
                "valu_frac": round(t_valu / (kms * 1e-3), 4),
                "bound": "valu" if t_valu >= t_exp else "exp"}
    return res


# Reduced-precision forms timed beside the FP32 step: (name, score precision,
# log-density precision). The first keeps the FP32 score and evaluates the
# log-density cross term on the vector ALU (FP32 FMAs). The last one mixes
# forms: the log-density error is set by the KDE's cross term far more than by
# the score's, so a single-FP16 score under the FP16-split KDE stays within
# the error bound.
PRECISION_FORMS = (("fp32_valu", "fp32_valu", "fp32_valu"),
                   ("fp16x3", "fp16x3", "fp16x3"), ("fp16x3s", "fp16x3s", "fp16x3"),
                   ("bf16x6", "bf16x6", "bf16x6"), ("bf16x3", "bf16x3", "bf16x3"),
                   ("bf16x3s", "bf16x3s", "bf16x3"), ("fp16", "fp16", "fp16"),
                   ("bf16", "bf16", "bf16"), ("fp16+fp16x3", "fp16", "fp16x3"))
ERROR_BANDWIDTHS = (0.5, 1.0, 2.0)  # multiples of the workload's bandwidth
# Instruction mix of each score form per ordered pair: (16-bit MFMA flop, VALU
# ops besides the one exp). A 16x16x32 MFMA covers 256 pairs (64 flop each);
# data tiles go in pairs, so the single forms need 1 cross-term MFMA and half
# a phi.X MFMA per tile, the one-phi forms 2 + 1, the split x3 forms 2 + 1.5,
# BF16X6 4 + 1.5. VALU per pair: single forms half a v_cvt_pk and half a
# v_dot2 (row sums); one-phi forms the row-sum add and half a v_cvt_pk; split
# forms the row-sum add, and + sub (low part), half a v_perm and half a
# v_cvt_pk; the FP32-VALU form 16 FMAs of the cross term, the FMA applying c,
# the row-sum add and 16 FMAs of phi.X (no MFMA).
SCORE_ISSUE_MIX = {"fp16": (96.0, 1.0), "bf16": (96.0, 1.0), "fp16x3s": (192.0, 1.5),
                   "bf16x3s": (192.0, 1.5), "fp16x3": (224.0, 4.0), "bf16x3": (224.0, 4.0),
                   "bf16x6": (352.0, 4.0), "fp32_valu": (0.0, 34.0)}


def score_variant(prec):
    """The FP32 score runs on the symmetric kernel (as the FP32 step), the
    16-bit forms on the ordered one."""
    return "symmetric" if prec == "fp32" else "ordered"


def negligible_tile_frac(xd, h, t_max, samples=96):
    """Sampled share of the (16-row tile, 32-point tile pair) blocks of the
    data in spatial order whose every exponent -|x_i - x_j|^2 log2(e)/(2h^2)
    is at or below t_max (FP64 distances; a sample of row tiles against all
    tile pairs)."""
    import torch

    from flash_sd_kde.spatial import spatial_order

    xs = xd[spatial_order(xd)].double()
    n = xs.shape[0]
    c = math.log2(math.e) / (2 * h * h)
    npair = n // 32
    far = total = 0
    g = torch.Generator().manual_seed(0)
    for r in torch.randint(0, n // 16, (samples,), generator=g).tolist():
        d2 = torch.cdist(xs[16 * r:16 * r + 16], xs[:npair * 32]) ** 2
        blk = (-c * d2).reshape(16, npair, 32).amax(dim=(0, 2))
        far += int((blk <= t_max).sum())
        total += npair
    return far / total


def precision_forms(xd, qd, h, fp32_ms, fp32_step, steps, warmup, n_err, rates=None):
    """Reduced-precision SD-KDE forms on the same workload: score with 16-bit
    MFMA operands (FP32 accumulate), FP64 debias, log-sum-exp KDE. Errors: max /
    median relative log-density error on the first n_err queries against the
    FP64 GPU reference pipeline (FP64 direct-difference score, debias and
    log-sum-exp), at 0.5, 1 and 2 times the bandwidth; a form is within the
    bound when its max error is at every one of them. Time is for the
    workload's bandwidth. The score kernel's exp/VALU fraction uses the
    algorithmic minimum per ordered pair and the issue rates measured in this
    run."""
    import math

    import torch

    from flash_sd_kde import _native, _ops, precision

    n, m = xd.shape[0], qd.shape[0]
    sub = qd[:n_err]

    def forms_at(hh):
        """Log-densities of the FP32 step and of every form at bandwidth hh."""
        log_norm = -math.log(n) - 16 * math.log(hh) - 8 * math.log(2 * math.pi)
        res = {"fp32": torch.log(fp32_step(hh)[:n_err].double()) + log_norm}
        for name, sp, ep in PRECISION_FORMS:
            res[name] = precision.sdkde_logdensity(xd, sub, hh, precision=ep, score_precision=sp,
                                                   device=xd.device)
        return res

    errs = {}
    t0 = time.perf_counter()
    for mult in ERROR_BANDWIDTHS:
        ref = precision.ref64_sdkde_logdensity(xd, sub, h * mult, device=xd.device)
        for name, lp in forms_at(h * mult).items():
            errs.setdefault(name, {})[mult] = precision.rel_log_error(lp, ref)
        del ref
    torch.cuda.synchronize()
    err_s = time.perf_counter() - t0

    def err_fields(name):
        e = errs[name]
        worst = max(v[0] for v in e.values())
        return {"max_rel_logp_err": e[1.0][0], "median_rel_logp_err": e[1.0][1],
                "max_rel_logp_err_by_h": {f"{k}x": v[0] for k, v in e.items()},
                "within_1e-3": worst <= 1e-3}

    out = {"reference": f"FP64 GPU pipeline on {n_err} queries at "
                        f"{'/'.join(str(k) for k in ERROR_BANDWIDTHS)} x h "
                        f"(all forms' errors in {err_s:.1f} s)",
           "fp32": {"ms_per_step": round(fp32_ms, 3), "speedup_vs_fp32": 1.0, **err_fields("fp32")}}
    lib = _native.lib()
    pairs = float(m) * n + float(n) * n
    dense_ms, tile_frac = {}, {}
    for name, sp, ep in PRECISION_FORMS:
        def step():
            return precision.sdkde_logdensity(xd, qd, h, precision=ep, score_precision=sp,
                                              device=xd.device)
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        lib.fsk_profile_enable(1)
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / steps * 1e3
        score_ms = lib.fsk_profile_last_ms(_native.OP_SCORE16_SYM if score_variant(sp) ==
                                           "symmetric" else _native.OP_SCORE16_ORDERED)
        kde_ms = lib.fsk_profile_last_ms(_native.OP_KDE16_LOGDENSITY)
        lib.fsk_profile_enable(0)
        out[name] = {"score_precision": sp, "kde_precision": ep,
                     "ms_per_step": round(ms, 3), "gpair_evals_per_s": round(pairs / ms / 1e6, 3),
                     "speedup_vs_fp32": round(fp32_ms / ms, 3), "score_kernel_ms": round(score_ms, 3),
                     "kde_kernel_ms": round(kde_ms, 3), **err_fields(name)}
        if sp in precision.SKIP_PRECISIONS or ep in precision.LSE_SKIP_PRECISIONS:
            # the same kernels evaluating every term (FSK_FLAG_DENSE)
            if (sp, ep) not in dense_ms:
                lib.fsk_profile_enable(1)
                precision.sdkde_logdensity(xd, qd, h, precision=ep, score_precision=sp, dense=True,
                                           device=xd.device)
                torch.cuda.synchronize()
                dense_ms[(sp, ep)] = (lib.fsk_profile_last_ms(_native.OP_SCORE16_ORDERED),
                                      lib.fsk_profile_last_ms(_native.OP_KDE16_LOGDENSITY))
                lib.fsk_profile_enable(0)
            out[name]["kde_kernel_dense_ms"] = round(dense_ms[(sp, ep)][1], 3)
        if sp in precision.SKIP_PRECISIONS:
            # the sampled share of tile pairs below the score kernel's threshold
            t_skip = -27.0 if sp == "fp16" else -(math.ceil(math.log2(n)) + 29.0)
            out[name]["score_kernel_dense_ms"] = round(dense_ms[(sp, ep)][0], 3)
            if t_skip not in tile_frac:
                tile_frac[t_skip] = negligible_tile_frac(xd, h, t_skip)
            out[name]["negligible_tile_frac_sampled"] = round(tile_frac[t_skip], 4)
            score_ms = dense_ms[(sp, ep)][0]  # the issue fractions below are the dense kernel's
        if (rates and rates.get("exp_per_s") and rates.get("fma_per_s")
                and rates.get("bf16_mfma_flops") and score_ms > 0 and sp in SCORE_ISSUE_MIX):
            # issue time of the score kernel at the rates measured in this run,
            # from its instruction mix per ordered pair (SCORE_ISSUE_MIX): the
            # MFMA pipe's time and the VALU's (exps + other VALU ops).
            # score_issue_frac lets the two overlap fully, max(MFMA, exp +
            # VALU); score_additive_issue_frac adds them, which is what the
            # mix microbenchmarks show on this GPU (16-bit MFMA + FMA and exp +
            # FMA run in about the sum of their parts: measured_rates).
            units = float(n) * n
            flop, v_ops = SCORE_ISSUE_MIX[sp]
            t_mfma = flop * units / rates["bf16_mfma_flops"]
            t_valu = units / rates["exp_per_s"] + v_ops * units / rates["fma_per_s"]
            out[name]["score_issue_bound_ms"] = {"mfma": round(t_mfma * 1e3, 2),
                                                 "exp_valu": round(t_valu * 1e3, 2)}
            out[name]["score_issue_frac"] = round(max(t_mfma, t_valu) / (score_ms * 1e-3), 4)
            out[name]["score_additive_issue_frac"] = round((t_mfma + t_valu) / (score_ms * 1e-3), 4)
    ok = [p for p in out if p != "reference" and out[p]["within_1e-3"]]
    best = max(ok, key=lambda p: out[p]["speedup_vs_fp32"])
    out["best_within_1e-3"] = {"form": best, "speedup_vs_fp32": out[best]["speedup_vs_fp32"],
                               "gpair_evals_per_s": out[best].get("gpair_evals_per_s",
                                                                  round(pairs / fp32_ms / 1e6, 3)),
                               "max_rel_logp_err_by_h": out[best]["max_rel_logp_err_by_h"]}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested): this build has "
              "no multi-GPU path; run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from flash_sd_kde import _native, _ops
    from flash_sd_kde.bandwidth import scott_mean_std_bandwidth
    from flash_sd_kde.datasets import gmm_isotropic_split

    dev = torch.device("cuda:0")
    _native.require_gpu(dev, "bench")
    n, m = args.n, args.m
    x, q, _ = gmm_isotropic_split(n, m, k=8, mean_scale=3.0, seed=0)
    h = scott_mean_std_bandwidth(x)
    xd = torch.as_tensor(x, device=dev)
    qd = torch.as_tensor(q, device=dev)

    def step(xs=xd, qs=qd):
        pdf, w = _ops.score16(xs, h, variant=args.variant)
        deb = _ops.debias(xs, pdf, w, h, 1e-12)
        return _ops.kde16_sums(deb, qs, h)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    lib = _native.lib()
    lib.fsk_profile_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dens = step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    op_score = _native.OP_SCORE16_SYM if args.variant == "symmetric" else _native.OP_SCORE16_ORDERED
    score_ms = lib.fsk_profile_last_ms(op_score)
    kde_ms = lib.fsk_profile_last_ms(_native.OP_KDE16_SPLITK)
    lib.fsk_profile_enable(0)

    ms_per_step = dt / args.steps * 1e3
    pairs = float(m) * n + float(n) * n
    value = pairs / (dt / args.steps) / 1e9

    # dominant kernel: the score pass. Algorithmic MFMA work per launch
    # (symmetric): n(n+1)/2 unique pairs x 96 flop (cross term 2d, phi X both
    # ways 4d); ordered: n^2 x 64 flop.
    if args.variant == "symmetric":
        units = n * (n + 1) / 2.0
        flops = units * 96.0
        # per unique pair: 1 exp + 2 VALU ops (row-sum add, column-pdf FMA)
        exp_valu_s = units / R_EXP + units * 2.0 / R_VALU
    else:
        units = float(n) * n
        flops = units * 64.0
        # per ordered pair: 1 exp + 1 VALU op (row-sum add; factored norm terms)
        exp_valu_s = units / R_EXP + units * 1.0 / R_VALU
    achieved = flops / (score_ms * 1e-3) / 1e12 if score_ms > 0 else None
    kde_flops = float(m) * n * 32.0
    kde_tflops = kde_flops / (kde_ms * 1e-3) / 1e12 if kde_ms > 0 else None
    kde_exp_valu_s = float(m) * n / R_EXP + float(m) * n * 3.0 / R_VALU

    finite = bool(torch.isfinite(dens).all().item())

    # one more step with the inputs handed over as host (numpy) arrays and the
    # densities copied back, as a caller of the drop-in API sees it
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dens_host = step(torch.as_tensor(x, device=dev), torch.as_tensor(q, device=dev)).cpu()
    dt_host = time.perf_counter() - t0
    del dens_host
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gpair-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"sdkde16_gmm8_n{n}_m{m}",
            "n_train": n,
            "n_query": m,
            "bandwidth": h,
            "bandwidth_rule": "mean_std*(4/(d+2))^(1/(d+4))*n^(-1/(d+4))",
            "score_kernel": args.variant,
            "precision": "fp32 (MFMA f32 cross term and phi.X on mean-centred points, "
                         "FP64 combine/debias)",
            "parallelism": "single_gpu",
        },
        "roofline": {
            "bound": "mfma",
            "kernel": f"score16_{args.variant}",
            "achieved": round(achieved, 3) if achieved else None,
            "peak": PEAK_FP32_MFMA_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_MFMA_TFLOPS, 4) if achieved else None,
            "traffic": pmc_traffic_bytes(args.variant, n),
            "kernel_ms": round(score_ms, 3),
            "exp_valu_frac": round(exp_valu_s / (score_ms * 1e-3), 4) if score_ms > 0 else None,
        },
        "kde_eval_kernel": {
            "kernel_ms": round(kde_ms, 3),
            "mfma_tflops": round(kde_tflops, 3) if kde_tflops else None,
            "mfma_frac": round(kde_tflops / PEAK_FP32_MFMA_TFLOPS, 4) if kde_tflops else None,
            "exp_valu_frac": round(kde_exp_valu_s / (kde_ms * 1e-3), 4) if kde_ms > 0 else None,
        },
        "outputs_finite": finite,
        "pcie_inclusive": {
            "ms_per_step": round(dt_host * 1e3, 3),
            "gpair_evals_per_s": round(pairs / dt_host / 1e9, 3),
            "note": "one step with X and Q copied from pageable host memory and the "
                    "densities copied back inside the timed region",
        },
    }
    rates = measure_rates()
    line["measured_rates"] = rates
    if rates.get("exp_per_s") and rates.get("fma_per_s") and score_ms > 0:
        # exp/VALU roofline with the measured issue rates (per ordered pair of
        # the ordered kernel in factored form: 1 exp + 1 VALU op, the row-sum
        # add; symmetric: per unique pair 1 exp + 2, the row-sum add and the
        # column-pdf FMA)
        v_ops = 2.0 if args.variant == "symmetric" else 1.0
        t = units / rates["exp_per_s"] + units * v_ops / rates["fma_per_s"]
        line["roofline"]["exp_valu_frac_measured_rates"] = round(t / (score_ms * 1e-3), 4)
        line["roofline"]["mfma_frac_of_measured_peak"] = round(
            flops / (score_ms * 1e-3) / rates["fp32_mfma_flops"], 4)
        if rates.get("fp32_mfma_flops"):
            # FP32 MFMA never co-executes with VALU work on this GPU (counter
            # SQ_VALU_MFMA_COEXEC_CYCLES = 0), so the issue bound of a score
            # kernel is the SUM of its MFMA time and its VALU time. Ordered:
            # per ordered pair 64 flop MFMA, 1 exp and 1 other VALU op (the
            # row-sum add; the factored form starts the accumulator from the
            # data-side norm term, 4 subtractions per 2048 pairs). Symmetric:
            # per unique pair 96 flop, 1 exp, 2 VALU ops, plus per data tile
            # of 2048 unique pairs the column-pdf reduction (5 DPP adds,
            # 6 selects) and the 4 subtractions of the factored form.
            v_pair = 2.0 + 15.0 / 2048 if args.variant == "symmetric" else 1.0 + 4.0 / 2048
            t_add = (flops / rates["fp32_mfma_flops"] + units / rates["exp_per_s"]
                     + v_pair * units / rates["fma_per_s"])
            line["roofline"]["additive_issue_bound_frac"] = round(t_add / (score_ms * 1e-3), 4)
    if not args.no_batched:
        line["batched_loo"] = batched_loo(rates, args.steps)
    if not args.no_precision_forms:
        def fp32_step(hh):
            pdf, w = _ops.score16(xd, hh, variant=args.variant)
            return _ops.kde16_sums(_ops.debias(xd, pdf, w, hh, 1e-12), qd, hh)

        line["precision_forms"] = precision_forms(xd, qd, h, ms_per_step, fp32_step, args.steps,
                                                  args.warmup, args.error_queries, rates)
    if not args.no_one_d:
        line["one_d_sdkde"] = one_d_sdkde(rates, args.steps)
    if not args.no_surrogate:
        del xd, qd
        torch.cuda.empty_cache()
        line["surrogate_form"] = surrogate_form(value)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(x, q, h, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
