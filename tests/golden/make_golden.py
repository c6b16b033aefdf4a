"""Write the golden fixtures (inputs + expected outputs) from the reference's
own CPU functions, flash_sd_kde/reference.py of Flash-SD-KDE.

Run in the build container only, where the reference is mounted read-only:
    python tests/golden/make_golden.py [/root/reference]
It imports the reference's reference.py (numpy only) and writes
tests/golden/reference_vectors.npz. Nothing at test time imports the
reference; the tests read the .npz.

Cases mirror the reference tests' seeds, shapes and bandwidths
(tests/small/test_kde_eval_small.py, tests/large/test_consistency_large.py)
plus the small configurations of the benchmark workloads.
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))


def main(ref_root: str) -> None:
    sys.path.insert(0, ref_root)
    import flash_sd_kde.reference as ref  # the reference's module (numpy only)

    # this repo's generators, loaded by path (the name flash_sd_kde is taken
    # by the reference package imported above)
    import importlib.util

    def load(name):
        path = os.path.join(REPO, "flash-sd-kde_amd", "flash_sd_kde", name + ".py")
        spec = importlib.util.spec_from_file_location("fsk_" + name, path)
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        return mod

    datasets, bandwidth = load("datasets"), load("bandwidth")
    gmm_isotropic_split, plain_normal_16d = datasets.gmm_isotropic_split, datasets.plain_normal_16d
    scott_mean_std_bandwidth = bandwidth.scott_mean_std_bandwidth

    out: dict[str, np.ndarray] = {}

    def put(name, **arrays):
        for k, v in arrays.items():
            out[f"{name}/{k}"] = np.asarray(v)

    # --- reference small tests (seed, shapes, h) ---
    rng = np.random.default_rng(0)
    x = rng.normal(size=256).astype(np.float32)
    q = rng.normal(size=128).astype(np.float32)
    put("kde1d", x=x, q=q, h=0.7, density=ref.kde_eval_1d_numpy(q, x, 0.7))

    rng = np.random.default_rng(1)
    x = rng.normal(size=(256, 16)).astype(np.float32)
    q = rng.normal(size=(64, 16)).astype(np.float32)
    put("kde16", x=x, q=q, h=1.2, density=ref.kde_eval_nd_numpy(q, x, 1.2))

    rng = np.random.default_rng(5)
    x = rng.normal(size=(192, 16)).astype(np.float32)
    q = rng.normal(size=(48, 16)).astype(np.float32)
    put("kde16_lin", x=x, q=q, h=1.05, density=ref.kde_eval_nd_linearized_numpy(q, x, 1.05))

    rng = np.random.default_rng(6)
    x = rng.normal(size=128).astype(np.float32)
    q = rng.normal(size=64).astype(np.float32)
    put("kde1d_lin", x=x, q=q, h=0.85, density=ref.kde_eval_1d_linearized_numpy(q, x, 0.85))

    rng = np.random.default_rng(7)
    x = rng.normal(size=(128, 16)).astype(np.float32)
    q = rng.normal(size=(32, 16)).astype(np.float32)
    put("kde16_lin_b", x=x, q=q, h=1.15, density=ref.kde_eval_nd_linearized_numpy(q, x, 1.15))

    for name, seed, n, h in (("score_a", 2, 128, 0.9), ("score_b", 3, 256, 1.1),
                             ("score_large", 10, 1024, 1.0)):
        rng = np.random.default_rng(seed)
        x = rng.normal(size=(n, 16)).astype(np.float32)
        pdf, w, score = ref.empirical_score_nd_numpy(x, h)
        put(name, x=x, h=h, pdf=pdf, weighted=w, score=score)

    rng = np.random.default_rng(4)
    x = rng.normal(size=(128, 16)).astype(np.float32)
    put("transform", x=x, h=1.0, debiased=ref.empirical_sd_kde_transform_nd_numpy(x, 1.0))

    rng = np.random.default_rng(11)
    x = rng.normal(size=(512, 16)).astype(np.float32)
    q = rng.normal(size=(1024, 16)).astype(np.float32)
    put("kde_chunk", x=x, q=q, h=1.1, density=ref.kde_eval_nd_numpy(q, x, 1.1))

    # --- plain 16-D KDE workload: 4096 train x 1024 queries, seed 0 ---
    x, q = plain_normal_16d(4096, 1024, seed=0)
    h_b = scott_mean_std_bandwidth(x)
    h_s = ref.silverman_bandwidth_nd(x)
    put("cfg_plain", x=x, q=q, h_baseline=h_b, h_silverman=h_s,
        density_baseline=ref.kde_eval_nd_numpy(q, x, h_b),
        density_silverman=ref.kde_eval_nd_numpy(q, x, h_s))

    # --- small mixture SD-KDE workload: n = 4096, 1024 queries ---
    x, q, _ = gmm_isotropic_split(4096, 1024, k=8, mean_scale=3.0, seed=0)
    h = scott_mean_std_bandwidth(x)
    pdf, w, score = ref.empirical_score_nd_numpy(x, h)
    deb = ref.empirical_sd_kde_transform_nd_numpy(x, h)
    put("cfg_gmm_small", x=x, q=q, h=h, pdf=pdf, weighted=w, score=score, debiased=deb,
        density=ref.kde_eval_nd_numpy(q, deb, h), silverman=ref.silverman_bandwidth_nd(x))

    # --- 1-D bandwidth / score path ---
    rng = np.random.default_rng(21)
    x1 = rng.normal(size=512).astype(np.float32)
    h1 = ref.silverman_bandwidth(x1)
    put("one_d", x=x1, h_silverman=h1, score=ref.kde_score_eval(x1, x1, h1),
        pdf=ref.kde_pdf_eval(x1, x1, h1))
    deb1, hdeb = ref.one_step_debiased_data_emp_kde(x1)
    put("one_d_sd", x=x1, debiased=deb1, h=hdeb)

    path = os.path.join(HERE, "reference_vectors.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
