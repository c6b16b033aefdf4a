"""Batched leave-one-out bandwidth grid: CPU validation paths and GPU parity
against the FP64 brute-force oracle, hipGraph replays, autotuner agreement."""

import numpy as np
import pytest
import torch

from flash_sd_kde.batched import LooBandwidthGrid, loo_loglik_grid
from flash_sd_kde.bandwidth import scott_mean_std_bandwidth
from oracle import fp64


def test_batched_rejects_bad_input():
    with pytest.raises(ValueError, match="at least 2"):
        LooBandwidthGrid([np.zeros((1, 16))])
    with pytest.raises(ValueError, match="shape"):
        LooBandwidthGrid([np.zeros((5, 3))])
    with pytest.raises(ValueError, match="datasets"):
        LooBandwidthGrid([])
    with pytest.raises(ValueError, match="positive"):
        LooBandwidthGrid([np.zeros((5, 16))], bandwidths=np.array([[0.0]]))
    with pytest.raises(ValueError, match="bandwidths"):
        LooBandwidthGrid([np.zeros((5, 16))], bandwidths=np.ones((1, 33)))
    bad = np.zeros((5, 16))
    bad[3, 7] = np.nan
    with pytest.raises(ValueError, match="NaN"):
        LooBandwidthGrid([np.zeros((4, 16)), bad])
    with pytest.raises(ValueError, match="tune"):
        LooBandwidthGrid([np.random.default_rng(0).normal(size=(5, 16))], tune=(4, 3, 1))


def _sets(sizes, seed):
    rng = np.random.default_rng(seed)
    out = []
    for n in sizes:
        means = rng.normal(size=(4, 16)) * 2.0
        comp = rng.integers(0, 4, size=n)
        out.append((rng.normal(size=(n, 16)) + means[comp]).astype(np.float32))
    return out


def test_spatial_order_is_a_permutation_that_keeps_clusters_together():
    from flash_sd_kde.batched import spatial_order

    rng = np.random.default_rng(4)
    a = rng.normal(size=(300, 16)) + 40.0
    b = rng.normal(size=(200, 16)) - 40.0
    x = np.concatenate([a, b])[rng.permutation(500)]
    perm = spatial_order(torch.as_tensor(x, dtype=torch.float32)).numpy()
    assert np.array_equal(np.sort(perm), np.arange(500))
    side = x[perm, 0] > 0  # cluster of each point in the new order
    assert np.count_nonzero(side[1:] != side[:-1]) == 1
    for n in (1, 2):
        assert spatial_order(torch.zeros(n, 16)).tolist() == list(range(n))


@pytest.mark.gpu
def test_spatial_order_on_the_gpu(gpu):
    """The library's Morton codes (fsk_spatial_codes16, used for float32
    (n, 16) CUDA tensors): a permutation that keeps two clusters apart, and on
    points along one line the order follows the line up to one cell of the
    10-bit grid over the mean +- 4 standard deviations (either direction: the
    sign of an axis is arbitrary)."""
    from flash_sd_kde.spatial import spatial_order

    rng = np.random.default_rng(4)
    a = rng.normal(size=(3000, 16)) + 40.0
    b = rng.normal(size=(2000, 16)) - 40.0
    x = np.concatenate([a, b])[rng.permutation(5000)]
    perm = spatial_order(torch.as_tensor(x, dtype=torch.float32, device=gpu)).cpu().numpy()
    assert np.array_equal(np.sort(perm), np.arange(5000))
    side = x[perm, 0] > 0
    assert np.count_nonzero(side[1:] != side[:-1]) == 1

    u = rng.normal(size=16)
    u /= np.linalg.norm(u)
    t = rng.normal(size=20000) * 3.0
    line = (t[:, None] * u + 0.5).astype(np.float32)
    perm = spatial_order(torch.as_tensor(line, device=gpu)).cpu().numpy()
    assert np.array_equal(np.sort(perm), np.arange(20000))
    ts = t[perm] if t[perm][0] < t[perm][-1] else -t[perm]
    ts = ts[np.abs(ts - t.mean()) < 3.5 * t.std()]  # past 4 sd: the end cells
    cell = 8.0 * t.std() / 1023
    assert np.all(np.diff(ts) >= -1.5 * cell)
    with pytest.raises(ValueError):
        from flash_sd_kde import _ops
        _ops.spatial_codes16(torch.as_tensor(line, device=gpu), bits=11)


@pytest.mark.gpu
def test_batched_spatial_order_changes_only_rounding(gpu):
    """The spatial order of the points only changes the summation order and
    which negligible (< 2^-(ceil(log2 n) + 28) relative) terms are skipped: the log-likelihoods
    agree with the given order, and fewer exponentials are evaluated."""
    sets = _sets([3000, 700, 65], seed=6)
    factors = np.geomspace(0.05, 2.0, 32)
    grids = [LooBandwidthGrid(sets, factors=factors, device=gpu, autotune=False, graph=False,
                              spatial=sp) for sp in (False, True)]
    ll = [g.run().cpu().numpy() for g in grids]
    np.testing.assert_allclose(ll[1], ll[0], rtol=1e-7)
    done = [g.exps_evaluated() for g in grids]
    logical = grids[0].pair_count() * 32
    print("evaluated exps: given order %.3f, spatial order %.3f" % (done[0] / logical,
                                                                    done[1] / logical))
    assert done[1] < done[0] <= logical * 1.1


@pytest.mark.gpu
def test_batched_matches_fp64_bruteforce(gpu):
    sets = _sets([2, 3, 17, 100, 333, 64], seed=1)
    factors = np.geomspace(0.05, 2.0, 32)
    ll = loo_loglik_grid(sets, factors=factors, device=gpu)
    worst = 0.0
    for k, x in enumerate(sets):
        hs = scott_mean_std_bandwidth(x) * factors
        ref = fp64.loo_loglik(x, hs)
        err = np.max(np.abs(ll[k] - ref) / np.abs(ref))
        worst = max(worst, err)
        np.testing.assert_allclose(ll[k], ref, rtol=1e-5)
    print("batched LOO max rel err vs FP64:", worst)


@pytest.mark.gpu
def test_batched_graph_replays_and_autotune(gpu):
    sets = _sets([900, 257, 4000, 31], seed=2)
    grid = LooBandwidthGrid(sets, factors=np.geomspace(0.1, 1.5, 12), device=gpu, autotune=True,
                            graph=True)
    eager = grid.run(use_graph=False).clone()
    for _ in range(3):
        torch.ones(1000, device=gpu).sum().item()  # unrelated eager work between replays
        out = grid.run().clone()
        torch.cuda.synchronize()
        np.testing.assert_allclose(out.cpu().numpy(), eager.cpu().numpy(), rtol=1e-12)
    from flash_sd_kde.batched import CANDIDATES

    assert grid.tune in CANDIDATES
    assert set(grid.tune_times) == set(CANDIDATES)
    # every tile shape gives the same result up to the order of its FP32 and
    # FP64 additions (eager, no graph)
    for cand in CANDIDATES:
        grid._launch(cand)
        np.testing.assert_allclose(grid.out.cpu().numpy(), eager.cpu().numpy(), rtol=1e-9)
    # the chosen bandwidth is the grid maximum of the FP64 oracle's curve
    for k, x in enumerate(sets):
        ref = fp64.loo_loglik(x, grid.bandwidths[k])
        assert np.argmax(ref) == np.argmax(eager[k].cpu().numpy())


@pytest.mark.gpu
def test_batched_agrees_with_score_kernel_identity(gpu):
    """Cross-kernel check: the leave-one-out density is the score kernel's
    diagonal-inclusive sum minus the self term, LOO_i = (pdf_sum_i - 1) / (n - 1)
    times the Gaussian normaliser, so the LOO grid and the ordered score kernel
    (two different kernels) must give the same mean log-likelihood."""
    from flash_sd_kde import _ops

    x = _sets([1500], seed=5)[0]
    h0 = scott_mean_std_bandwidth(x)
    # bandwidths wide enough that every point has neighbours: pdf_sum - 1 is
    # then far above the FP32 resolution of pdf_sum (at small h the LOO grid's
    # shifted form is exact where this difference is not)
    factors = np.array([2.0, 4.0])
    ll = loo_loglik_grid([x], factors=factors, device=gpu)[0]
    n = x.shape[0]
    for f, got in zip(factors, ll):
        h = h0 * f
        pdf, _ = _ops.score16(torch.as_tensor(x, device=gpu), h, variant="ordered")
        s = pdf.double().cpu().numpy() - 1.0
        assert s.min() > 0.05
        log_norm = -np.log(n - 1) - 16 * np.log(h) - 8 * np.log(2 * np.pi)
        ref = float(np.mean(np.log(s) + log_norm))
        assert abs(got - ref) <= 1e-5 * abs(ref), (f, got, ref)


@pytest.mark.gpu
def test_batched_offset_datasets_match_fp64(gpu):
    """Datasets far from the origin (every coordinate shifted by +50, |x|^2 ~
    4e4): each dataset is centred on its own mean before the FP32 cross term,
    so the result matches FP64 as for data near 0, over the whole grid down
    to 0.05 h where the exponent coefficient is 400 times that at h."""
    sets = [x + np.float32(50.0) for x in _sets([3, 40, 333, 1000], seed=3)]
    sets[1] -= np.float32(120.0)  # another dataset on the other side of the origin
    factors = np.geomspace(0.05, 2.0, 32)
    ll = loo_loglik_grid(sets, factors=factors, device=gpu)
    for k, x in enumerate(sets):
        ref = fp64.loo_loglik(x, scott_mean_std_bandwidth(x) * factors)
        np.testing.assert_allclose(ll[k], ref, rtol=1e-5)


def test_morton_bit_spread_matches_bit_loop():
    """spatial_order's code puts bit b of axis a at bit 3b + a; the magic-mask
    spread equals moving each bit one by one."""
    from flash_sd_kde.spatial import _spread3

    v = torch.arange(1024, dtype=torch.int32)
    ref = torch.zeros_like(v)
    for b in range(10):
        ref |= ((v >> b) & 1) << (3 * b)
    assert torch.equal(_spread3(v), ref)
