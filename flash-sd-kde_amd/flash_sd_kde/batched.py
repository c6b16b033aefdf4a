"""Batched leave-one-out bandwidth selection (no reference counterpart; the
caller it serves is the validation-likelihood bandwidth search of the
reference's benchmarks/mnist_fashion_pca16_ood.py:63-92, done here as an
exact LOO log-likelihood over a bandwidth grid for many datasets at once).

    grid = LooBandwidthGrid(datasets, factors=np.geomspace(0.05, 2, 32))
    ll = grid.run()            # (n_sets, n_bw) float64 torch tensor on the GPU
    best = grid.best_bandwidths()

All datasets and bandwidths are evaluated by ONE persistent kernel launch
(fsk_loo_bandwidth_grid) fed from a device work queue. Its tile shape and
launch shape -- waves per workgroup, row tiles of 16 rows per wave (the task
tile is 16 x row_tiles x waves rows against all of a dataset's points) and
workgroups per CU -- are chosen by `autotune`, which times every candidate on
the actual batch, checks that all candidates agree, and caches the winner per
batch signature. The whole sequence (queue reset, centring, persistent
kernel, finish) is captured once in a HIP graph and replayed by run().

The kernel skips, per 16-point data tile and group of four bandwidths, terms
below 2^-(ceil(log2 n) + 28) of a row's nearest-neighbour term (n the
dataset's size; together below 2^-28 of the row's sum). It skips the more the closer
together the points of a tile lie, so each dataset is first put in a spatial
order (`spatial_order`): the Morton (Z-order) code of its coordinates along
its three leading principal axes. The log-likelihoods are sums over all rows
and pairs, so the order changes them only by rounding.
"""

from __future__ import annotations

import ctypes
import threading
from typing import Sequence

import numpy as np
import torch

from flash_sd_kde import _native as nat
from flash_sd_kde.bandwidth import scott_mean_std_bandwidth
from flash_sd_kde.spatial import spatial_order

MAX_SETS = 64
MAX_BW = 32
# (waves per workgroup, row tiles per wave, workgroups per CU). Resident waves
# per SIMD are bounded by registers: 3 for one row tile (137 VGPRs), 2 for two
# (~204 VGPRs), so larger workgroup counts would only queue behind these.
CANDIDATES = [(4, 1, 2), (4, 1, 3), (8, 1, 1), (8, 1, 2), (4, 2, 1), (4, 2, 2), (8, 2, 1)]
DEFAULT_TUNE = (4, 1, 3)

_tune_cache: dict[tuple, tuple[tuple[int, int, int], dict]] = {}
_tune_lock = threading.Lock()


class LooBandwidthGrid:
    def __init__(self, datasets: Sequence, *, factors: Sequence[float] | None = None,
                 bandwidths: np.ndarray | None = None, device="cuda", autotune: bool = True,
                 graph: bool = True, tune: tuple[int, int, int] | None = None,
                 spatial: bool = True):
        self.device = torch.device(device)
        arrays = [np.asarray(d, dtype=np.float32) for d in datasets]
        if not 1 <= len(arrays) <= MAX_SETS:
            raise ValueError(f"need 1..{MAX_SETS} datasets, got {len(arrays)}")
        for k, a in enumerate(arrays):
            if a.ndim != 2 or a.shape[1] != 16:
                raise ValueError(f"dataset {k}: expected shape (n, 16), got {a.shape}")
            if a.shape[0] < 2:
                raise ValueError(f"dataset {k}: leave-one-out needs at least 2 points")
            if not np.all(np.isfinite(a)):
                raise ValueError(f"dataset {k}: points contain NaN or infinite values.")
        self.sizes = [a.shape[0] for a in arrays]
        self.offsets = np.zeros(len(arrays) + 1, dtype=np.int64)
        self.offsets[1:] = np.cumsum(self.sizes)
        if bandwidths is None:
            if factors is None:
                factors = np.geomspace(0.05, 2.0, MAX_BW)
            base = np.array([scott_mean_std_bandwidth(a) for a in arrays])
            bandwidths = base[:, None] * np.asarray(factors, dtype=np.float64)[None, :]
        bandwidths = np.asarray(bandwidths, dtype=np.float64)
        if bandwidths.shape[0] != len(arrays) or not 1 <= bandwidths.shape[1] <= MAX_BW:
            raise ValueError(f"bandwidths must be (n_sets, 1..{MAX_BW}), got {bandwidths.shape}")
        if not np.all(bandwidths > 0):
            raise ValueError("bandwidths must be positive")
        if tune is not None and tuple(tune) not in CANDIDATES:
            raise ValueError(f"tune must be one of {CANDIDATES}, got {tune}")
        self.bandwidths = bandwidths
        self.n_sets, self.n_bw = bandwidths.shape
        nat.require_gpu(self.device, "LooBandwidthGrid")
        parts = [torch.as_tensor(a, device=self.device) for a in arrays]
        if spatial:
            parts = [p[spatial_order(p)] for p in parts]
        self.x = torch.cat(parts, dim=0).contiguous()
        self.bw = torch.as_tensor(bandwidths, device=self.device).contiguous()
        self.out = torch.empty((self.n_sets, self.n_bw), dtype=torch.float64, device=self.device)
        self._off_c = self.offsets.ctypes.data_as(ctypes.c_void_p)
        nbytes = int(nat.lib().fsk_loo_workspace_bytes(self._off_c, self.n_sets, self.n_bw))
        self.ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=self.device)
        self.ws_bytes = nbytes
        self.tune = tuple(tune) if tune is not None else (
            self.autotune() if autotune else DEFAULT_TUNE)
        self.graph = None
        if graph:
            self._capture()

    # ---- launching ----
    def _launch(self, tune: tuple[int, int, int] | None = None,
                exps_done: torch.Tensor | None = None) -> None:
        w, rt, g = tune or self.tune
        st = nat.lib().fsk_loo_bandwidth_grid(
            self.x.data_ptr(), self.x.stride(0), self._off_c, self.n_sets, self.bw.data_ptr(),
            self.n_bw, w, rt, g, self.ws.data_ptr(), self.ws_bytes, self.out.data_ptr(),
            None if exps_done is None else exps_done.data_ptr(), nat.stream_handle(self.device))
        nat.check(st, "fsk_loo_bandwidth_grid")

    def signature(self) -> tuple:
        return (tuple(self.sizes), self.n_bw)

    def autotune(self, repeats: int = 2) -> tuple[int, int, int]:
        """Time every tile/launch-shape candidate on this batch; keep the
        fastest. All candidates must produce the same log-likelihoods: they
        differ in the order of FP64 atomic additions, and in which terms below
        2^-(ceil(log2 n) + 28) of a row's nearest-neighbour term their tiles
        skip."""
        key = self.signature()
        with _tune_lock:
            if key in _tune_cache:
                best, self.tune_times = _tune_cache[key]
                return best
            times, results = {}, {}
            for cand in CANDIDATES:
                self._launch(cand)  # warm-up
                start = torch.cuda.Event(enable_timing=True)
                stop = torch.cuda.Event(enable_timing=True)
                start.record()
                for _ in range(repeats):
                    self._launch(cand)
                stop.record()
                stop.synchronize()
                times[cand] = start.elapsed_time(stop) / repeats
                results[cand] = self.out.clone()
            ref = results[CANDIDATES[0]]
            for cand, r in results.items():
                if not torch.allclose(r, ref, rtol=1e-9, atol=1e-9):
                    raise RuntimeError(f"autotune candidate {cand} disagrees with {CANDIDATES[0]}")
            best = min(times, key=times.get)
            self.tune_times = times
            _tune_cache[key] = (best, times)
            return best

    def exps_evaluated(self) -> int:
        """Exponentials one launch evaluates with the chosen shape (the
        logical count is pair_count() * n_bw; the kernel skips tile groups
        whose terms are all below 2^-(ceil(log2 n) + 28) of the
        nearest-neighbour term)."""
        cnt = torch.zeros(1, dtype=torch.int64, device=self.device)
        self._launch(exps_done=cnt)
        return int(cnt.item())

    def _capture(self) -> None:
        self._launch()  # eager run before capture (library state, allocator)
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream(device=self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            with torch.cuda.graph(g, stream=side):
                self._launch()
        torch.cuda.current_stream(self.device).wait_stream(side)
        torch.cuda.synchronize(self.device)
        self.graph = g

    def run(self, *, use_graph: bool = True) -> torch.Tensor:
        """(n_sets, n_bw) mean leave-one-out log-likelihoods (float64, device)."""
        if use_graph and self.graph is not None:
            self.graph.replay()
        else:
            self._launch()
        return self.out

    def best_bandwidths(self) -> np.ndarray:
        ll = self.run().cpu().numpy()
        return self.bandwidths[np.arange(self.n_sets), np.argmax(ll, axis=1)]

    def pair_count(self) -> int:
        """Ordered pairs evaluated (phase 2, j != i) summed over datasets."""
        return int(sum(n * (n - 1) for n in self.sizes))


def loo_loglik_grid(datasets: Sequence, *, factors=None, bandwidths=None, device="cuda",
                    autotune: bool = True) -> np.ndarray:
    """One-shot helper: (n_sets, n_bw) LOO log-likelihoods as numpy."""
    grid = LooBandwidthGrid(datasets, factors=factors, bandwidths=bandwidths, device=device,
                            autotune=autotune, graph=False)
    return grid.run().cpu().numpy()
