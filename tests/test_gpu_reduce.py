"""GPU parity of the public split-K combine (kernels/reduce_partials.py, the
reference's reduce_pdf_partials / reduce_weighted_partials at
kernels/reduce_partials.py:61-108): per-split partial sums (S, m) and
(S, m, 16), contiguous and as strided views, against an FP64 numpy sum of the
same FP32 inputs. The combine adds in FP64 and rounds once, so every output
is within one FP32 ulp of the FP64 total (the reference adds in FP32)."""

import numpy as np
import pytest
import torch

from kernels.reduce_partials import reduce_pdf_partials, reduce_weighted_partials

pytestmark = pytest.mark.gpu


def _within_one_ulp(got, parts64_sum):
    ref = parts64_sum.astype(np.float32)
    ulp = np.spacing(np.abs(ref)).astype(np.float64)
    err = np.abs(got.astype(np.float64) - parts64_sum)
    assert np.all(err <= ulp + 1e-45), float(np.max(err / np.maximum(ulp, 1e-45)))


@pytest.mark.parametrize("S,m", [(1, 1), (3, 5), (17, 1000), (64, 4097)])
def test_reduce_pdf_partials_matches_fp64_sum(gpu, S, m):
    rng = np.random.default_rng(S * 1000 + m)
    p = (rng.standard_normal((S, m)) * 10.0 ** rng.integers(-3, 4, size=(S, m))).astype(np.float32)
    got = reduce_pdf_partials(torch.as_tensor(p, device=gpu)).cpu().numpy()
    assert got.shape == (m,) and got.dtype == np.float32
    _within_one_ulp(got, p.astype(np.float64).sum(0))


def test_reduce_pdf_partials_strided_views(gpu):
    rng = np.random.default_rng(7)
    big = torch.as_tensor(rng.standard_normal((9, 1300)).astype(np.float32), device=gpu)
    view = big[:, 100:1100]                    # row stride 1300, read in place
    got = reduce_pdf_partials(view).cpu().numpy()
    _within_one_ulp(got, view.cpu().numpy().astype(np.float64).sum(0))
    tview = big[:, :700].t().contiguous().t()  # column-major storage, copied first
    assert tview.stride(1) != 1
    got = reduce_pdf_partials(tview).cpu().numpy()
    _within_one_ulp(got, tview.cpu().numpy().astype(np.float64).sum(0))


def test_reduce_partials_add_in_fp64(gpu):
    """1e8 + 1 - 1e8 is 1 in FP64 and 0 in FP32 left to right."""
    p = torch.tensor([[1e8, 3.0], [1.0, 1e8], [-1e8, -1e8]], device=gpu)
    assert reduce_pdf_partials(p).cpu().tolist() == [1.0, 3.0]
    w = p[:, :, None].expand(3, 2, 16).contiguous()
    assert reduce_weighted_partials(w).cpu().numpy().tolist() == [[1.0] * 16, [3.0] * 16]


@pytest.mark.parametrize("S,m", [(1, 3), (4, 257), (33, 2048)])
def test_reduce_weighted_partials_matches_fp64_sum(gpu, S, m):
    rng = np.random.default_rng(S + 31 * m)
    p = rng.standard_normal((S, m, 16)).astype(np.float32)
    got = reduce_weighted_partials(torch.as_tensor(p, device=gpu)).cpu().numpy()
    assert got.shape == (m, 16)
    _within_one_ulp(got, p.astype(np.float64).sum(0))
    # (S, m, 16) view of wider per-split slabs: split stride (m + 5) * 16
    big = torch.as_tensor(rng.standard_normal((S, m + 5, 16)).astype(np.float32), device=gpu)
    view = big[:, 2:m + 2]
    got = reduce_weighted_partials(view).cpu().numpy()
    _within_one_ulp(got, view.cpu().numpy().astype(np.float64).sum(0))


def test_reduce_partials_reject_bad_shapes(gpu):
    with pytest.raises(ValueError):
        reduce_pdf_partials(torch.zeros(3, device=gpu))
    with pytest.raises(ValueError):
        reduce_weighted_partials(torch.zeros(2, 3, 8, device=gpu))
    with pytest.raises(ValueError):
        reduce_pdf_partials(torch.zeros(2, 3))  # CPU tensor: no CPU compute path
