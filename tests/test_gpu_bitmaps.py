"""Term-presence bitmaps of the device layout. A shape whose dense diagonals
have many zero cells keeps those cells as cleared bits of a per-row bitmap,
not as sign-0 remainder entries (the reference's compensation, TaiChi.cu:
500-520), so the kernels add only the present terms. Checked on banded
matrices of density 0.7 (bands of 3 to 64 diagonals: one and two bitmap
words a row), alone and with random entries beside the band: the layout's
counts, integer x exact against the CPU sum, random x within the FP64 bound
for k = 1..8 and BF16 x, every k = 1 tile configuration, and batched results
equal to the plain call."""
import numpy as np
import pytest
import torch

import oracle as O
import taichi_amd as T

from helpers import check_float

pytestmark = pytest.mark.gpu


def _band_plus_random(n, w, extra, seed):
    rp, col = T.gen_banded(n, w, 0.7, seed)
    if extra == 0:
        return rp, col
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp))
    er = rng.integers(0, n, extra)
    ec = rng.integers(0, n, extra)
    key = np.unique(np.concatenate([rows * n + col, er * n + ec]))
    r2, c2 = key // n, (key % n).astype(np.int32)
    rp2 = np.zeros(n + 1, np.int32)
    np.add.at(rp2, r2 + 1, 1)
    return np.cumsum(rp2).astype(np.int32), c2


CASES = [(30000, 3, 0), (40000, 24, 0), (30000, 40, 5000), (20000, 64, 0), (25000, 33, 20000)]


@pytest.mark.parametrize("n,w,extra", CASES)
def test_layout_counts(n, w, extra, dev):
    rp, col = _band_plus_random(n, w, extra, 7)
    fmt, _ = T.partition_shapes(n, n, rp, col)
    A = T.Matrix(fmt, nnz=len(col))
    st = A.stats()
    R = int(fmt.csrRow[-1])
    zeros = int(np.count_nonzero((np.asarray(fmt.csrCol[:R]) & 1) == 0))
    assert st["holes"] + st["remainder"] == R
    assert st["holes"] <= zeros
    if w >= 8:  # wide bands of density 0.7: most sign-0 entries become bits
        assert st["bitmap_shapes"] >= 1 and st["holes"] > zeros // 2
    # the footprint the server admits by is the allocation
    assert T.footprint(fmt) == st["device_bytes"]


@pytest.mark.parametrize("n,w,extra", CASES)
def test_integer_exact_and_float(n, w, extra, dev):
    rp, col = _band_plus_random(n, w, extra, 11)
    fmt, _ = T.partition_shapes(n, n, rp, col)
    A = T.Matrix(fmt, nnz=len(col))
    rng = np.random.default_rng(n + w)
    xi = rng.integers(0, 8, n).astype(np.float32)
    want = O.csr_spmv_f32(rp, col, xi)
    y = torch.full((n,), float("nan"), device=dev)
    for rows in (256, 512, 1024):
        for flags in (0, 2, 8):
            A.set_config(1, T.F32, rows, 4096, 0, flags)
            A.spmv(torch.from_numpy(xi).to(dev), y)
            torch.cuda.synchronize()
            assert np.array_equal(y.cpu().numpy(), want), (rows, flags)
    xf = rng.random(n).astype(np.float32)
    A.spmv(torch.from_numpy(xf).to(dev), y)
    torch.cuda.synchronize()
    check_float(y.cpu().numpy(), rp, col, xf.astype(np.float64))


@pytest.mark.parametrize("n,w,extra", CASES[1:4])
def test_multi_rhs_and_bf16(n, w, extra, dev):
    rp, col = _band_plus_random(n, w, extra, 13)
    fmt, _ = T.partition_shapes(n, n, rp, col)
    A = T.Matrix(fmt, nnz=len(col))
    g = torch.Generator(device=dev).manual_seed(w)
    X8 = torch.randn(n, 8, device=dev, generator=g)
    for k in (2, 3, 4, 5, 7, 8):
        X = X8[:, :k].contiguous()
        Y = torch.full((n, k), float("nan"), device=dev)
        A.spmv(X, Y)
        torch.cuda.synchronize()
        Xh, Yh = X.cpu().numpy().astype(np.float64), Y.cpu().numpy()
        for c in range(k):
            check_float(Yh[:, c], rp, col, Xh[:, c])
        # each column equals its k = 1 result
        for c in range(k):
            y1 = torch.empty(n, device=dev)
            A.spmv(X[:, c].contiguous(), y1)
            torch.cuda.synchronize()
            assert torch.equal(Y[:, c], y1), (k, c)
    Xb = X8[:, :4].to(torch.bfloat16)
    Yb = torch.empty(n, 4, device=dev)
    A.spmv(Xb, Yb)
    torch.cuda.synchronize()
    Xq = Xb.float().cpu().numpy().astype(np.float64)
    for c in range(4):
        check_float(Yb[:, c].cpu().numpy(), rp, col, Xq[:, c])


def test_batched_equals_alone(dev):
    mats, items, alone = [], [], []
    for i, (n, w, extra) in enumerate(CASES):
        rp, col = _band_plus_random(n, w, extra, 17 + i)
        fmt, _ = T.partition_shapes(n, n, rp, col)
        A = T.Matrix(fmt, nnz=len(col))
        mats.append(A)
        for k in (1, 3, 8):
            X = torch.rand(n, k, device=dev) if k > 1 else torch.rand(n, device=dev)
            Ya = torch.empty(n, k, device=dev) if k > 1 else torch.empty(n, device=dev)
            A.spmv(X, Ya)
            Yb = torch.full_like(Ya, float("nan"))
            items.append((A, X, Yb))
            alone.append(Ya)
    T.spmv_batch(items)
    torch.cuda.synchronize()
    for (_, _, Yb), Ya in zip(items, alone):
        assert torch.equal(Yb, Ya)
