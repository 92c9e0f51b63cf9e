"""The wide-rows tiles' near window. When most remainder entries of a matrix
(24 or more a row) lie within 64 rows of their own row, the k > 1 tiles
(k <= 8) stage the x rows around each tile in LDS once and read those
entries there; the others still gather from HBM. A sparse band (density 0.4:
below the partition's 3/5, so no dense diagonals and about 32 remainder
entries a row) turns it on. Checked: the flag, integer x exact against the
CPU sum for every k, each column equal to its k = 1 result (k = 1 never uses
the window), masked chunks (k = 3, 5, 6, 7 of an 8-column buffer) and BF16
x, long rows beside the window, and a batched launch equal to the plain
one."""
import numpy as np
import pytest
import torch

import oracle as O
import taichi_amd as T

from helpers import check_float

pytestmark = pytest.mark.gpu


def _sparse_band(n, w, far, nlong, seed):
    """Band of w diagonals at density 0.4, `far` uniform entries a row (read
    from HBM), and nlong rows of 1,500 uniform entries (long rows)."""
    rp, col = T.gen_banded(n, w, 0.4, seed)
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp))
    er = [np.repeat(np.arange(n, dtype=np.int64), far)]
    ec = [rng.integers(0, n, n * far)]
    for r in rng.integers(0, n, nlong):
        er.append(np.full(1500, r, np.int64))
        ec.append(rng.integers(0, n, 1500))
    key = np.unique(np.concatenate([rows * n + col] + [a * n + b for a, b in zip(er, ec)]))
    r2, c2 = key // n, (key % n).astype(np.int32)
    rp2 = np.zeros(n + 1, np.int64)
    np.add.at(rp2, r2 + 1, 1)
    return np.cumsum(rp2).astype(np.int32), c2


CASES = [(30011, 80, 2, 0), (20000, 72, 1, 3)]


@pytest.mark.parametrize("n,w,far,nlong", CASES)
def test_near_window_results(n, w, far, nlong, dev):
    rp, col = _sparse_band(n, w, far, nlong, n + w)
    fmt, _ = T.partition_shapes(n, n, rp, col)
    A = T.Matrix(fmt, nnz=len(col), long_threshold=1000 if nlong else 0)
    st = A.stats()
    assert st["near_window"] == 1 and st["dia_shapes"] == 0
    assert (st["long_rows"] > 0) == (nlong > 0)
    rng = np.random.default_rng(w)
    xi = rng.integers(0, 8, (n, 8)).astype(np.float32)
    X8 = torch.from_numpy(xi).to(dev)
    cols1 = []
    for c in range(8):
        y = torch.empty(n, device=dev)
        A.spmv(X8[:, c].contiguous(), y)
        cols1.append(y)
    torch.cuda.synchronize()
    for c in range(8):
        assert np.array_equal(cols1[c].cpu().numpy(), O.csr_spmv_f32(rp, col, xi[:, c])), c
    items = []
    for k in range(2, 9):
        for X, Y in ((X8[:, :k].contiguous(), torch.full((n, k), float("nan"), device=dev)),
                     (X8[:, :k], torch.full((n, 8), float("nan"), device=dev)[:, :k])):
            A.spmv(X, Y)
            torch.cuda.synchronize()
            for c in range(k):
                assert torch.equal(Y[:, c], cols1[c]), (k, c, X.stride())
            Yb = torch.full_like(Y, float("nan")) if Y.is_contiguous() else \
                torch.full((n, 8), float("nan"), device=dev)[:, :k]
            items.append((A, X, Yb, Y))
    T.spmv_batch([(a, x, yb) for a, x, yb, _ in items])
    torch.cuda.synchronize()
    for _, _, Yb, Y in items:
        assert torch.equal(Yb, Y)
    # random x: within the FP64 bound; BF16 x against the FP64 sum of its values
    g = torch.Generator(device=dev).manual_seed(n)
    Xf = torch.randn(n, 8, device=dev, generator=g)
    Yf = torch.empty(n, 8, device=dev)
    A.spmv(Xf, Yf)
    Xb = Xf.to(torch.bfloat16)
    Yb = torch.empty(n, 8, device=dev)
    A.spmv(Xb, Yb)
    torch.cuda.synchronize()
    Xh, Xq = Xf.cpu().numpy().astype(np.float64), Xb.float().cpu().numpy().astype(np.float64)
    for c in (0, 3, 7):
        check_float(Yf[:, c].cpu().numpy(), rp, col, Xh[:, c])
        check_float(Yb[:, c].cpu().numpy(), rp, col, Xq[:, c])


def test_near_window_stays_off_for_scattered_and_light_remainders(dev):
    # R-MAT (scattered) and a light band (6 diagonals at density 0.5: no
    # dense diagonals, 3 near entries a row) keep gathering from HBM: the
    # window costs more than it saves there
    for M, rp, col in ((1 << 14, *T.gen_rmat(1 << 14, 14, 16 << 14, 0.57, 0.19, 0.19, 3)),
                       (4096, *T.gen_banded(4096, 6, 0.5, 1))):
        fmt, _ = T.partition_shapes(M, M, rp, col)
        assert T.Matrix(fmt, nnz=len(col)).stats()["near_window"] == 0


def _dense_and_near(n, seed):
    """Dense diagonals -6..5 at density 0.75 (a shape whose zero cells become
    presence-bitmap bits) inside a sparse band of offsets -60..59 at density
    0.3 (about 32 remainder entries a row, all within 64 rows: the near
    window)"""
    rng = np.random.default_rng(seed)
    r = np.repeat(np.arange(n, dtype=np.int64), 120)
    o = np.tile(np.arange(-60, 60, dtype=np.int64), n)
    c = r + o
    dense = (o >= -6) & (o <= 5)
    keep = (c >= 0) & (c < n) & (rng.random(len(r)) < np.where(dense, 0.75, 0.3))
    r, c = r[keep], c[keep]
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, r + 1, 1)
    return np.cumsum(rp).astype(np.int32), c.astype(np.int32)


@pytest.mark.parametrize("grid", [0, 7])
def test_near_window_with_bitmap_shape(grid, dev):
    """A tile stages its shape's x window (after kWideZero floats of +0.0
    that absent terms read) and then the near window over the same LDS; with
    a grid of 7 workgroups each walks many tiles, so the next tile's shape
    restages the window and its zero slot after the near window. Integer x
    must give the CPU sums exactly at every k."""
    n = 40000
    rp, col = _dense_and_near(n, 17)
    fmt, _ = T.partition_shapes(n, n, rp, col)
    A = T.Matrix(fmt, nnz=len(col))
    st = A.stats()
    assert st["near_window"] == 1 and st["dia_shapes"] >= 1 and st["bitmap_shapes"] >= 1
    if grid:
        for k in range(2, 9):
            A.set_config(k, T.F32, 0, 0, grid, 0)
    xi = np.random.default_rng(3).integers(0, 8, (n, 8)).astype(np.float32)
    X8 = torch.from_numpy(xi).to(dev)
    want = [O.csr_spmv_f32(rp, col, np.ascontiguousarray(xi[:, c])) for c in range(8)]
    for k in range(2, 9):
        Y = torch.full((n, 8), float("nan"), device=dev)[:, :k]
        A.spmv(X8[:, :k], Y)
        torch.cuda.synchronize()
        Yh = Y.cpu().numpy()
        for c in range(k):
            assert np.array_equal(Yh[:, c], want[c]), (k, c)
