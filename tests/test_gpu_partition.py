"""Device partition build (tc_partition_device): the TaiChi_new it returns is
identical, field by field and array by array, to the oracle's serial
partition_shapes (TaiChi.cu:313-638) on matrices with one shape, several
shapes sharing rows, wide bands, no shape at all, and duplicate entries."""
import numpy as np
import pytest
import torch

import oracle as O
import taichi_amd as T
from taichi_amd import workloads as W

from helpers import BAND_CASES, band_csr, band_format_equals, band_goldens, rmat

pytestmark = pytest.mark.gpu

ARRAYS = ("xStart", "yStart", "xStop", "nDiasPtr", "neg_offsets", "csrRow", "csrCol")


def _same_as_oracle(fmt, M, N, rp, col):
    of, _ = O.partition(M, N, rp, col)
    for f in O.Format.FIELDS:
        assert getattr(fmt, f) == getattr(of, f), f
    for f in ARRAYS:
        assert np.array_equal(getattr(fmt, f), getattr(of, f)), f
    return of


def _csr(M, rows, cols):
    """CSR of the (row, col) pairs in the given order within each row."""
    rows = np.asarray(rows, np.int64)
    cols = np.asarray(cols, np.int64)
    order = np.argsort(rows, kind="stable")
    rp = np.zeros(M + 1, np.int32)
    rp[1:] = np.cumsum(np.bincount(rows, minlength=M))
    return rp, cols[order].astype(np.int32)


def _three_bands(n=60000, seed=3):
    """Main band (|c - r| <= 4, density 0.9) plus bands at offsets +20000
    (rows [0, 40000)) and -20000 (rows [20000, 60000)). A contour's column is
    used as its row start (TaiChi.cu:368-369), so each side band's contour
    yields the square of the other: three shapes, rows overlapping."""
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for o in range(-4, 5):
        r = np.arange(max(0, -o), min(n, n - o))
        keep = rng.random(len(r)) < 0.9
        rows.append(r[keep])
        cols.append(r[keep] + o)
    for base, lo, hi in ((20000, 0, 40000), (-20000, 20000, 60000)):
        for o in range(-3, 4):
            r = np.arange(max(lo, -(base + o)), min(hi, n - (base + o)))
            keep = rng.random(len(r)) < 0.85
            rows.append(r[keep])
            cols.append(r[keep] + base + o)
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    # CSR order within a row: by column
    order = np.lexsort((cols, rows))
    return _csr(n, rows[order], cols[order])


def _with_duplicates(n, rp, col, on_diagonal, every=7):
    """Copy of the CSR with entry (r, c) repeated for rows r % every == 0:
    c = r (a dense diagonal) or the row's last column (off the band)."""
    rows = np.repeat(np.arange(n), np.diff(rp))
    extra_r, extra_c = [], []
    for r in range(0, n, every):
        seg = col[rp[r]:rp[r + 1]]
        if on_diagonal and r in seg:
            extra_r.append(r)
            extra_c.append(r)
        elif not on_diagonal and len(seg):
            extra_r.append(r)
            extra_c.append(int(seg.max()) + 1000 if int(seg.max()) + 1000 < n else int(seg.min()) - 1000)
    rows = np.concatenate([rows, np.array(extra_r, np.int64)])
    cols = np.concatenate([col.astype(np.int64), np.array(extra_c, np.int64)])
    if not on_diagonal:  # far column twice: the second copy is the duplicate
        rows = np.concatenate([rows, np.array(extra_r, np.int64)])
        cols = np.concatenate([cols, np.array(extra_c, np.int64)])
    order = np.lexsort((cols, rows))
    return _csr(n, rows[order], cols[order])


def _device(M, N, rp, col, dev):
    rpd = torch.from_numpy(rp).to(dev)
    cold = torch.from_numpy(col if len(col) else np.zeros(1, np.int32)).to(dev)
    return T.partition_device(M, N, rp, col, rpd, cold)


def _cases():
    out = [("cfg1", *W.diag5_random())]
    for s in (14, 16, 18):
        out.append((f"rmat{s}", *rmat(s, seed=40 + s)))
    for n, w, dens, seed in ((60000, 17, 0.7, 9), (150000, 24, 0.7, 5), (20000, 150, 0.7, 6)):
        rp, col = T.gen_banded(n, w, dens, seed)
        out.append((f"banded{n}_w{w}", n, n, rp, col))
    rp, col = T.gen_rmat(20000, 15, 200000, 0.25, 0.25, 0.25, 4)   # uniform: no shape
    out.append(("uniform", 20000, 20000, rp, col))
    rp, col = _three_bands()
    out.append(("three_bands", 60000, 60000, rp, col))
    rp, col = T.gen_banded(70000, 9, 0.8, 11)   # tall: N < M, columns clipped
    keep_rp, keep_col = [0], []
    for r in range(70000):
        seg = col[rp[r]:rp[r + 1]]
        seg = seg[seg < 50000]
        keep_col.extend(seg.tolist())
        keep_rp.append(len(keep_col))
    out.append(("tall", 70000, 50000, np.array(keep_rp, np.int32), np.array(keep_col, np.int32)))
    return out


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_device_build_equals_oracle(case, dev):
    name, M, N, rp, col = case
    fmt = _device(M, N, rp, col, dev)
    of = _same_as_oracle(fmt, M, N, rp, col)
    if name == "three_bands":
        assert of.dia_shapes >= 2
    if name == "uniform":
        assert of.dense_nnz == 0  # the alias path: the input CSR is the format
    # partition_shapes takes the same device build
    ps, _ = T.partition_shapes(M, N, rp, col)
    for f in ARRAYS:
        assert np.array_equal(getattr(ps, f), getattr(fmt, f)), f


def test_duplicates_on_a_dense_diagonal_use_the_host_build(dev):
    n = 60000
    rp, col = T.gen_banded(n, 17, 0.7, 9)
    rp2, col2 = _with_duplicates(n, rp, col, on_diagonal=True)
    assert len(col2) > len(col)
    with pytest.raises(T.TaiChiError) as e:
        _device(n, n, rp2, col2, dev)
    assert e.value.status == -104
    ps, _ = T.partition_shapes(n, n, rp2, col2)      # host build instead
    _same_as_oracle(ps, n, n, rp2, col2)


def test_duplicates_off_the_dense_diagonals_stay_on_the_device(dev):
    n = 60000
    rp, col = T.gen_banded(n, 17, 0.7, 9)
    rp2, col2 = _with_duplicates(n, rp, col, on_diagonal=False)
    assert len(col2) > len(col)
    fmt = _device(n, n, rp2, col2, dev)
    _same_as_oracle(fmt, n, n, rp2, col2)


# ---- the hand-derived band goldens (tests/golden/make_golden.py, BANDS_NOTE) ----


@pytest.mark.parametrize("name", BAND_CASES)
def test_device_build_equals_band_golden(name, dev):
    """The contour-corner transposition pinned on the device build: one band
    at +-2048 gives no shape; the pair gives two shapes, each from the other
    band's contour."""
    g = band_goldens()[name]
    M, N = g["M"], g["N"]
    rp, col = band_csr(M, g["bands"])
    band_format_equals(_device(M, N, rp, col, dev), g["format"], rp, col)
    ps, _ = T.partition_shapes(M, N, rp, col)
    band_format_equals(ps, g["format"], rp, col)


def _spmv_parity(M, N, rp, col, fmt, dev):
    """integer x bit-exact against the CPU FP32 CSR sums at k = 1..8 (every
    column) and through the batched launch; U[0,1) x within the FP64 bound"""
    A = T.Matrix(fmt, nnz=len(col))
    rng = np.random.default_rng(8)
    for k in range(1, 9):
        X = rng.integers(0, 8, (N, k)).astype(np.float32)
        Xd = torch.from_numpy(X).to(dev)
        Y = torch.full((M, k), float("nan"), device=dev)
        if k == 1:
            A.spmv(Xd[:, 0].contiguous(), Y[:, 0])
        else:
            A.spmv(Xd, Y)
        torch.cuda.synchronize()
        Yh = Y.cpu().numpy()
        for c in range(k):
            assert np.array_equal(Yh[:, c], O.csr_spmv_f32(rp, col, np.ascontiguousarray(X[:, c]))), (k, c)
        Yb = torch.full((M, k), float("nan"), device=dev)
        T.spmv_batch([(A, Xd if k > 1 else Xd[:, 0].contiguous(), Yb if k > 1 else Yb[:, 0])])
        torch.cuda.synchronize()
        assert torch.equal(Yb, Y), k
    x = rng.random(N).astype(np.float32)
    y = torch.empty(M, device=dev)
    A.spmv(torch.from_numpy(x).to(dev), y)
    torch.cuda.synchronize()
    y64 = O.csr_spmv_f64(rp, col, x.astype(np.float64))
    assert O.rel_err(y.cpu().numpy(), y64, O.csr_absrow_f64(rp, col, x.astype(np.float64))) <= 1e-6
    return A.stats()


def test_spmv_on_band_pair_golden(dev):
    """Two shapes of one dense offset each (the reference's hybrid2 class,
    TaiChi.cu:1044-1129 dispatched at :1288-1290), rows covered by both
    shapes, zero cells on one of them"""
    g = band_goldens()["band_pair_2048"]
    M, N = g["M"], g["N"]
    rp, col = band_csr(M, g["bands"])
    fmt, _ = T.partition_shapes(M, N, rp, col)
    assert (fmt.dia_shapes, fmt.num_negs) == (2, 2)
    st = _spmv_parity(M, N, rp, col, fmt, dev)
    # shape 0's zero cells lie in rows shape 1 also covers: the layout turns
    # them into bitmap bits (an FP32 cancellation there would be inexact)
    assert st["bitmap_shapes"] >= 1 and st["holes"] == 875


def test_spmv_on_three_bands(dev):
    """three shapes sharing rows (S = 3), each with several dense offsets"""
    rp, col = _three_bands()
    n = 60000
    fmt, _ = T.partition_shapes(n, n, rp, col)
    assert fmt.dia_shapes == 3
    _spmv_parity(n, n, rp, col, fmt, dev)
