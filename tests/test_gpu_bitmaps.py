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


def _corners_format(seed=3):
    """Hand-built TaiChi_new for the layout planner's corners:
    shape 0, rows [100, 2100), 40 offsets stored out of order, many zero cells
    cancelled (gets a bitmap of two words a row); shape 1, rows [2500, 4500),
    70 offsets (too many for a bitmap: its cancelled cells stay entries);
    shape 2, rows [1000, 1500) inside shape 0's rows, 4 offsets, few zero cells
    (a bitmap all the same: its rows also carry shape 0's terms, so a
    cancelled cell would not cancel exactly in FP32). Rows also hold positive entries, a sign-0 entry on no
    diagonal, and some cancelled cells listed twice (the second stays)."""
    M = N = 6000
    rng = np.random.default_rng(seed)
    offs0 = rng.permutation(np.arange(-20, 20)).astype(np.int32)
    offs1 = np.arange(-35, 35, dtype=np.int32)
    offs2 = np.array([0, 1, 2, 3], np.int32)
    xs, xe, ys = [100, 2500, 1000], [2100, 4500, 1500], [100, 2600, 3000]
    f = T.Format(M, N)
    f.xStart, f.xStop, f.yStart = (np.array(v, np.int32) for v in (xs, xe, ys))
    f.nDiasPtr = np.array([0, 40, 110, 114], np.int32)
    f.neg_offsets = np.concatenate([offs0, offs1, offs2])
    rows, absorbed, pos, neg = [], 0, 0, 0
    for r in range(M):
        ent = []
        for c in rng.integers(0, N, int(rng.integers(0, 4))):
            ent.append((int(c) << 1) | 1)
        for s, (offs, p) in enumerate(((offs0, 0.3), (offs1, 0.1), (offs2, 0.2))):
            if not (xs[s] <= r < xe[s]):
                continue
            for o in offs:
                c = r - xs[s] + ys[s] + int(o)
                if ys[s] <= c < ys[s] + (xe[s] - xs[s]) and rng.random() < p:
                    ent.append(c << 1)
                    if s != 1:
                        absorbed += 1
                    if s == 0 and rng.random() < 0.05:
                        ent.append(c << 1)  # listed twice: only the first is absorbed
        if r % 97 == 0:
            ent.append(((r + 3000) % N) << 1)  # on no diagonal
        ent = [ent[i] for i in rng.permutation(len(ent))]
        pos += sum(e & 1 for e in ent)
        neg += sum(1 - (e & 1) for e in ent)
        rows.append(np.array(ent, np.int32))
    rp = np.zeros(M + 1, np.int32)
    rp[1:] = np.cumsum([len(e) for e in rows])
    f.csrRow, f.csrCol = rp, np.concatenate(rows).astype(np.int32)
    f.dia_shapes, f.num_negs = 3, 114
    f.dense_nnz = 1  # any positive count marks the columns as encoded
    f.sparse_nnz, f.zero_elements = pos, neg
    f.row_start, f.row_stop = 0, M
    return f, absorbed


def test_hand_made_layout_corners(dev):
    f, absorbed = _corners_format()
    A = T.Matrix(f)
    st = A.stats()
    assert st["bitmap_shapes"] == 2
    assert st["holes"] == absorbed
    assert st["holes"] + st["remainder"] == int(f.csrRow[-1])
    assert T.footprint(f) == st["device_bytes"]
    rng = np.random.default_rng(5)
    xi = rng.integers(0, 8, f.N).astype(np.float32)
    want = O.taichi_spmv(f, xi, np.float32)
    for k in (1, 4, 8):
        X = torch.from_numpy(np.repeat(xi[:, None], k, 1)).to(dev)
        Y = torch.full((f.M, k), float("nan"), device=dev)
        A.spmv(X if k > 1 else X[:, 0].contiguous(), Y if k > 1 else Y[:, 0])
        torch.cuda.synchronize()
        for c in range(k):
            assert np.array_equal(Y[:, c].cpu().numpy(), want), (k, c)
    xf = rng.standard_normal(f.N)
    y = torch.empty(f.M, device=dev)
    A.spmv(torch.from_numpy(xf.astype(np.float32)).to(dev), y)
    torch.cuda.synchronize()
    x64 = xf.astype(np.float32).astype(np.float64)
    y64 = O.taichi_spmv(f, x64)
    fa = T.Format(f.M, f.N)
    for name in T.Format.FIELDS + T.Format.ARRAYS:
        setattr(fa, name, getattr(f, name))
    fa.csrCol = f.csrCol | 1  # every term with a plus sign: the bound's denominator
    den = np.maximum(np.abs(y64), O.taichi_spmv(fa, np.abs(x64)))
    den[den == 0] = 1
    assert float(np.max(np.abs(y.cpu().numpy() - y64) / den)) <= 1e-6


@pytest.mark.parametrize("n,w,extra", [(30000, 24, 0), (20000, 64, 0), (30000, 40, 5000)])
def test_band_offset_order(n, w, extra, dev):
    """A band's offsets stored in increasing order take the contiguous-offset
    loop (compile-time offsets, one bit-select a term); the same format with
    each shape's offsets stored in another order takes the general loop.
    Both give the CPU sums exactly for integer x (k = 1..8, FP32 and BF16 x)
    and stay within the FP64 bound for real x."""
    rp, col = _band_plus_random(n, w, extra, 23)
    fmt, _ = T.partition_shapes(n, n, rp, col)
    assert fmt.dia_shapes >= 1
    offs = np.asarray(fmt.neg_offsets).copy()
    ptr = np.asarray(fmt.nDiasPtr)
    assert all(np.array_equal(offs[ptr[s]:ptr[s + 1]], np.arange(offs[ptr[s]], offs[ptr[s]] + ptr[s + 1] - ptr[s]))
               for s in range(fmt.dia_shapes))  # bands: consecutive increasing offsets
    rng = np.random.default_rng(9)
    perm = offs.copy()
    for s in range(fmt.dia_shapes):
        perm[ptr[s]:ptr[s + 1]] = rng.permutation(offs[ptr[s]:ptr[s + 1]])
    fb = T.Format(fmt.M, fmt.N)
    for name in T.Format.FIELDS + T.Format.ARRAYS:
        setattr(fb, name, getattr(fmt, name))
    fb.neg_offsets = perm.astype(np.int32)
    A, B = T.Matrix(fmt, nnz=len(col)), T.Matrix(fb, nnz=len(col))
    assert A.stats()["holes"] == B.stats()["holes"]
    xi = T.gen_vector(n * 8, 1, 31).reshape(n, 8)
    want = [O.csr_spmv_f32(rp, col, np.ascontiguousarray(xi[:, c])) for c in range(8)]
    for dt in (torch.float32, torch.bfloat16):
        Xd = torch.from_numpy(xi).to(dev).to(dt)  # integers 0..7 are exact in BF16
        for k in (1, 2, 3, 4, 5, 8):
            outs = []
            for M_ in (A, B):
                Y = torch.full((n, 8), float("nan"), device=dev)
                if k == 1:
                    y = torch.empty(n, device=dev)
                    M_.spmv(Xd[:, 0].contiguous(), y)
                    torch.cuda.synchronize()
                    outs.append(y.cpu().numpy()[:, None])
                else:
                    M_.spmv(Xd[:, :k], Y[:, :k])
                    torch.cuda.synchronize()
                    outs.append(Y.cpu().numpy()[:, :k])
            for c in range(k):
                assert np.array_equal(outs[0][:, c], want[c]), (dt, k, c)
                assert np.array_equal(outs[1][:, c], want[c]), (dt, k, c)
    xr = np.random.default_rng(4).standard_normal(n).astype(np.float32)
    for M_ in (A, B):
        y = torch.empty(n, device=dev)
        M_.spmv(torch.from_numpy(xr).to(dev), y)
        torch.cuda.synchronize()
        check_float(y.cpu().numpy(), rp, col, xr.astype(np.float64))
