"""Edge cases of the SpMV path through the C ABI: empty matrices (no rows,
no columns, no entries), non-square shapes, a row of a million entries and
the long-row threshold around rows of exactly its length, and columns at the
top of the encoded range (N = 2^30 - 1, col<<1|sign still an int32) in a
format with overlapping shapes. Integer x must match the CPU sums exactly;
random x within the FP64 bound."""
import numpy as np
import pytest
import torch

import oracle as O
import taichi_amd as T

from helpers import check_float

pytestmark = pytest.mark.gpu


def _csr(M, rows):
    rp = np.zeros(M + 1, np.int32)
    rp[1:] = np.cumsum([len(c) for c in rows])
    col = np.concatenate(rows).astype(np.int32) if rows and rp[-1] else np.zeros(0, np.int32)
    return rp, col


@pytest.mark.parametrize("M,N,nnz", [(0, 0, 0), (0, 100, 0), (100, 0, 0), (600, 600, 0), (5000, 3000, 0),
                                     (5000, 3000, 3)])
def test_empty_and_nearly_empty(M, N, nnz, dev):
    rows = [np.zeros(0, np.int32) for _ in range(M)]
    for i in range(nnz):  # a few isolated entries
        rows[(i * 1237) % M] = np.array([(i * 911) % N], np.int32)
    rp, col = _csr(M, rows)
    fmt, _ = T.partition_shapes(M, N, rp, col)
    A = T.Matrix(fmt)
    for k in (1, 4):
        for dt in (torch.float32, torch.bfloat16):
            X = torch.full((N, k), 3.0, device=dev, dtype=dt)
            Y = torch.full((M, k), float("nan"), device=dev)
            A.spmv(X if k > 1 else X[:, 0].contiguous(), Y if k > 1 else Y[:, 0])
            torch.cuda.synchronize()
            want = np.repeat(O.csr_spmv_f32(rp, col, np.full(N, 3.0, np.float32))[:, None], k, 1) if M else \
                np.zeros((0, k), np.float32)
            got = Y.cpu().numpy() if k > 1 else Y[:, :1].cpu().numpy()
            if k == 1:
                want = want[:, :1]
            assert np.array_equal(got, want), (k, dt)


@pytest.mark.parametrize("k", [3, 5, 7])
def test_single_row_x_with_masked_widths(k, dev):
    """N = 1 (X one row): 3, 5 or 7 columns, which a masked pass would take,
    run as unmasked chunks, alone and batched; columns past k stay
    untouched, and NaN past column k of X is never read into a result."""
    M, N = 3000, 1
    rows = [np.array([0], np.int32) if r % 3 else np.zeros(0, np.int32) for r in range(M)]
    rp, col = _csr(M, rows)
    fmt, _ = T.partition_shapes(M, N, rp, col)
    A = T.Matrix(fmt)
    want = (np.arange(M) % 3 != 0).astype(np.float32)[:, None] * np.arange(1, k + 1, dtype=np.float32)[None, :]
    for dt in (torch.float32, torch.bfloat16):
        buf = torch.full((1, 8), float("nan"), device=dev, dtype=dt)
        buf[0, :k] = torch.arange(1, k + 1, device=dev, dtype=torch.float32).to(dt)
        for batched in (False, True):
            Y = torch.full((M, 8), float("nan"), device=dev)
            if batched:
                T.spmv_batch([(A, buf[:, :k], Y[:, :k])])
            else:
                A.spmv(buf[:, :k], Y[:, :k])
            torch.cuda.synchronize()
            assert np.array_equal(Y[:, :k].cpu().numpy(), want), (dt, batched)
            assert torch.isnan(Y[:, k:]).all(), (dt, batched)


@pytest.mark.parametrize("shape", ["tall", "wide"])
def test_non_square(shape, dev):
    M, N = (30000, 1000) if shape == "tall" else (1000, 300000)
    rng = np.random.default_rng(11)
    rows = []
    for r in range(M):
        c = set(rng.integers(0, N, 3).tolist())
        for o in range(-2, 3):  # a band along r * N / M
            cc = r * N // M + o
            if 0 <= cc < N and rng.random() < 0.9:
                c.add(cc)
        rows.append(np.array(sorted(c), np.int32))
    rp, col = _csr(M, rows)
    fmt, _ = T.partition_shapes(M, N, rp, col)
    A = T.Matrix(fmt)
    for kind, seed in ((1, 3), (2, 4)):
        x = T.gen_vector(N, kind, seed)
        xd = torch.from_numpy(x).to(dev)
        yd = torch.full((M,), float("nan"), device=dev)
        A.spmv(xd, yd)
        torch.cuda.synchronize()
        y = yd.cpu().numpy()
        if kind == 1:
            assert np.array_equal(y, O.csr_spmv_f32(rp, col, x))
        else:
            check_float(y, rp, col, x.astype(np.float64))


def test_million_entry_row_and_long_threshold(dev):
    """Row 12345 holds 10^6 entries. The threshold is also set to the exact
    length of a remainder row and one below it, so that row is short in one
    matrix and long (a wavefront) in the other."""
    n = 1 << 21
    rng = np.random.default_rng(5)
    rp_b, col_b = T.gen_banded(n, 5, 0.9, 2)
    rows = [col_b[rp_b[r]:rp_b[r + 1]] for r in range(n)]
    rows[12345] = np.unique(np.concatenate([rows[12345], rng.choice(n, 1_000_000, replace=False)])).astype(np.int32)
    for r in range(1000, 1100):  # rows of 60..70 remainder entries around the default threshold
        extra = rng.choice(n, 60 + (r % 11), replace=False)
        rows[r] = np.unique(np.concatenate([rows[r], extra])).astype(np.int32)
    rp, col = _csr(n, rows)
    fmt, _ = T.partition_shapes(n, n, rp, col)
    lens = np.diff(fmt.csrRow)
    probe = int(lens[1000:1100].max())
    xi = T.gen_vector(n, 1, 6)
    xr = T.gen_vector(n, 2, 7)
    want_i = O.csr_spmv_f32(rp, col, xi)
    for lt in (0, probe - 1, probe, 1, 1023):
        A = T.Matrix(fmt, long_threshold=lt)
        assert A.stats()["long_rows"] >= 1
        for x, exact in ((xi, True), (xr, False)):
            xd = torch.from_numpy(x).to(dev)
            yd = torch.full((n,), float("nan"), device=dev)
            A.spmv(xd, yd)
            torch.cuda.synchronize()
            y = yd.cpu().numpy()
            if exact:
                assert np.array_equal(y, want_i), lt
            else:
                check_float(y, rp, col, x.astype(np.float64))
        A.close()


def _top_columns_format(seed=0):
    """Hand-built TaiChi_new with N = 2^30 - 1: shape 0 is rows [0, 2048)
    against the last 2048 columns (offsets -1, 0, 1), shape 1 is rows
    [1024, 4096) at column 5 (offsets 0, 2), so rows 1024..2047 sit in both;
    remainder entries of both signs, many within 5000 of the top column."""
    M, N = 4096, (1 << 30) - 1
    rng = np.random.default_rng(seed)
    f = T.Format(M, N)
    f.xStart = np.array([0, 1024], np.int32)
    f.xStop = np.array([2048, 4096], np.int32)
    f.yStart = np.array([N - 2048, 5], np.int32)
    f.nDiasPtr = np.array([0, 3, 5], np.int32)
    f.neg_offsets = np.array([-1, 0, 1, 0, 2], np.int32)
    rows, pos, neg = [], 0, 0
    for r in range(M):
        k = int(rng.integers(0, 7))
        c = np.where(rng.random(k) < 0.5, N - 1 - rng.integers(0, 5000, k), rng.integers(0, N, k))
        s = rng.integers(0, 2, k)
        pos += int(s.sum())
        neg += int(k - s.sum())
        rows.append(((c.astype(np.int64) << 1) | s).astype(np.int32))
    f.csrRow, f.csrCol = _csr(M, rows)
    f.dia_shapes, f.num_negs = 2, 5
    f.dense_nnz = 3 * 2048 + 2 * 3072  # any positive count marks the columns as encoded
    f.sparse_nnz, f.zero_elements = pos, neg
    f.row_start, f.row_stop = 0, M
    return f


def _used_columns(f):
    cols = [f.csrCol >> 1, np.arange(f.yStart[0], f.yStart[0] + 2048),
            np.arange(f.yStart[1], f.yStart[1] + 3072 + 2)]
    return np.unique(np.concatenate(cols))


def test_columns_at_the_top_of_the_encoded_range(dev):
    f = _top_columns_format()
    N = f.N
    used = _used_columns(f)
    rng = np.random.default_rng(1)
    A = T.Matrix(f)
    xd = torch.zeros(N, device=dev)
    for exact in (True, False):
        vals = (rng.integers(0, 8, len(used)) if exact else rng.standard_normal(len(used))).astype(np.float32)
        x32 = np.zeros(N, np.float32)  # pages not written stay unallocated
        x32[used] = vals
        xd[torch.from_numpy(used).to(dev)] = torch.from_numpy(vals).to(dev)
        yd = torch.full((f.M,), float("nan"), device=dev)
        A.spmv(xd, yd)
        torch.cuda.synchronize()
        y = yd.cpu().numpy()
        if exact:
            assert np.array_equal(y, O.taichi_spmv(f, x32, np.float32))
        else:
            x64 = np.zeros(N, np.float64)
            x64[used] = vals
            y64 = O.taichi_spmv(f, x64)
            fa = T.Format(f.M, N)
            for k in T.Format.FIELDS + T.Format.ARRAYS:
                setattr(fa, k, getattr(f, k))
            fa.csrCol = f.csrCol | 1  # every term with a plus sign
            x64[used] = np.abs(vals)
            absrow = O.taichi_spmv(fa, x64)
            den = np.maximum(np.abs(y64), absrow)
            den[den == 0] = 1
            assert float(np.max(np.abs(y - y64) / den)) <= 1e-6
