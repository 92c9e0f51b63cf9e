"""GPU parity of the hot path: thumbnail kernel (a) and the fused SpMV kernels
(b)+(c) against the oracle, through the C ABI."""
import numpy as np
import pytest
import torch

import oracle as O
import taichi_amd as T
from helpers import cfg1, check_float, host_format, rmat, run_spmv

pytestmark = pytest.mark.gpu


def _thumb_gpu(M, N, rp, col, dev):
    rc, H, W, _, _ = T.thumb_dims(M, N)
    assert rc == 0
    rpd = torch.from_numpy(rp).to(dev)
    cd = torch.from_numpy(col if len(col) else np.zeros(1, np.int32)).to(dev)
    img = torch.empty(H * W, dtype=torch.uint8, device=dev)
    T.thumbnail_gpu(M, N, rpd, cd, img)
    torch.cuda.synchronize()
    return img.cpu().numpy().reshape(H, W)


@pytest.mark.parametrize("case", ["cfg1", "rmat16", "rmat20", "wide", "odd"])
def test_thumbnail_bit_exact(case, dev):
    if case == "cfg1":
        M, N, rp, col = cfg1()
    elif case == "rmat16":
        M, N, rp, col = rmat(16)
    elif case == "rmat20":
        M, N, rp, col = rmat(20)
    elif case == "wide":  # N > M: W > 512
        M, N = 3000, 9000
        rng = np.random.default_rng(3)
        lens = rng.integers(0, 12, M)
        rp = np.zeros(M + 1, np.int32)
        rp[1:] = np.cumsum(lens)
        col = rng.integers(0, N, int(rp[-1])).astype(np.int32)
    else:  # non-power-of-two M (W may round to 511), rows past 512*pM dropped
        M, N, rp, col = rmat(14, n=13001)
    img = _thumb_gpu(M, N, rp, col, dev)
    ref = O.thumbnail(M, N, rp, col)
    assert np.array_equal(img, ref)


def test_partition_shapes_gpu_matches_oracle(dev):
    M, N, rp, col = rmat(18)
    fmt, _ = T.partition_shapes(M, N, rp, col)
    of, _ = O.partition(M, N, rp, col)
    for f in O.Format.FIELDS:
        assert getattr(fmt, f) == getattr(of, f), f
    for f in ("xStart", "yStart", "xStop", "nDiasPtr", "neg_offsets", "csrRow", "csrCol"):
        assert np.array_equal(getattr(fmt, f), getattr(of, f)), f


@pytest.mark.parametrize("case", ["cfg1", "rmat18", "banded", "stencil", "small"])
def test_spmv_integer_bit_exact(case, dev):
    if case == "cfg1":
        M, N, rp, col = cfg1()
    elif case == "rmat18":
        M, N, rp, col = rmat(18)
    elif case == "banded":
        M = N = 150000
        rp, col = T.gen_banded(M, 24, 0.7, 9)
    elif case == "stencil":
        M = N = 300 * 300
        rp, col = T.gen_stencil5(300, 300, 0.02, 2)
    else:  # M <= 512: no thumbnail, plain CSR path
        M = N = 500
        rp, col = T.gen_rmat(500, 9, 4000, 0.25, 0.25, 0.25, 4)
    fmt = host_format(M, N, rp, col)
    x = T.gen_vector(N, 1, 3)
    y, _ = run_spmv(fmt, x, dev)
    assert np.array_equal(y, O.csr_spmv_f32(rp, col, x))


@pytest.mark.parametrize("case", ["cfg1", "rmat18", "banded"])
def test_spmv_random_fp64(case, dev):
    if case == "cfg1":
        M, N, rp, col = cfg1()
    elif case == "rmat18":
        M, N, rp, col = rmat(18)
    else:
        M = N = 150000
        rp, col = T.gen_banded(M, 24, 0.7, 9)
    fmt = host_format(M, N, rp, col)
    for kind in (0, 2):
        x = T.gen_vector(N, kind, 5)
        y, _ = run_spmv(fmt, x, dev)
        check_float(y, rp, col, x.astype(np.float64))


def test_long_rows_wavefront(dev):
    """Rows far above the long-row threshold take the wavefront path."""
    n = 4096
    rng = np.random.default_rng(0)
    rows = []
    for r in range(n):
        if r % 97 == 0:
            cols = np.unique(rng.integers(0, n, 3000))
        else:
            cols = np.unique(rng.integers(0, n, rng.integers(0, 9)))
        rows.append(np.sort(cols))
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(c) for c in rows])
    col = np.concatenate(rows).astype(np.int32)
    fmt = host_format(n, n, rp, col)
    x = T.gen_vector(n, 1, 8)
    y, A = run_spmv(fmt, x, dev)
    assert A.stats()["long_rows"] > 0
    assert np.array_equal(y, O.csr_spmv_f32(rp, col, x))
    xr = T.gen_vector(n, 2, 9)
    y, _ = run_spmv(fmt, xr, dev)
    check_float(y, rp, col, xr.astype(np.float64))


def _offdiag_band(n, off, extra, seed):
    """One full off-diagonal (row r, column r + off) plus `extra` uniform
    entries per row: a single shape whose x window is not 4-aligned."""
    rng = np.random.default_rng(seed)
    rows = []
    for r in range(n):
        c = set(rng.integers(0, n, extra).tolist())
        if 0 <= r + off < n:
            c.add(r + off)
        rows.append(np.array(sorted(c), np.int32))
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(c) for c in rows])
    return rp, np.concatenate(rows).astype(np.int32)


def _local_mix(n, extra, spread, long_every, long_len, long_spread, seed):
    """Main diagonal, `extra` entries per row within +-spread of it, and every
    long_every-th row with long_len entries within +-long_spread (long rows)."""
    rng = np.random.default_rng(seed)
    rows = []
    for r in range(n):
        if long_every and r % long_every == 0:
            m, w = long_len, long_spread
        else:
            m, w = extra, spread
        c = set(np.clip(r + rng.integers(-w, w + 1, m), 0, n - 1).tolist())
        c.add(r)
        rows.append(np.array(sorted(c), np.int32))
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(c) for c in rows])
    return rp, np.concatenate(rows).astype(np.int32)


@pytest.mark.parametrize("case", ["rmat17", "offdiag", "long", "plain_csr", "wide_groups"])
def test_k1_kernels_agree(case, dev):
    """Every k = 1 kernel (quad rows at 512 and 1024-row tiles, one row per
    lane at 256..4096) and stage size gives the oracle's y: bit-exact for
    integer x, and bit-identical to each other for random x."""
    if case == "rmat17":
        M, N, rp, col = rmat(17)
    elif case == "offdiag":
        M = N = 70001
        rp, col = _offdiag_band(M, 3, 1, 21)
    elif case == "long":
        M = N = 50003
        rp, col = _local_mix(M, 1, 40, 509, 200, 20000, 22)
    elif case == "plain_csr":
        M = N = 3000  # thumbnail blocks of 5 rows: no shape survives, S = 0
        rp, col = T.gen_rmat(M, 12, 30000, 0.25, 0.25, 0.25, 23)
    else:  # 64-row groups above 255 entries: 16-bit local offsets
        M = N = 120000
        rp, col = _local_mix(M, 6, 300, 0, 0, 0, 24)
    fmt = host_format(M, N, rp, col)
    A = T.Matrix(fmt)
    xi = torch.from_numpy(T.gen_vector(N, 1, 31)).to(dev)
    xr = torch.from_numpy(T.gen_vector(N, 2, 32)).to(dev)
    want = O.csr_spmv_f32(rp, col, xi.cpu().numpy())
    ref = None
    for rows in (256, 512, 1024, 2048, 4096):
        for capv in (1024, 4096):
            for flags in (0, 2):
                A.set_config(1, T.F32, rows, capv, 0, flags)
                y = torch.full((M,), float("nan"), device=dev)
                A.spmv(xi, y)
                torch.cuda.synchronize()
                assert np.array_equal(y.cpu().numpy(), want), (rows, capv, flags)
                A.spmv(xr, y)
                torch.cuda.synchronize()
                if ref is None:
                    ref = y.clone()
                    check_float(y.cpu().numpy(), rp, col, xr.cpu().numpy().astype(np.float64))
                assert torch.equal(y, ref), (rows, capv, flags)
    if case == "offdiag":
        assert fmt.dia_shapes == 1 and fmt.num_negs == 1
    if case == "long":
        assert A.stats()["long_rows"] > 0 and fmt.dia_shapes == 1 and fmt.num_negs == 1
    if case == "plain_csr":
        assert fmt.dia_shapes == 0
    if case == "wide_groups":
        assert A.stats()["offset_bytes"] == 2 and fmt.dia_shapes == 1 and fmt.num_negs == 1


@pytest.mark.parametrize("k", [1, 3, 4, 16, 64])
def test_multi_rhs(k, dev):
    M, N, rp, col = rmat(16)
    fmt = host_format(M, N, rp, col)
    X = T.gen_vector(N * k, 2, 10 + k).reshape(N, k)
    Y, _ = run_spmv(fmt, X, dev)
    Y64 = O.csr_spmm_f64(rp, col, X.astype(np.float64))
    absX = O.csr_spmm_f64(rp, col, np.abs(X.astype(np.float64)))
    den = np.maximum(np.abs(Y64), absX)
    den[den == 0] = 1
    assert float(np.max(np.abs(Y - Y64) / den)) <= 1e-6
    # every column equals the single-vector result bit for bit
    y0, _ = run_spmv(fmt, np.ascontiguousarray(X[:, 0]), dev)
    assert np.array_equal(Y[:, 0], y0)


@pytest.mark.parametrize("case", ["banded", "long"])
def test_k_columns_wide_equals_staged(case, dev):
    """For every k the wide-rows kernel serves (2..64), each tile size of it
    and the staged kernel give identical Y; integer X is bit-exact against the
    CPU CSR per column; BF16 X is exact to its quantised FP64 product."""
    if case == "banded":  # one shape, 17 offsets, 16-bit local offsets
        M = N = 30011
        rp, col = T.gen_banded(M, 17, 0.7, 41)
    else:  # a diagonal shape plus long rows
        M = N = 50003
        rp, col = _local_mix(M, 1, 40, 509, 200, 20000, 42)
    fmt = host_format(M, N, rp, col)
    A = T.Matrix(fmt)
    if case == "long":
        assert A.stats()["long_rows"] > 0
    for k in (2, 4, 8, 16, 32, 64):
        Xi = T.gen_vector(N * k, 1, 50 + k).reshape(N, k)
        Xd = torch.from_numpy(Xi).to(dev)
        outs = []
        cpl = 8 if k == 8 else min(k, 4)  # FP32 columns a lane owns (wide_cpl)
        for rows, flags in [(256 * cpl // k, 0), (512 * cpl // k, 0), (512 // k, 4)]:
            A.set_config(k, T.F32, rows, 4096, 0, flags)
            Y = torch.full((M, k), float("nan"), device=dev)
            A.spmv(Xd, Y)
            torch.cuda.synchronize()
            outs.append(Y.cpu().numpy())
        for o in outs[1:]:
            assert np.array_equal(o, outs[0]), k
        for c in (0, k - 1):
            assert np.array_equal(outs[0][:, c], O.csr_spmv_f32(rp, col, np.ascontiguousarray(Xi[:, c]))), (k, c)
    k = 8
    Xr = T.gen_vector(N * k, 2, 77).reshape(N, k)
    Xb = torch.from_numpy(Xr).to(dev).to(torch.bfloat16)
    Y = torch.empty(M, k, device=dev)
    A.spmv(Xb, Y)
    torch.cuda.synchronize()
    Xq = Xb.float().cpu().numpy().astype(np.float64)
    Y64 = O.csr_spmm_f64(rp, col, Xq)
    den = np.maximum(np.abs(Y64), O.csr_spmm_f64(rp, col, np.abs(Xq)))
    den[den == 0] = 1
    assert float(np.max(np.abs(Y.cpu().numpy() - Y64) / den)) <= 1e-6


@pytest.mark.parametrize("dt", [torch.bfloat16, torch.float16])
def test_half_storage(dt, dev):
    M, N, rp, col = rmat(16)
    fmt = host_format(M, N, rp, col)
    x = T.gen_vector(N, 2, 21)
    xq = torch.from_numpy(x).to(dt).float().numpy()  # RNE quantised x
    y, _ = run_spmv(fmt, x, dev, dtype=dt)
    check_float(y, rp, col, xq.astype(np.float64))


@pytest.mark.parametrize("kind", ["rmat", "stencil", "band"])
@pytest.mark.parametrize("dt", [torch.bfloat16, torch.float16])
def test_half_x_at_odd_offset(kind, dt, dev):
    """2-byte x gathers load the aligned 4-byte word that holds the element
    and pick its half: x starting at an odd element (the word's upper half
    for even i) and at an even one give the same y, bit for bit, for every
    k = 1 tile size, and integer x matches the CPU sum exactly."""
    if kind == "rmat":
        M, N, rp, col = rmat(15)
    elif kind == "stencil":  # 5-point stencil + 2 % uniform: all in the remainder
        M = N = 300 * 300
        rp, col = T.gen_stencil5(300, 300, 0.02, 2)
    else:
        M = N = 40000
        rp, col = T.gen_banded(M, 40, 0.7, 5)
    fmt, _ = T.partition_shapes(M, N, rp, col)
    A = T.Matrix(fmt)
    xi = np.random.default_rng(N).integers(0, 8, N).astype(np.float32)
    want = O.csr_spmv_f32(rp, col, xi)
    buf = torch.full((N + 3,), 7.0, device=dev, dtype=dt)  # neighbours of x hold 7
    code = T.BF16 if dt == torch.bfloat16 else T.F16
    for rows, fl in ((256, 2), (512, 2), (1024, 2), (2048, 2), (512, 0), (1024, 0)):
        A.set_config(1, code, rows, 4096, 0, fl)
        ys = []
        for off in (1, 2):
            buf[off:off + N] = torch.from_numpy(xi).to(dev).to(dt)
            y = torch.full((M,), float("nan"), device=dev)
            A.spmv(buf[off:off + N], y)
            torch.cuda.synchronize()
            ys.append(y.cpu().numpy())
            buf.fill_(7.0)
        assert np.array_equal(ys[0], want), (rows, fl)
        assert np.array_equal(ys[1], want), (rows, fl)


def test_batch_matches_alone(dev):
    mats = []
    for i, (scale, ef) in enumerate([(12, 4), (14, 16), (16, 8), (13, 32)]):
        n = (1 << scale) - 37 * i
        rp, col = T.gen_rmat(n, scale, ef * n, 0.499, 0.001, 0.001, 100 + i)
        mats.append((n, rp, col, host_format(n, n, rp, col)))
    items, alone = [], []
    for i, (n, rp, col, fmt) in enumerate(mats):
        k = [1, 2, 5, 8][i]
        X = torch.from_numpy(T.gen_vector(n * k, 0, 200 + i).reshape(n, k)).to(dev)
        if i == 1:
            X = X.to(torch.bfloat16)
        A = T.Matrix(fmt)
        Yb = torch.full((n, k), float("nan"), device=dev)
        Ya = torch.full((n, k), float("nan"), device=dev)
        A.spmv(X, Ya)
        items.append((A, X, Yb))
        alone.append(Ya)
    T.spmv_batch(items)
    torch.cuda.synchronize()
    for (A, X, Yb), Ya in zip(items, alone):
        assert torch.equal(Yb, Ya)


@pytest.mark.parametrize("count", [3, 5])
def test_batch_one_column_tile_sizes(count, dev):
    """One-column groups of 4 M rows or more run 1024-row tiles, smaller ones
    512-row tiles (3 and 5 matrices of about 1 M rows fall either side):
    every y equals the matrix run alone, bit for bit, and integer x equals the
    CPU sum."""
    items, want = [], []
    for i in range(count):
        n = (1 << 20) - 77 * i
        rp, col = T.gen_rmat(n, 20, 8 * n, 0.499, 0.001, 0.001, 300 + i)
        A = T.Matrix(host_format(n, n, rp, col))
        x = T.gen_vector(n, 1, 400 + i)  # integers 0..7
        X = torch.from_numpy(x).to(dev)
        Ya = torch.full((n,), float("nan"), device=dev)
        A.spmv(X, Ya)
        items.append((A, X, torch.full((n,), float("nan"), device=dev)))
        want.append((Ya, O.csr_spmv_f32(rp, col, x)))
    assert (sum(A.stats()["M"] for A, _, _ in items) >= 4 << 20) == (count == 5)
    T.spmv_batch(items)
    torch.cuda.synchronize()
    for (A, X, Yb), (Ya, yc) in zip(items, want):
        assert torch.equal(Yb, Ya)
        assert np.array_equal(Yb.cpu().numpy(), yc)


@pytest.mark.parametrize("ks", [(3, 7, 13, 20), (64, 1, 6, 11)])
def test_batch_column_chunks(dev, ks):
    """Requests wider than a chunk are cut into 8/4/2/1-column pieces, x types
    mixed, strided column views, a one-offset shape (single diagonal) and long
    rows: every Y equals the request run alone."""
    mats = []
    n0 = 9000
    rp, col = T.gen_banded(n0, 1, 1.0, 7)                      # one shape, one offset
    mats.append((n0, host_format(n0, n0, rp, col)))
    rp, col = T.gen_banded(n0 + 123, 9, 0.7, 8)                # one shape, 9 offsets
    mats.append((n0 + 123, host_format(n0 + 123, n0 + 123, rp, col)))
    n2, _, rp, col = rmat(14)
    mats.append((n2, host_format(n2, n2, rp, col)))
    n3 = 5000                                                  # long rows
    rows = [np.arange(0, n3, 3, dtype=np.int32) if r % 97 == 0 else np.array([r], np.int32) for r in range(n3)]
    rp = np.zeros(n3 + 1, np.int32)
    rp[1:] = np.cumsum([len(c) for c in rows])
    col = np.concatenate(rows).astype(np.int32)
    mats.append((n3, host_format(n3, n3, rp, col)))
    items, alone = [], []
    for i, ((n, fmt), k) in enumerate(zip(mats, ks)):
        dt = [torch.float32, torch.bfloat16, torch.float16, torch.float32][i]
        big = torch.from_numpy(T.gen_vector(n * (k + 3), 2, 300 + i).reshape(n, k + 3)).to(dev).to(dt)
        X = big[:, 2:2 + k]                                    # ldx = k + 3
        A = T.Matrix(fmt)
        Yb = torch.full((n, k + 1), float("nan"), device=dev)[:, :k]
        Ya = torch.full((n, k), float("nan"), device=dev)
        A.spmv(X, Ya)
        items.append((A, X, Yb))
        alone.append(Ya)
    T.spmv_batch(items)
    torch.cuda.synchronize()
    for (A, X, Yb), Ya in zip(items, alone):
        assert torch.equal(Yb, Ya)


def test_graph_replay_matches_eager(dev):
    M, N, rp, col = rmat(16)
    fmt = host_format(M, N, rp, col)
    A = T.Matrix(fmt)
    x = torch.from_numpy(T.gen_vector(N, 0, 4)).to(dev)
    ye = torch.empty(M, device=dev)
    A.spmv(x, ye)
    yg = torch.full((M,), float("nan"), device=dev)
    s = torch.cuda.Stream()
    g = T.Graph(A, x, yg, iters=5, stream=s.cuda_stream)
    for rep in range(3):
        yg.fill_(float("nan"))
        torch.cuda.synchronize()
        g.launch(stream=s.cuda_stream)
        s.synchronize()
        assert torch.equal(yg, ye), rep
        torch.cuda.synchronize()


def test_taichi_spmv_host_path(dev):
    M, N, rp, col = cfg1()
    fmt, _ = T.partition_shapes(M, N, rp, col)
    x = np.ones(N, np.float32)  # main.cu:54-55
    y, t_ms, s_ms = T.taichi_spmv(fmt, x, warmup=2, runs=5)
    assert np.array_equal(y, O.csr_spmv_f32(rp, col, x))
    assert s_ms > 0


@pytest.mark.parametrize("case,dt", [("banded", torch.float32), ("banded", torch.bfloat16),
                                     ("banded", torch.float16), ("long", torch.float32)],
                         ids=["banded-f32", "banded-bf16", "banded-f16", "long-f32"])
def test_masked_column_chunks(case, dt, dev):
    """k = 3, 5, 6, 7 (and 1..8) columns of an 8-column buffer run as one
    masked wide-rows launch: columns past k are neither read nor written,
    each column equals the single-vector result, and the batched launch over
    the same views equals the plain one. With 2-byte x one lane owns all 8
    columns, so 5, 6 and 7 columns are one masked 8-wide pass (its last-row
    loads, and stores of the first 5 or 6 columns) also on a band with more
    than two dense diagonals, in tc_spmv and in the batch."""
    if case == "banded":
        M = N = 20011
        rp, col = T.gen_banded(M, 13, 0.7, 61)
    else:
        M = N = 50003
        rp, col = _local_mix(M, 1, 40, 509, 200, 20000, 62)
    fmt = host_format(M, N, rp, col)
    A = T.Matrix(fmt)
    if case == "banded":
        assert A.stats()["num_negs"] > 2
    # integers 0..7: exact in BF16 and FP16 too
    X8 = torch.from_numpy(T.gen_vector(N * 8, 1, 70).reshape(N, 8)).to(dev).to(dt)
    cols1 = []
    for c in range(8):
        y = torch.empty(M, device=dev)
        A.spmv(X8[:, c].contiguous(), y)
        cols1.append(y)
    torch.cuda.synchronize()
    items = []
    for k in range(1, 9):
        Y8 = torch.full((M, 8), float("nan"), device=dev)
        A.spmv(X8[:, :k], Y8[:, :k])
        torch.cuda.synchronize()
        assert torch.isnan(Y8[:, k:]).all(), k                # nothing written past k
        for c in range(k):
            assert torch.equal(Y8[:, c], cols1[c]), (k, c)
        assert np.array_equal(Y8[:, 0].cpu().numpy(),
                              O.csr_spmv_f32(rp, col, X8[:, 0].float().cpu().numpy()))
        Yb = torch.full((M, 8), float("nan"), device=dev)
        items.append((A, X8[:, :k], Yb[:, :k], Y8, Yb, k))
    T.spmv_batch([(a, x, y) for a, x, y, _, _, _ in items])
    torch.cuda.synchronize()
    for _, _, _, Y8, Yb, k in items:
        assert torch.equal(Yb[:, :k], Y8[:, :k]), k
        assert torch.isnan(Yb[:, k:]).all(), k
    # the last rows of a tight buffer: X of exactly k columns at an offset
    # that keeps rows aligned (ldx = 8, no column past k exists in memory)
    for k in (3, 5, 7):
        Xt = X8[:, :k].contiguous()                           # ldx = k: unaligned -> staged
        Yt = torch.full((M, k), float("nan"), device=dev)
        A.spmv(Xt, Yt)
        torch.cuda.synchronize()
        for c in range(k):
            assert torch.equal(Yt[:, c], cols1[c]), (k, c)
    # masked path over a buffer that ends at column k of X's last row, with
    # NaN in every slot past column k: a masked chunk must not use them
    for k in (1, 3, 5, 6, 7):
        flat = torch.full(((N - 1) * 8 + k,), float("nan"), device=dev, dtype=dt)
        Xs = flat.as_strided((N, k), (8, 1))
        Xs.copy_(X8[:, :k])
        Ys = torch.full((M, 8), float("nan"), device=dev)[:, :k]
        A.spmv(Xs, Ys)
        Yb = torch.full((M, 8), float("nan"), device=dev)[:, :k]
        T.spmv_batch([(A, Xs, Yb)])
        torch.cuda.synchronize()
        for c in range(k):
            assert torch.equal(Ys[:, c], cols1[c]), (k, c)
            assert torch.equal(Yb[:, c], cols1[c]), (k, c)


@pytest.mark.parametrize("kind", ["rmat", "stencil", "band"])
def test_strided_single_column(kind, dev):
    """k = 1 on one column of row-major X and Y (ldx = ldy = 8, as a served
    request's column is) runs the one-row-per-lane kernel with strides and
    gives the plain vector's result, bit for bit, for every tile size and x
    type; integer x matches the CPU sum exactly."""
    if kind == "rmat":
        M, N, rp, col = rmat(15)
    elif kind == "stencil":
        M = N = 128 * 128
        rp, col = T.gen_diagonals_random(M, [-128, -1, 0, 1, 128], M // 50, 3)
    else:
        M = N = 40000
        rp, col = T.gen_banded(M, 40, 0.7, 5)
    fmt, _ = T.partition_shapes(M, N, rp, col)
    A = T.Matrix(fmt)
    rng = np.random.default_rng(M)
    xi = rng.integers(0, 8, N).astype(np.float32)
    want = O.csr_spmv_f32(rp, col, xi)
    for dt in (torch.float32, torch.bfloat16):
        X8 = torch.zeros(N, 8, device=dev, dtype=dt)
        X8[:, 3] = torch.from_numpy(xi).to(dev).to(dt)
        xv = X8[:, 3].contiguous()
        for rows in (256, 512, 1024):
            A.set_config(1, T.BF16 if dt == torch.bfloat16 else T.F32, rows, 4096, 0, 2)
            Y8 = torch.full((M, 8), float("nan"), device=dev)
            A.spmv(X8[:, 3:4], Y8[:, 3:4])
            yv = torch.full((M,), float("nan"), device=dev)
            A.spmv(xv, yv)
            torch.cuda.synchronize()
            assert torch.equal(Y8[:, 3], yv), (dt, rows)
            assert torch.isnan(Y8[:, :3]).all() and torch.isnan(Y8[:, 4:]).all(), (dt, rows)
            assert np.array_equal(yv.cpu().numpy(), want), (dt, rows)
