"""SpMV parity at the benchmark configurations' full sizes, against the
oracle. The reference checks y against a host CSR SpMV (main.cu:53-86, x = 1,
the matrix it ran); here every configuration is checked at its real size:
integer-valued x bit for bit against the CPU FP32 CSR sums of main.cu:61-69,
real x within 1e-6 componentwise of an FP64 product.

  R-MAT 2^22 (a,b,c,d = .499,.001,.001,.499, edge factor 16, seed 1): the
      device partition equals the oracle's field by field; every k = 1
      kernel configuration gives the oracle's y
  2048^2 5-point stencil + 2% uniform ones (seed 2): k in {1, 4, 16, 64},
      FP32 and BF16 X
  256 uneven R-MAT matrices (seeds 100..355) in one batched launch
  the 64-matrix serving mix through the server, FP32 and BF16 x, k 1..8,
      coalesced launches and one launch per request
"""
import numpy as np
import pytest
import torch

import oracle as O
import taichi_amd as T
from taichi_amd import workloads as W

from helpers import REL_TOL

pytestmark = pytest.mark.gpu

ARRAYS = ("xStart", "yStart", "xStop", "nDiasPtr", "neg_offsets", "csrRow", "csrCol")


def _bound(Y, rp, col, Xq):
    """max componentwise error of Y (M x k) against the FP64 product of the
    binary CSR with Xq (the x values the kernel read), denominator
    max(|A Xq|, A |Xq|)"""
    Xq = np.asarray(Xq, np.float64)
    if Xq.ndim == 1:
        Xq = Xq[:, None]
        Y = np.asarray(Y).reshape(-1, 1)
    Y64 = O.csr_spmm_f64(rp, col, np.ascontiguousarray(Xq))
    den = np.maximum(np.abs(Y64), O.csr_spmm_f64(rp, col, np.ascontiguousarray(np.abs(Xq))))
    den[den == 0] = 1.0
    return float(np.max(np.abs(np.asarray(Y, np.float64) - Y64) / den))


# ---- R-MAT 2^22 (the headline) ----

@pytest.fixture(scope="module")
def rmat22(dev):
    n = 1 << 22
    rp, col = T.gen_rmat(n, 22, 16 * n, 0.499, 0.001, 0.001, 1)
    fmt, _ = T.partition_shapes(n, n, rp, col)  # the drop-in call: device build
    return n, rp, col, fmt


def test_rmat22_partition_equals_oracle(rmat22):
    n, rp, col, fmt = rmat22
    of, _ = O.partition(n, n, rp, col)
    for f in O.Format.FIELDS:
        assert getattr(fmt, f) == getattr(of, f), f
    for f in ARRAYS:
        assert np.array_equal(getattr(fmt, f), getattr(of, f)), f
    assert fmt.dia_shapes == 1 and fmt.num_negs == 1


def test_rmat22_integer_x_exact_every_config(rmat22, dev):
    n, rp, col, fmt = rmat22
    A = T.Matrix(fmt, nnz=len(col))
    x = T.gen_vector(n, 1, 17)  # integers 0..7
    want = O.csr_spmv_f32(rp, col, x)
    xd = torch.from_numpy(x).to(dev)
    y = torch.empty(n, device=dev)
    # (rows_per_tile, lds_values, grid, flags): the default, the quad-row
    # kernel, every rows-tile size, a persistent grid
    for cfg in [(0, 0, 0, 0), (512, 4096, 0, 2), (1024, 4096, 0, 0), (512, 2048, 0, 0), (256, 4096, 0, 2),
                (1024, 4096, 0, 2), (2048, 8192, 0, 2), (4096, 8192, 0, 2), (512, 4096, 2048, 2)]:
        if cfg != (0, 0, 0, 0):
            A.set_config(1, T.F32, *cfg)
        y.fill_(float("nan"))
        A.spmv(xd, y)
        torch.cuda.synchronize()
        assert np.array_equal(y.cpu().numpy(), want), cfg


def test_rmat22_real_x_within_bound(rmat22, dev):
    n, rp, col, fmt = rmat22
    A = T.Matrix(fmt, nnz=len(col))
    x = T.gen_vector(n, 0, 12345)  # U[0,1)
    y = torch.empty(n, device=dev)
    A.spmv(torch.from_numpy(x).to(dev), y)
    torch.cuda.synchronize()
    assert _bound(y.cpu().numpy(), rp, col, x) <= REL_TOL


# ---- 2048^2 stencil, multi-RHS ----

@pytest.fixture(scope="module")
def stencil(dev):
    M, N, rp, col = W.stencil5()
    fmt, _ = T.partition_shapes(M, N, rp, col)
    return M, N, rp, col, T.Matrix(fmt, nnz=len(col)), fmt


@pytest.mark.parametrize("xdt", ["f32", "bf16"])
@pytest.mark.parametrize("k", [1, 4, 16, 64])
def test_stencil_multi_rhs_within_bound(stencil, k, xdt, dev):
    M, N, rp, col, A, fmt = stencil
    X = np.random.default_rng(1000 + k).standard_normal((N, k)).astype(np.float32)
    Xd = torch.from_numpy(X).to(dev)
    if xdt == "bf16":
        Xd = Xd.to(torch.bfloat16)
    Xq = Xd.float().cpu().numpy()  # what the kernel reads (RNE-quantised for BF16)
    Y = torch.full((M, k), float("nan"), device=dev)
    if k == 1:
        A.spmv(Xd[:, 0].contiguous(), Y[:, 0])
    else:
        A.spmv(Xd, Y)
    torch.cuda.synchronize()
    Yh = Y.cpu().numpy()
    assert not np.isnan(Yh).any()
    assert _bound(Yh, rp, col, Xq) <= REL_TOL


@pytest.mark.parametrize("xdt", ["f32", "bf16"])
def test_stencil_integer_columns_exact(stencil, xdt, dev):
    M, N, rp, col, A, fmt = stencil
    k = 16
    X = np.random.default_rng(7).integers(0, 8, (N, k)).astype(np.float32)
    Xd = torch.from_numpy(X).to(dev)
    if xdt == "bf16":
        Xd = Xd.to(torch.bfloat16)  # small integers are exact in BF16
    Y = torch.empty(M, k, device=dev)
    A.spmv(Xd, Y)
    torch.cuda.synchronize()
    Yh = Y.cpu().numpy()
    for c in range(k):
        assert np.array_equal(Yh[:, c], O.csr_spmv_f32(rp, col, np.ascontiguousarray(X[:, c]))), c


# ---- 256 uneven R-MAT matrices, one batched launch ----

def test_uneven_256_batched_exact(dev):
    items, refs, csrs = [], [], []
    for (n, _, rp, col, s) in W.uneven_rmat():
        fmt, _ = T.partition_shapes(n, n, rp, col)
        x = T.gen_vector(n, 1, s)  # integers 0..7
        items.append((T.Matrix(fmt, nnz=len(col)), torch.from_numpy(x).to(dev),
                      torch.full((n,), float("nan"), device=dev)))
        refs.append(O.csr_spmv_f32(rp, col, x))
        csrs.append((rp, col))
    assert len(items) == 256
    b = T.Batch(items)
    for launch in (b.launch, b.launch_static):
        for _, _, y in items:
            y.fill_(float("nan"))
        launch()
        torch.cuda.synchronize()
        for i, ((_, _, y), want) in enumerate(zip(items, refs)):
            assert np.array_equal(y.cpu().numpy(), want), (launch.__name__, i)
    # real x: every matrix within the bound
    xs = []
    for i, (A, _, y) in enumerate(items):
        x = T.gen_vector(A.stats()["N"], 0, 500 + i)
        xs.append(x)
        items[i] = (A, torch.from_numpy(x).to(dev), y)
    T.Batch(items).launch()
    torch.cuda.synchronize()
    for (A, _, y), x, (rp, col) in zip(items, xs, csrs):
        assert _bound(y.cpu().numpy(), rp, col, x) <= REL_TOL


# ---- the 64-matrix serving mix ----

@pytest.fixture(scope="module")
def serving_mix(dev):
    out = []
    for kind, n, rp, col in W.serving_mix():
        fmt, _ = T.partition_shapes(n, n, rp, col)
        out.append((kind, n, rp, col, fmt))
    return out


@pytest.mark.parametrize("launch", [0, 2])
def test_serving_mix_vs_oracle(serving_mix, launch, dev):
    """launch 0 (the default): every batch of requests goes through one
    coalesced batched launch group; 2: one by one. A request for each (x
    type, k in 1..8) on a spread of the 64 matrices, all in flight together:
    each y within the FP64 bound and equal to tc_spmv alone; integer x
    bit-exact."""
    srv = T.Server(hbm_cap_bytes=200 << 30, max_batch=64, streams=3, max_inflight=256, launch=launch)
    ids = [srv.add_matrix(fmt, nnz=len(col)) for (_, _, _, col, fmt) in serving_mix]
    reqs = []
    for dt in (torch.float32, torch.bfloat16):
        for k in range(1, 9):
            j = (11 * k + (29 if dt == torch.bfloat16 else 0)) % len(serving_mix)
            kind, n, rp, col, fmt = serving_mix[j]
            integer = (k + (dt == torch.bfloat16)) % 3 == 0
            X = (np.random.default_rng(j * 100 + k).integers(0, 8, (n, k)) if integer
                 else np.random.default_rng(j * 100 + k).random((n, k))).astype(np.float32)
            Xd = torch.from_numpy(X).to(dev).to(dt)
            Y = torch.full((n, k), float("nan"), device=dev)
            reqs.append((j, k, integer, Xd, Y))
    tickets = [srv.submit(ids[j], Xd if k > 1 else Xd[:, 0], Y if k > 1 else Y[:, 0])
               for j, k, _, Xd, Y in reqs]
    for t in tickets:
        srv.wait(t)
    st = srv.stats()
    assert st["completed"] == len(reqs)
    assert st["batches"] > 0
    kinds = set()
    for j, k, integer, Xd, Y in reqs:
        kind, n, rp, col, fmt = serving_mix[j]
        kinds.add(kind)
        Yh = Y.cpu().numpy()
        Xq = Xd.float().cpu().numpy()
        if integer:
            for c in range(k):
                assert np.array_equal(Yh[:, c], O.csr_spmv_f32(rp, col, np.ascontiguousarray(Xq[:, c]))), (j, k, c)
        else:
            assert _bound(Yh, rp, col, Xq) <= REL_TOL, (j, k)
        # the same request alone
        A = T.Matrix(fmt, nnz=len(col))
        Ya = torch.empty(n, k, device=dev)
        if k == 1:
            A.spmv(Xd[:, 0], Ya[:, 0])
        else:
            A.spmv(Xd, Ya)
        torch.cuda.synchronize()
        assert torch.equal(Ya, Y), (j, k)
        A.close()
    assert kinds == {"banded", "rmat"}
    srv.close()
