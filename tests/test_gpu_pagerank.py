"""Iterative driver: PageRank by power iteration on the device. Ranks match an
FP64 PageRank of the same CSR (the definition the driver implements) within
1e-5 relative, sum to 1, and a captured graph replays to bit-identical ranks,
also after other work ran in between."""
import numpy as np
import pytest
import torch

import taichi_amd as T
from helpers import host_format, rmat

pytestmark = pytest.mark.gpu


def _pagerank64(M, rp, col, alpha, iters):
    """FP64 power iteration: A[i][j] = 1 for each entry (i, j) of the CSR."""
    deg = np.bincount(col, minlength=M).astype(np.float64)
    inv = np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0)
    rows = np.repeat(np.arange(M), np.diff(rp))
    r = np.full(M, 1.0 / M)
    for _ in range(iters):
        z = r * inv
        dang = r[deg == 0].sum()
        y = np.bincount(rows, weights=z[col], minlength=M)
        r = (1 - alpha) / M + alpha * (y + dang / M)
    return r


@pytest.mark.parametrize("case", ["rmat16", "banded", "odd_n", "unaligned_r"])
def test_pagerank_matches_fp64(case, dev):
    if case == "rmat16":  # pages j % 97 == 0 lose their out-links (dangling)
        M, N, rp, col = rmat(16)
        keep = col % 97 != 0
        rows = np.repeat(np.arange(M), np.diff(rp))[keep]
        col = col[keep].astype(np.int32)
        rp = np.zeros(M + 1, np.int32)
        rp[1:] = np.cumsum(np.bincount(rows, minlength=M))
    else:  # odd_n: n % 4 = 1 (vector tail); unaligned_r: r 4 bytes off (scalar path)
        M = N = 40001 if case == "odd_n" else 40000
        rp, col = T.gen_banded(M, 9, 0.7, 13)
    fmt = host_format(M, N, rp, col)
    A = T.Matrix(fmt)
    inv = torch.from_numpy(T.inverse_out_degree(N, col)).to(dev)
    pr = T.PageRank(A, inv, alpha=0.85)
    iters = 30
    r = torch.empty(M + 1, device=dev)[1:] if case == "unaligned_r" else torch.empty(M, device=dev)
    pr.run(r, iters)
    torch.cuda.synchronize()
    want = _pagerank64(M, rp, col, 0.85, iters)
    got = r.cpu().numpy().astype(np.float64)
    assert np.max(np.abs(got - want) / want) <= 1e-5
    assert abs(got.sum() - 1.0) <= 1e-5
    if case == "rmat16":
        assert (np.bincount(col, minlength=M) == 0).any()  # dangling pages exercised


def test_pagerank_graph_replays_identically(dev):
    M, N, rp, col = rmat(15)
    fmt = host_format(M, N, rp, col)
    A = T.Matrix(fmt)
    inv = torch.from_numpy(T.inverse_out_degree(N, col)).to(dev)
    pr = T.PageRank(A, inv)
    r_eager = torch.empty(M, device=dev)
    pr.run(r_eager, 20)
    torch.cuda.synchronize()
    r = torch.full((M,), float("nan"), device=dev)
    s = torch.cuda.Stream()
    g = pr.capture(r, 20, stream=s.cuda_stream)
    for rep in range(3):
        r.fill_(float("nan"))
        torch.cuda.synchronize()
        g.launch(stream=s.cuda_stream)
        s.synchronize()
        assert torch.equal(r, r_eager), rep
        # unrelated eager work between replays
        other = torch.empty(M, device=dev)
        A.spmv(inv, other)
        torch.cuda.synchronize()
    g.close()
