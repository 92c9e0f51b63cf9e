"""tc_spmv_batch_static, the batched launch that allocates nothing (its
descriptor table is a fixed per-GPU array of the library): called through
the C ABI, its results must equal tc_spmv_batch's bit for bit, on two calls
in a row -- the second smaller than the first -- and a batch whose
descriptors do not fit the array is refused before anything runs."""
import numpy as np
import pytest
import torch

import taichi_amd as T
from helpers import host_format

TC_ERR_CAPACITY = -105


def _mats(specs):
    out = []
    for kind, n, a, seed in specs:
        if kind == "rmat":
            scale = int(np.ceil(np.log2(n)))
            rp, col = T.gen_rmat(n, scale, a * n, 0.499, 0.001, 0.001, seed)
        else:
            rp, col = T.gen_banded(n, a, 0.7, seed)
        out.append((n, T.Matrix(host_format(n, n, rp, col), nnz=len(col))))
    return out


def _items(mats, ks, dev, seed):
    items = []
    for i, ((n, A), k) in enumerate(zip(mats, ks)):
        X = torch.from_numpy(T.gen_vector(n * k, 0, seed + i).reshape(n, k)).to(dev)
        if i % 3 == 1:
            X = X.to(torch.bfloat16)
        items.append((A, X if k > 1 else X[:, 0].contiguous(), None, k, n))
    return items


def _run_both(items, dev):
    """(tc_spmv_batch's Ys, tc_spmv_batch_static's Ys); every Y starts as NaN"""
    outs = []
    for static in (False, True):
        trip = []
        for A, X, _, k, n in items:
            Y = torch.full((n, k), float("nan"), device=dev)
            trip.append((A, X, Y if k > 1 else Y[:, 0]))
        b = T.Batch(trip)
        (b.launch_static if static else b.launch)()
        torch.cuda.synchronize()
        outs.append([y.clone() for _, _, y in trip])
    return outs


@pytest.mark.gpu
def test_static_batch_equals_batch_twice(dev):
    # first call: 9 requests, k in 1..16, banded (presence bitmaps) and R-MAT
    big = _mats([("rmat", 70000, 16, 1), ("band", 50000, 24, 2), ("rmat", 30011, 8, 3),
                 ("band", 20000, 5, 4), ("rmat", 65536, 4, 5), ("band", 40000, 40, 6),
                 ("rmat", 9000, 32, 7), ("band", 12345, 3, 8), ("rmat", 100000, 16, 9)])
    ks = [1, 3, 8, 16, 5, 1, 7, 2, 12]
    ref, got = _run_both(_items(big, ks, dev, 100), dev)
    for i, (a, b) in enumerate(zip(ref, got)):
        assert not torch.isnan(a).any(), i
        assert torch.equal(a, b), i
    # second call right after, smaller: fewer, smaller matrices and requests
    small = _mats([("rmat", 5000, 8, 11), ("band", 3000, 9, 12), ("rmat", 4100, 16, 13)])
    ref, got = _run_both(_items(small, [2, 1, 4], dev, 300), dev)
    for i, (a, b) in enumerate(zip(ref, got)):
        assert not torch.isnan(a).any(), i
        assert torch.equal(a, b), i


@pytest.mark.gpu
def test_static_batch_refuses_oversize(dev):
    # 700 requests of 64 columns are 5,600 chunk descriptors: more than the
    # fixed table holds; the call is refused before anything is enqueued
    (n, A), = _mats([("rmat", 4096, 4, 21)])
    X = torch.zeros(n, 64, device=dev)
    Y = torch.full((n, 64), float("nan"), device=dev)
    b = T.Batch([(A, X, Y)] * 700)
    with pytest.raises(T.TaiChiError) as e:
        b.launch_static()
    assert e.value.status == TC_ERR_CAPACITY
    torch.cuda.synchronize()
    assert torch.isnan(Y).all()
    # a call that fits still runs afterwards
    T.Batch([(A, X, Y)]).launch_static()
    torch.cuda.synchronize()
    assert torch.equal(Y, torch.zeros_like(Y))
