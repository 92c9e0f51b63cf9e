"""Serving runtime: continuous batching over resident matrices. Each result
must be bit-identical alone (tc_spmv) and in the mix; admission follows the
HBM footprint."""
import numpy as np
import pytest
import torch

import taichi_amd as T
from helpers import host_format
from taichi_amd import workloads as W

pytestmark = pytest.mark.gpu


def _mats(count=6):
    out = []
    for i in range(count):
        if i % 2:
            n = 20000 + 1111 * i
            rp, col = T.gen_banded(n, 5 + 3 * i, 0.7, 40 + i)
        else:
            n, _, rp, col = W.rmat(30000 + 777 * i, 8, 60 + i)
        out.append((n, rp, col, host_format(n, n, rp, col)))
    return out


@pytest.mark.parametrize("launch", [0, 1, 2])  # coalesced (default), coalesced, one by one
def test_served_results_equal_alone(launch, dev):
    mats = _mats()
    srv = T.Server(hbm_cap_bytes=1 << 30, max_batch=16, streams=3, max_inflight=64, launch=launch)
    ids = [srv.add_matrix(fmt) for (_, _, _, fmt) in mats]
    alone = [T.Matrix(fmt) for (_, _, _, fmt) in mats]
    rng = np.random.default_rng(0)
    reqs = []
    for q in range(120):
        i = int(rng.integers(len(mats)))
        n = mats[i][0]
        k = int(rng.integers(1, 9))
        dt = [torch.float32, torch.bfloat16, torch.float16][q % 3]
        X = torch.from_numpy(T.gen_vector(n * k, 2, 1000 + q).reshape(n, k)).to(dev).to(dt)
        Y = torch.full((n, k), float("nan"), device=dev)
        reqs.append((i, X, Y, srv.submit(ids[i], X, Y)))
    for i, X, Y, t in reqs:
        srv.wait(t)
    torch.cuda.synchronize()
    for i, X, Y, t in reqs:
        Ya = torch.full_like(Y, float("nan"))
        alone[i].spmv(X, Ya)
        torch.cuda.synchronize()
        assert torch.equal(Y, Ya)
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["batches"] >= 1
    srv.close()


def test_lone_requests_equal_alone(dev):
    """Requests submitted one at a time, each waited for, make batches of one,
    which launch as tc_spmv: every k in 1..8 and x type equals the matrix run
    alone, and the padding columns of Y stay untouched."""
    mats = _mats(2)
    srv = T.Server(hbm_cap_bytes=1 << 30, max_batch=16, streams=2, max_inflight=16)
    ids = [srv.add_matrix(fmt) for (_, _, _, fmt) in mats]
    for i, (n, rp, col, fmt) in enumerate(mats):
        A = T.Matrix(fmt)
        for k in range(1, 9):
            for dt in (torch.float32, torch.bfloat16):
                X = torch.from_numpy(T.gen_vector(n * 8, 2, 50 * k + i).reshape(n, 8)).to(dev).to(dt)
                Y = torch.full((n, 8), float("nan"), device=dev)
                srv.wait(srv.submit(ids[i], X[:, :k], Y[:, :k]))
                torch.cuda.synchronize()
                Ya = torch.full((n, 8), float("nan"), device=dev)
                A.spmv(X[:, :k], Ya[:, :k])
                torch.cuda.synchronize()
                assert torch.equal(Y[:, :k], Ya[:, :k]), (i, k, dt)
                assert torch.isnan(Y[:, k:]).all(), (i, k, dt)
    st = srv.stats()
    assert st["completed"] == 2 * 8 * 2 and st["batches"] == st["completed"]
    srv.close()


def test_admission_by_footprint(dev):
    mats = _mats(4)
    fps = [T.footprint(fmt) for (_, _, _, fmt) in mats]
    cap = fps[0] + fps[1] + fps[2] // 2
    srv = T.Server(hbm_cap_bytes=cap)
    a = srv.add_matrix(mats[0][3])
    b = srv.add_matrix(mats[1][3])
    assert srv.stats()["resident_bytes"] == fps[0] + fps[1]
    with pytest.raises(T.TaiChiError) as e:
        srv.add_matrix(mats[2][3])
    assert e.value.status == -105  # TC_ERR_CAPACITY
    assert srv.stats()["resident_bytes"] == fps[0] + fps[1]
    srv.remove_matrix(a)
    c = srv.add_matrix(mats[2][3])
    assert srv.stats()["resident_bytes"] == fps[1] + fps[2]
    # the footprint is what the device allocation really is
    for (_, _, _, fmt), fp in zip(mats, fps):
        assert T.Matrix(fmt).stats()["device_bytes"] == fp
    del b, c
    srv.close()


def test_load_generator_open_and_closed(dev):
    mats = _mats(4)
    srv = T.Server(hbm_cap_bytes=1 << 30, max_batch=32, streams=2, max_inflight=64)
    items, keep = [], []
    for (n, _, _, fmt) in mats:
        mid = srv.add_matrix(fmt)
        for dt, code in ((torch.float32, T.F32), (torch.bfloat16, T.BF16)):
            X = torch.rand(n, 8, device=dev).to(dt)
            keep.append(X)
            for k in (1, 4, 8):
                items.append((mid, X.data_ptr(), 8, code, k))
    nmax = max(m[0] for m in mats)
    ybufs = [torch.empty(nmax, 8, device=dev) for _ in range(64)]
    r = srv.run_load(items, ybufs, 8, count=400, rate=20000.0, seed=1)
    assert r["completed"] == 400 and 0 < r["p50_us"] <= r["p99_us"] <= r["max_us"]
    r = srv.run_load(items, ybufs, 8, count=400, rate=0.0, concurrency=48, seed=2)
    assert r["completed"] == 400 and r["req_per_s"] > 0 and r["mean_batch"] >= 1
    p = srv.run_plain(items, ybufs, 8, count=100, seed=3)
    assert p["completed"] == 100
    srv.close()


@pytest.mark.parametrize("k,ld", [(1, 1), (3, 4), (6, 8)])
def test_load_generator_per_request_strides(k, ld, dev):
    """ldy = 0: each request's y has the row stride of its own x (the serving
    bench's per-request buffers, rows padded to the kernels' vector widths).
    One item, so every request computes the same y: each output buffer
    equals tc_spmv alone, and the padding columns stay untouched."""
    n, rp, col, fmt = _mats(2)[1]  # a band
    srv = T.Server(hbm_cap_bytes=1 << 30, max_batch=16, streams=2, max_inflight=32)
    mid = srv.add_matrix(fmt)
    X = torch.from_numpy(T.gen_vector(n * ld, 2, 77).reshape(n, ld)).to(dev)
    ybufs = [torch.full((n, ld), float("nan"), device=dev) for _ in range(8)]
    r = srv.run_load([(mid, X.data_ptr(), ld, T.F32, k)], ybufs, 0, count=64, rate=0.0, concurrency=8, seed=5)
    assert r["completed"] == 64
    p = srv.run_plain([(mid, X.data_ptr(), ld, T.F32, k)], ybufs, 0, count=16, seed=6)
    assert p["completed"] == 16
    torch.cuda.synchronize()
    Ya = torch.full((n, ld), float("nan"), device=dev)
    T.Matrix(fmt).spmv(X[:, :k], Ya[:, :k])
    torch.cuda.synchronize()
    for Y in ybufs:
        assert torch.equal(Y[:, :k], Ya[:, :k])
        assert torch.isnan(Y[:, k:]).all()
    srv.close()
