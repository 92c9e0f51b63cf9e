"""Shared test helpers: workloads from the generators, device runs, checks."""
import json
import os

import numpy as np

import oracle as O
import taichi_amd as T

# FP32 tolerance against the FP64 oracle, componentwise:
# max_i |y_i - y64_i| / max(|y64_i|, sum_j |x_j| over row i)
REL_TOL = 1e-6


def cfg1():
    """n=16,384, ones on offsets -2..2 plus 16,384 uniform draws (seed 0)."""
    n = 16384
    rp, col = T.gen_diagonals_random(n, [-2, -1, 0, 1, 2], n, 0)
    return n, n, rp, col


def rmat(scale, ef=16, seed=1, n=None):
    n = n or (1 << scale)
    rp, col = T.gen_rmat(n, scale, ef * (1 << scale), 0.499, 0.001, 0.001, seed)
    return n, n, rp, col


def host_format(M, N, rp, col):
    """Product host partition from the oracle's thumbnail (CPU only)."""
    img = O.thumbnail(M, N, rp, col)
    fmt, _ = T.partition_from_image(M, N, rp, col, img=img)
    return fmt


def check_float(y, rp, col, x64, tol=REL_TOL):
    y64 = O.csr_spmv_f64(rp, col, x64)
    absrow = O.csr_absrow_f64(rp, col, x64)
    err = O.rel_err(y, y64, absrow)
    assert err <= tol, f"relative error {err:.3e} > {tol:.0e}"
    return err


def run_spmv(fmt, x, dev, k=None, dtype=None):
    """y = A x on the GPU; x numpy float32 [N] or [N, k]."""
    import torch
    A = T.Matrix(fmt)
    xd = torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    if dtype is not None:
        xd = xd.to(dtype)
    shape = (fmt.M,) if x.ndim == 1 else (fmt.M, x.shape[1])
    yd = torch.full(shape, float("nan"), dtype=torch.float32, device=dev)
    A.spmv(xd, yd)
    torch.cuda.synchronize()
    return yd.cpu().numpy(), A


def morph_images():
    """Thumbnails for the morphology checks: (name, uint8 [H, W], 0 = hit)."""
    out = []
    for seed in range(6):
        rng = np.random.default_rng(300 + seed)
        H, W = 512, int(rng.integers(300, 513))
        a = np.full((H, W), 255, np.uint8)
        for _ in range(int(rng.integers(1, 6))):        # diagonal segments
            r0, c0, ln = rng.integers(0, 400), rng.integers(0, W - 50), rng.integers(60, 300)
            t = np.arange(int(ln))
            keep = (r0 + t < H) & (c0 + t < W)
            a[r0 + t[keep], c0 + t[keep]] = 0
        a[rng.random((H, W)) < rng.uniform(0.0, 0.05)] = 0
        if seed % 2:
            a[100:160, 200:260] = 0                       # a dense block
        out.append((f"random{seed}", a))
    a = np.full((512, 512), 255, np.uint8)                # nested closed rings: holes in holes
    for k in range(0, 240, 12):
        a[k, k:512 - k] = 0
        a[511 - k, k:512 - k] = 0
        a[k:512 - k, k] = 0
        a[k:512 - k, 511 - k] = 0
    out.append(("rings", a))
    a = np.full((512, 512), 255, np.uint8)                # rings with alternating gaps: a maze
    for i, k in enumerate(range(2, 250, 4)):
        a[k, k:512 - k] = 0
        a[511 - k, k:512 - k] = 0
        a[k:512 - k, k] = 0
        a[k:512 - k, 511 - k] = 0
        g = k + 2 if i % 2 else 509 - k
        a[k, g] = 255                                     # gap on the top side, alternating ends
    out.append(("maze", a))
    yy, xx = np.mgrid[0:512, 0:512]
    out.append(("checker", (((yy + xx) % 2) * 255).astype(np.uint8)))
    out.append(("all_hit", np.zeros((512, 512), np.uint8)))
    out.append(("empty", np.full((512, 512), 255, np.uint8)))
    rng = np.random.default_rng(7)
    for W in (300, 700):
        a = np.full((512, W), 255, np.uint8)
        for i in range(min(512, W)):
            a[i, i] = 0
        a[rng.random((512, W)) < 0.02] = 0
        out.append((f"wide{W}", a))
    return out


# ---- hand-derived band goldens (tests/golden/make_golden.py, BANDS_NOTE) ----
BAND_CASES = ["band_plus_2048", "band_minus_2048", "band_pair_2048"]


def band_goldens():
    """the hand-derived band partitions of tests/golden/band_partitions.json"""
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "band_partitions.json")) as f:
        return json.load(f)["cases"]


def band_csr(n, bands):
    """the fixture's input: the union of bands (offset, first row, end row,
    rows r % 7 == 3 dropped or not), columns ascending within a row"""
    rows = [[] for _ in range(n)]
    for o, r0, r1, gaps in bands:
        for r in range(r0, r1):
            if not (gaps and r % 7 == 3):
                rows[r].append(r + o)
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    col = np.array([c for r in rows for c in sorted(r)], np.int32)
    return rp, col


def band_format_equals(fmt, want, rp, col):
    """fmt equals a band golden's format; the aliased case keeps the input
    CSR unencoded, the others hold the listed remainder entries"""
    for k in O.Format.FIELDS:
        assert getattr(fmt, k) == want[k], k
    for k in ("xStart", "yStart", "xStop", "nDiasPtr", "neg_offsets"):
        assert np.array_equal(np.asarray(getattr(fmt, k), np.int64).ravel(), np.array(want[k], np.int64)), k
    if want["aliased"]:
        assert np.array_equal(fmt.csrRow, rp) and np.array_equal(fmt.csrCol, col)
        return
    M = len(rp) - 1
    rem = np.array(want["remainder"], np.int64).reshape(-1, 2)
    crow = np.zeros(M + 1, np.int64)
    crow[1:] = np.cumsum(np.bincount(rem[:, 0], minlength=M))
    assert np.array_equal(fmt.csrRow, crow)
    assert np.array_equal(fmt.csrCol, rem[:, 1])
