"""Pins for the oracle's thumbnail, partition arithmetic and format
(morohutils.cu:153-200; TaiChi.cu:280-311, 313-638, 677-916), its CSR SpMV
(main.cu:61-69) and format interpreter, checked against hand-derived tables,
brute force and invariants; then the product host partition against the
oracle on the same inputs. CPU only."""
import json
import os

import numpy as np
import pytest

import oracle as O
import taichi_amd as T

from helpers import BAND_CASES, band_csr, band_format_equals, band_goldens

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---- width tiers (TaiChi.cu:689-699): hand table at every tier edge ----
@pytest.mark.parametrize("sd,w", [
    (9, 0), (10, 1), (4088, 408), (9999, 999),          # x 0.1
    (10000, 100), (16352, 163), (99999, 999),            # x 0.01
    (100000, 100), (516096, 516), (999999, 999),         # x 0.001
    (1000000, 10), (4186112, 41), (9999999, 99),         # x 1e-5
    (10000000, 10), (40000000, 40)])                     # x 1e-6
def test_width_tiers(sd, w):
    assert O.width_for(sd) == w


# ---- cal_start_end (TaiChi.cu:881-916) equals the rows a diagonal covers ----
@pytest.mark.parametrize("Ms,Ns,width", [(10, 10, 1), (50, 50, 5), (4088, 4088, 408),
                                         (300, 200, 30), (200, 300, 20), (64, 63, 6)])
def test_cal_start_end_is_the_diagonal_row_range(Ms, Ns, width):
    for off in range(-(width - 1), width):
        n = off + width - 1
        s, e = O.cal_start_end(Ms, Ns, width, off, n)
        rows = [r for r in range(Ms) if 0 <= r + off < Ns]   # brute force
        assert (s, e) == (rows[0], rows[-1] + 1), (off, s, e)


# ---- cal_shape_size (TaiChi.cu:280-311): float gain, clamps, square ----
@pytest.mark.parametrize("args,want", [
    # gain = 8192.0f; x,y,a,b scaled; a = b = min(a, b)
    ((4194304, 4194304, 512, 3, 5, 10, 12), (24576, 40960, 81920, 81920, 8192.0)),
    # the golden diagonal: gain 8.0f, 511 * 8 = 4088
    ((4096, 4096, 512, 0, 0, 511, 511), (0, 0, 4088, 4088, 8.0)),
    # (float)16777219 = 16777220, gain = 32768.0078125f; 511 * gain rounds
    # up to 16744452 in float (16744450 in double); then x + b > M clamps b
    ((16777219, 16777219, 512, 511, 0, 1, 1), (16744452, 0, 32767, 32767, 32768.0078125)),
    # gain 19.53125: x = 9765, y = 195, a = b = 1953, x + b > M -> b = 235
    ((10000, 10000, 512, 500, 10, 100, 100), (9765, 195, 235, 235, 19.53125)),
    # M < mF: divide by gain = 2.0
    ((256, 300, 512, 100, 40, 64, 30), (50, 20, 15, 15, 2.0)),
    # M == mF: unchanged, then the square
    ((512, 512, 512, 1, 2, 9, 7), (1, 2, 7, 7, 1.0)),
])
def test_cal_shape_size_table(args, want):
    got = O.cal_shape_size(*args)
    assert got[:4] == want[:4]
    assert np.float32(got[4]) == np.float32(want[4])


# ---- thumbnail (morohutils.cu:191-200, 153-180) ----
@pytest.mark.parametrize("M,N,want", [
    (4096, 4096, (0, 512, 512, 8, 8)),
    (13001, 13001, (0, 512, 512, 25, 25)),
    (561, 561, (0, 512, 511, 1, 1)),       # (int)(512.0/561*561) = 511
    (3000, 9000, (0, 512, 1536, 5, 5)),
    (512, 512, (1, None, None, None, None)),
])
def test_thumb_dims_closed_form(M, N, want):
    got = O.thumb_dims(M, N)
    assert got[0] == want[0]
    if want[0] == 0:
        assert got[1:] == want[1:]


@pytest.mark.parametrize("M,N,seed", [(4096, 4096, 0), (13001, 13001, 1), (561, 561, 2), (3000, 9000, 3)])
def test_thumbnail_brute_force(M, N, seed):
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 6, M)
    rp = np.zeros(M + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    col = rng.integers(0, N, int(rp[-1])).astype(np.int32)
    _, H, W, pM, pN = O.thumb_dims(M, N)
    want = np.full((H, W), 255, np.uint8)
    for i in range(M):
        for c in col[rp[i]:rp[i + 1]]:
            if i // pM < H and c // pN < W:
                want[i // pM, c // pN] = 0
    assert np.array_equal(O.thumbnail(M, N, rp, col), want)


# ---- CSR SpMV (main.cu:61-69): sequential FP32 in entry order ----
def test_csr_spmv_f32_is_sequential_float32():
    rng = np.random.default_rng(4)
    M = N = 300
    lens = rng.integers(0, 40, M)
    rp = np.zeros(M + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    col = rng.integers(0, N, int(rp[-1])).astype(np.int32)
    x = (rng.standard_normal(N) * 10 ** rng.uniform(-3, 3, N)).astype(np.float32)
    want = np.empty(M, np.float32)
    for i in range(M):
        s = np.float32(0)
        for j in range(rp[i], rp[i + 1]):
            s = np.float32(s + x[col[j]])
        want[i] = s
    assert np.array_equal(O.csr_spmv_f32(rp, col, x), want)
    # reversing a row's entries changes some sums: the order is pinned
    col_r = col.copy()
    for i in range(M):
        col_r[rp[i]:rp[i + 1]] = col[rp[i]:rp[i + 1]][::-1]
    assert not np.array_equal(O.csr_spmv_f32(rp, col_r, x), want)
    # the multi-threaded baseline splits rows, not sums: the same y
    for threads in (1, 3, 8):
        assert np.array_equal(O.csr_spmv_f32_mt(rp, col, x, threads), want)


# ---- golden partition derived by hand (tests/golden/make_golden.py) ----
def _gold():
    with open(os.path.join(GOLD, "diag_gaps_4096.json")) as f:
        return json.load(f)


def _format_equals(fmt, want):
    for k in O.Format.FIELDS:
        assert getattr(fmt, k) == want[k], k
    for k in ("xStart", "yStart", "xStop", "nDiasPtr", "neg_offsets", "csrRow", "csrCol"):
        assert np.array_equal(getattr(fmt, k), np.array(want[k], np.int32)), k


def test_golden_diag_gaps_pipeline():
    g = _gold()
    M, N = g["M"], g["N"]
    rp, col = np.array(g["rowptr"], np.int32), np.array(g["col"], np.int32)
    img = O.thumbnail(M, N, rp, col)
    line = np.full((512, 512), 255, np.uint8)
    line[np.arange(512), np.arange(512)] = 0
    assert np.array_equal(img, line)
    cs = O.find_contours(255 - img)
    assert len(cs) == 1 and len(cs[0][0]) == g["contour_len"]
    assert list(cs[0][0][0]) == g["contour_first"]
    assert list(cs[0][0][len(cs[0][0]) // 2]) == g["contour_mid"]
    sh = g["shape"]
    x, y, a, b, gain = O.cal_shape_size(M, N, 512, 0, 0, 511, 511)
    assert (x, y, a, b, gain) == (sh["x"], sh["y"], sh["a"], sh["b"], sh["gain"])
    assert O.width_for(min(a, b)) == sh["width"]
    fmt, _ = O.partition(M, N, rp, col)
    _format_equals(fmt, g["format"])


def test_golden_diag_gaps_product_partition():
    g = _gold()
    rp, col = np.array(g["rowptr"], np.int32), np.array(g["col"], np.int32)
    fmt, _ = T.partition_from_image(g["M"], g["N"], rp, col, img=O.thumbnail(g["M"], g["N"], rp, col))
    _format_equals(fmt, g["format"])


# ---- band goldens pinning the contour-corner transposition ----
# (hand derivations in tests/golden/make_golden.py, BANDS_NOTE)


@pytest.mark.parametrize("name", BAND_CASES)
def test_golden_bands_oracle_pipeline(name):
    g = band_goldens()[name]
    M, N = g["M"], g["N"]
    rp, col = band_csr(M, g["bands"])
    img = O.thumbnail(M, N, rp, col)
    cs = O.find_contours(255 - img)
    assert len(cs) == len(g["contours"])
    for (pts, _), want, sh in zip(cs, g["contours"], g["shapes"]):
        assert len(pts) == want["len"]
        assert list(pts[0]) == want["first"] and list(pts[len(pts) // 2]) == want["mid"]
        # the corners' x (a column) is the square's row start: x1 = P0.x
        x1, y1 = want["first"]
        a1, b1 = want["mid"][0] - x1, want["mid"][1] - y1
        assert O.cal_shape_size(M, N, 512, x1, y1, a1, b1) == (sh["x"], sh["y"], sh["a"], sh["b"], sh["gain"])
        assert O.width_for(min(sh["a"], sh["b"])) == sh["width"]
    fmt, _ = O.partition(M, N, rp, col)
    band_format_equals(fmt, g["format"], rp, col)


@pytest.mark.parametrize("name", BAND_CASES)
def test_golden_bands_product_partition(name):
    g = band_goldens()[name]
    rp, col = band_csr(g["M"], g["bands"])
    fmt, _ = T.partition_from_image(g["M"], g["N"], rp, col, img=O.thumbnail(g["M"], g["N"], rp, col))
    band_format_equals(fmt, g["format"], rp, col)


def test_golden_band_pair_interpreter_equals_csr():
    """On the two-shape golden the format's meaning (each shape's dense
    offset plus the signed remainder) gives the CSR sums exactly."""
    g = band_goldens()["band_pair_2048"]
    rp, col = band_csr(g["M"], g["bands"])
    fmt, _ = O.partition(g["M"], g["N"], rp, col)
    x = np.random.default_rng(4).integers(0, 8, g["N"]).astype(np.float32)
    assert np.array_equal(O.taichi_spmv(fmt, x, np.float32), O.csr_spmv_f32(rp, col, x))


# ---- the 0.6 density rule (TaiChi.cu:756-757) on both sides of it ----
@pytest.mark.parametrize("drop,dense", [(0, True), (1, False)])
def test_density_threshold_exactly_three_fifths(drop, dense):
    """M = N = 5120: the diagonal line image gives one square of side
    511 * 10 = 5110. Keeping the rows r % 5 in {0,1,2} of [0, 5110) gives
    3066 / 5110 = 0.6 exactly (float 0.6f >= 0.6): dense. One row fewer is
    below: no dense diagonal, the shape is dropped (S = 0)."""
    M = N = 5120
    keep = [r for r in range(M) if r >= 5110 or r % 5 in (0, 1, 2)]
    if drop:
        keep.remove(5)
    rp = np.zeros(M + 1, np.int32)
    for r in keep:
        rp[r + 1] = 1
    rp = np.cumsum(rp).astype(np.int32)
    col = np.array(keep, np.int32)
    full = np.arange(M + 1, dtype=np.int32)
    img = O.thumbnail(M, N, full, np.arange(M, dtype=np.int32))   # the diagonal line
    assert np.array_equal(img, O.thumbnail(M, N, rp, col))        # unchanged by the gaps
    for fmt in (O.partition(M, N, rp, col, img=img)[0],
                T.partition_from_image(M, N, rp, col, img=img)[0]):
        if dense:
            assert (fmt.dia_shapes, fmt.num_negs, fmt.dense_nnz) == (1, 1, 3066)
            assert fmt.zero_elements == 5110 - 3066 and fmt.sparse_nnz == 10
        else:
            assert (fmt.dia_shapes, fmt.dense_nnz) == (0, 0)
            assert fmt.sparse_nnz == len(keep)


# ---- format invariants and the interpreter, oracle vs product partition ----
def _cases():
    from taichi_amd import workloads as W
    out = [("cfg1", *W.diag5_random())]
    for s in (14, 16):
        out.append((f"rmat{s}", *W.rmat(1 << s, 16, 40 + s)))
    n = 60000
    rp, col = T.gen_banded(n, 17, 0.7, 9)
    out.append(("banded", n, n, rp, col))
    rp, col = T.gen_stencil5(200, 200, 0.02, 2)
    out.append(("stencil", 40000, 40000, rp, col))
    return out


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_format_invariants_and_interpreter(case):
    name, M, N, rp, col = case
    fmt, _ = O.partition(M, N, rp, col)
    nnz = len(col)
    assert fmt.dense_nnz + fmt.sparse_nnz == nnz
    R = int(fmt.csrRow[-1])
    assert R == fmt.zero_elements + fmt.sparse_nnz
    if fmt.dense_nnz > 0:
        signs = fmt.csrCol & 1
        assert int((signs == 0).sum()) == fmt.zero_elements
        assert int((signs == 1).sum()) == fmt.sparse_nnz
        # every dense diagonal's cells are its nonzeros plus its zero entries
        cells = 0
        for s in range(fmt.dia_shapes):
            side = int(fmt.xStop[s] - fmt.xStart[s])
            for o in fmt.neg_offsets[fmt.nDiasPtr[s]:fmt.nDiasPtr[s + 1]]:
                cells += sum(1 for r in range(side) if 0 <= r + o < side) if side < 3000 else \
                    side - abs(int(o))
        assert cells == fmt.dense_nnz + fmt.zero_elements
    x = T.gen_vector(N, 1, 17)
    assert np.array_equal(O.taichi_spmv(fmt, x, np.float32), O.csr_spmv_f32(rp, col, x))
    assert np.array_equal(O.taichi_spmv(fmt, x.astype(np.float64)), O.csr_spmv_f64(rp, col, x.astype(np.float64)))
    # the product's host partition builds the identical format
    pf, _ = T.partition_from_image(M, N, rp, col, img=O.thumbnail(M, N, rp, col))
    for k in O.Format.FIELDS:
        assert getattr(pf, k) == getattr(fmt, k), k
    for k in ("xStart", "yStart", "xStop", "nDiasPtr", "neg_offsets", "csrRow", "csrCol"):
        assert np.array_equal(getattr(pf, k), getattr(fmt, k)), k


def test_cfg1_band_count_closed_form():
    """5 diagonals (offsets -2..2) of n = 16384 hold 5n - 2*(1+2) = 81914
    cells; the 16384 uniform draws add the ones not on the band, deduped."""
    n = 16384
    rp, col = T.gen_diagonals_random(n, [-2, -1, 0, 1, 2], 0, 0)
    assert len(col) == 81914
    rows = np.repeat(np.arange(n), np.diff(rp))
    assert set(np.unique(col - rows).tolist()) == {-2, -1, 0, 1, 2}
    rp2, col2 = T.gen_diagonals_random(n, [-2, -1, 0, 1, 2], 16384, 0)
    rows2 = np.repeat(np.arange(n), np.diff(rp2))
    keys = rows2.astype(np.int64) * n + col2
    assert len(np.unique(keys)) == len(keys)                 # deduplicated
    assert 81914 < len(col2) <= 81914 + 16384
