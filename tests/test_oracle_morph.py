"""a3 -- morphology and contours. Each expectation is derived by hand from
OpenCV's documented semantics (anchor ksize/2, no reflection for dilate,
255/0 constant border for erode/dilate, Suzuki-Abe border following on the
zero-padded image); the product's implementation is then checked against
the oracle on the same and on random images."""
import numpy as np
import pytest

import oracle as O
import taichi_amd as T

from helpers import morph_images

H30 = (np.arange(30) - 15, np.zeros(30, int))


def img(h, w):
    return np.zeros((h, w), np.uint8)


def runs(row):
    """[(start, end_inclusive)] of 255 runs in a 1-D array."""
    out, s = [], None
    for i, v in enumerate(list(row) + [0]):
        if v and s is None:
            s = i
        if not v and s is not None:
            out.append((s, i - 1))
            s = None
    return out


def test_open30_shifts_a_30_run_right_by_one():
    # window [x-15, x+14]: erosion keeps only x = a+15, dilation then covers
    # [a+1, a+30] -- the output is the input shifted one pixel right
    a = img(1, 100)
    a[0, 40:70] = 255
    assert runs(O.open_(a, 0)[0]) == [(41, 70)]


def test_open30_removes_a_29_run():
    a = img(1, 100)
    a[0, 40:69] = 255
    assert runs(O.open_(a, 0)[0]) == []


def test_open30_left_border_reads_255_for_erosion():
    # run [0,20]: erosion keeps x in [0,6] (outside pixels are 255); the
    # dilation then covers [0, 21]
    a = img(1, 60)
    a[0, 0:21] = 255
    assert runs(O.open_(a, 0)[0]) == [(0, 21)]


def test_open_vertical_is_transposed_horizontal():
    a = img(100, 3)
    a[40:70, 1] = 255
    assert runs(O.open_(a, 1)[:, 1]) == [(41, 70)]


def test_eye5_open_keeps_5_removes_4():
    a = img(40, 40)
    for i in range(5):
        a[5 + i, 5 + i] = 255
    for i in range(4):
        a[20 + i, 30 + i] = 255
    out = O.open_(a, 2)
    assert set(zip(*np.nonzero(out))) == {(5 + i, 5 + i) for i in range(5)}


def test_erode_dilate_unreflected_offsets():
    a = img(1, 10)
    a[0, 5] = 255
    # dilate with element offsets {0, +1}: dst(x) = max(src(x), src(x+1))
    out = O.dilate(a, [0, 1], [0, 0])
    assert runs(out[0]) == [(4, 5)]


def test_imfill_fills_holes_only():
    ring = img(10, 10)
    ring[2:7, 2:7] = 255
    ring[3:6, 3:6] = 0
    out = O.imfill(ring)
    assert np.all(out[2:7, 2:7] == 255) and out.sum() == 25 * 255
    cshape = ring.copy()
    cshape[4, 6] = 0  # opening in the right wall: the inside touches the border
    assert np.array_equal(O.imfill(cshape), cshape)


@pytest.mark.parametrize("L", [5, 10, 50, 511])
def test_contour_of_diagonal_line(L):
    a = img(L + 4, L + 4)
    for i in range(L):
        a[2 + i, 1 + i] = 255
    cs = O.find_contours(a)
    assert len(cs) == 1
    pts, hole = cs[0]
    assert not hole and len(pts) == 2 * L - 2
    assert pts[0].tolist() == [1, 2]
    assert pts[L - 1].tolist() == [L, L + 1]          # far end at index L-1
    assert pts[-1].tolist() == [2, 3]                 # last point before closing


def test_contour_start_and_direction_of_rectangle():
    # 2x3 block: start at its top-left pixel, follow the left side downwards
    a = img(8, 8)
    a[2:5, 3:5] = 255
    pts, _ = O.find_contours(a)[0]
    assert pts[:4].tolist() == [[3, 2], [3, 3], [3, 4], [4, 4]]
    assert len(pts) == 6


def test_contour_order_is_reverse_raster():
    a = img(20, 20)
    a[1, 1] = 255          # found first
    a[10, 5:8] = 255       # found second
    a[10, 15] = 255        # found third
    cs = O.find_contours(a)
    assert [p[0].tolist() for p, _ in cs] == [[15, 10], [5, 10], [1, 1]]
    assert [len(p) for p, _ in cs] == [1, 4, 1]


def test_hole_follows_its_outer_contour():
    a = img(12, 20)
    a[2:7, 2:7] = 255
    a[4, 4] = 0            # hole in the first square
    a[2:5, 12:15] = 255    # second square, found after the first outer
    cs = O.find_contours(a)
    kinds = [(p[0].tolist(), h) for p, h in cs]
    # outer contours in reverse discovery order; the hole right after its parent
    assert kinds == [([12, 2], False), ([2, 2], False), ([3, 4], True)]
    hole = cs[2][0]
    assert sorted(map(tuple, hole.tolist())) == [(3, 4), (4, 3), (4, 5), (5, 4)]


def _product_contours(a):
    return [(p.tolist(), h) for p, h in T.contours(a)]


def _oracle_contours(a):
    return [(p.tolist(), h) for p, h in O.find_contours(a)]


def test_product_contours_match_oracle_on_hand_images():
    cases = []
    a = img(30, 30)
    for i in range(20):
        a[3 + i, 2 + i] = 255
    cases.append(a)
    b = img(12, 20)
    b[2:7, 2:7] = 255
    b[4, 4] = 0
    b[2:5, 12:15] = 255
    cases.append(b)
    c = img(16, 16)
    c[1:15, 1:15] = 255
    c[3:6, 3:6] = 0
    c[8:12, 8:12] = 0
    c[9:11, 9:11] = 255     # island inside a hole
    cases.append(c)
    for a in cases:
        assert _product_contours(a) == _oracle_contours(a)


@pytest.mark.parametrize("seed", range(12))
def test_product_contours_match_oracle_random(seed):
    rng = np.random.default_rng(seed)
    h, w = rng.integers(5, 60, 2)
    a = ((rng.random((h, w)) < rng.uniform(0.2, 0.7)) * 255).astype(np.uint8)
    assert _product_contours(a) == _oracle_contours(a)


@pytest.mark.parametrize("seed", range(8))
def test_product_sdf_matches_oracle_random(seed):
    """Whole getSdf image pipeline on random thumbnails with diagonal bands."""
    rng = np.random.default_rng(100 + seed)
    H, W = 512, int(rng.integers(300, 513))
    a = np.full((H, W), 255, np.uint8)
    for _ in range(int(rng.integers(1, 6))):  # diagonal segments (0 = hit)
        r0, c0, ln = rng.integers(0, 400), rng.integers(0, W - 50), rng.integers(60, 300)
        for t in range(int(ln)):
            if r0 + t < H and c0 + t < W:
                a[r0 + t, c0 + t] = 0
    noise = rng.random((H, W)) < rng.uniform(0.0, 0.05)
    a[noise] = 0
    if rng.random() < 0.5:
        a[100:160, 200:260] = 0  # a dense block for the rectangle removal
    assert np.array_equal(T.sdf_from_image(a), O.sdf_from_image(a))


def test_sdf_of_single_diagonal_thumbnail():
    """The thumbnail of diag_gaps_4096: a 512-pixel line -> one shape
    (0,0)-(511,511) (derivation in tests/golden/make_golden.py)."""
    a = np.full((512, 512), 255, np.uint8)
    for i in range(512):
        a[i, i] = 0
    for sdf in (O.sdf_from_image(a), T.sdf_from_image(a)):
        assert sdf.tolist() == [[0, 0, 511, 511]]


@pytest.mark.parametrize("case", morph_images(), ids=lambda c: c[0])
def test_product_morph_matches_oracle_steps(case):
    """tc_morph_from_image equals the oracle's steps composed as getSdf
    composes them (morohutils.cu:231-251), pixel for pixel."""
    _, a = case
    rev = (255 - a.astype(np.int32)).astype(np.uint8)
    o2 = O.open_(O.open_(O.imfill(rev), 0), 1)
    dele = np.where(rev > o2, rev.astype(np.int32) - o2, 0).astype(np.uint8)
    want = O.open_(dele, 2)
    assert np.array_equal(T.morph_from_image(a), want)
