#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/.

The reference ships no tests or fixtures (SURVEY: "the fixtures in the
reference's own tests are none"), so every expected value below is derived
by hand from the reference's source semantics and typed in literally; none
is computed by the oracle or the product.

  * *.mtx            Matrix Market inputs for readMtx (TaiChi.cu:120-278)
  * expected.json    CSR each reader must produce (row order = file order,
                     duplicates kept, symmetric mirrored, pattern -> 1.0),
                     or the return code (-1..-4)
  * diag_gaps_4096   a partition derived by hand: n=4096 identity with every
                     row i % 10 == 9 missing (derivation in PARTITION_NOTE)
  * band_partitions  three partitions derived by hand on n = 8192 that pin
                     the contour-corner transposition (a contour point's x,
                     a column, is used as the square's row start,
                     TaiChi.cu:368-371 with cal_shape_size :280-311): one
                     band at +2048, one at -2048 (both give no shape), and
                     the pair (each contour yields the other band's square:
                     two shapes of one dense offset each, the class the
                     reference runs with hybrid2, TaiChi.cu:1288-1290)
                     (derivation in BANDS_NOTE)
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))

MTX = {
    "general_real.mtx": (
        "%%MatrixMarket matrix coordinate real general\n"
        "% comment line\n"
        "%another comment\n"
        "4 5 6\n"
        "1 1 1.5\n"
        "3 2 -2.0\n"
        "1 5 3.25\n"
        "4 4 1e-3\n"
        "2 3 7\n"
        "3 1 0.5\n"),
    "symmetric_pattern.mtx": (
        "%%MatrixMarket matrix coordinate pattern symmetric\n"
        "3 3 4\n"
        "1 1\n"
        "2 1\n"
        "3 2\n"
        "3 3\n"),
    "integer_dups.mtx": (
        "%%MatrixMarket matrix coordinate integer general\n"
        "2 3 4\n"
        "1 2 5\n"
        "1 2 -3\n"
        "2 3 4\n"
        "1 1 2\n"),
    "skew.mtx": (
        "%%MatrixMarket matrix coordinate real skew-symmetric\n"
        "3 3 2\n"
        "2 1 1.0\n"
        "3 1 -4.0\n"),
    "mixed_case.mtx": (
        "%%MatrixMarket MATRIX Coordinate Pattern GENERAL\n"
        "2 2 1\n"
        "2 2\n"),
    "empty_rows.mtx": (
        "%%MatrixMarket matrix coordinate pattern general\n"
        "5 5 2\n"
        "5 1\n"
        "1 5\n"),
    "complex.mtx": (
        "%%MatrixMarket matrix coordinate complex general\n"
        "1 1 1\n"
        "1 1 1.0 2.0\n"),
    "bad_banner.mtx": (
        "%%MatrixMarkt matrix coordinate real general\n"
        "1 1 1\n"
        "1 1 1.0\n"),
    "bad_object.mtx": (
        "%%MatrixMarket tensor coordinate real general\n"
        "1 1 1\n"
        "1 1 1.0\n"),
    "short_banner.mtx": (
        "%%MatrixMarket matrix coordinate\n"
        "1 1 1\n"),
    "no_size.mtx": (
        "%%MatrixMarket matrix coordinate real general\n"
        "% only comments follow\n"),
    "array.mtx": (
        "%%MatrixMarket matrix array real general\n"
        "2 2\n"
        "1.0\n2.0\n3.0\n4.0\n"),
    "truncated.mtx": (
        "%%MatrixMarket matrix coordinate pattern general\n"
        "3 3 3\n"
        "1 1\n"
        "2 2\n"),
    "out_of_range.mtx": (
        "%%MatrixMarket matrix coordinate pattern general\n"
        "3 3 1\n"
        "5 1\n"),
}

# Hand-derived reader results. CSR rows keep file order (counting sort,
# TaiChi.cu:218-270); symmetric files add (c,r) right after (r,c) in the
# same pass (:240-253).
EXPECTED = {
    # rows: 0:(1,1)(1,5) 1:(2,3) 2:(3,2)(3,1) 3:(4,4)
    "general_real.mtx": {"rc": 0, "M": 4, "N": 5, "rowptr": [0, 2, 3, 5, 6],
                         "col": [0, 4, 2, 1, 0, 3], "val": [1.5, 3.25, 7.0, -2.0, 0.5, 0.001],
                         "typecode": "MCRG"},
    # (1,1) ; (2,1)+(1,2) ; (3,2)+(2,3) ; (3,3)
    "symmetric_pattern.mtx": {"rc": 0, "M": 3, "N": 3, "rowptr": [0, 2, 4, 6],
                              "col": [0, 1, 0, 2, 1, 2], "val": [1.0] * 6, "typecode": "MCPS"},
    # duplicates kept, integer -> float
    "integer_dups.mtx": {"rc": 0, "M": 2, "N": 3, "rowptr": [0, 3, 4],
                         "col": [1, 1, 0, 2], "val": [5.0, -3.0, 2.0, 4.0], "typecode": "MCIG"},
    # skew-symmetric is not mirrored (only 'S' and 'H' are, :162)
    "skew.mtx": {"rc": 0, "M": 3, "N": 3, "rowptr": [0, 0, 1, 2], "col": [0, 0],
                 "val": [1.0, -4.0], "typecode": "MCRK"},
    "mixed_case.mtx": {"rc": 0, "M": 2, "N": 2, "rowptr": [0, 0, 1], "col": [1], "val": [1.0],
                       "typecode": "MCPG"},
    "empty_rows.mtx": {"rc": 0, "M": 5, "N": 5, "rowptr": [0, 1, 1, 1, 1, 2], "col": [4, 0],
                       "val": [1.0, 1.0], "typecode": "MCPG"},
    "complex.mtx": {"rc": -3, "typecode": "MCCG"},
    "bad_banner.mtx": {"rc": -2, "banner_rc": 14},
    "bad_object.mtx": {"rc": -2, "banner_rc": 15},
    "short_banner.mtx": {"rc": -2, "banner_rc": 12},
    "no_size.mtx": {"rc": -4, "typecode": "MCRG"},
    # the reference would misread the two-number size line; rejected here
    "array.mtx": {"rc": -4, "typecode": "MARG"},
    # the reference reads past the data (undefined); rejected here
    "truncated.mtx": {"rc": -4, "typecode": "MCPG"},
    "out_of_range.mtx": {"rc": -4, "typecode": "MCPG"},
}

PARTITION_NOTE = """
diag_gaps_4096: M = N = 4096, A[i][i] = 1 unless i % 10 == 9.
thumbnail (morohutils.cu:191-200): pM = 4096/512 = 8, W = (int)(512.0/4096*4096) = 512,
  pN = 8; pixel (i/8, i/8) is hit for every i (each 8-row block holds a
  nonzero), so the image is a 512-pixel diagonal line, 0 on the line.
morphology: no holes; the 30x1 opening leaves nothing of a 1-pixel line,
  so rev - open = rev; eye(5) opening keeps the whole line (outside pixels
  read 255 for the erosion at both ends).
contours: one contour of 2*512-2 = 1022 points, contour[0] = (0,0),
  contour[511] = (511,511).
cal_shape_size: gain = 4096/512 = 8.0f, x = y = 0, a = b = 511*8 = 4088.
partition_dia_shape: sd = 4088 -> width = (int)(4088*0.1) = 408;
  offset 0 has rows [0,4088) minus the 408 rows 9,19,...,4079 -> 3680 of
  4088 = 0.90 >= 0.6 -> dense; every other offset is empty.
  zero rows: 9,19,...,4079 (408) -> zero entries (i<<1).
remainder: rows 4088..4095 keep their diagonal except 4089 -> 7 entries
  ((i<<1)|1).  nnz = 4096 - 409 = 3687, dense_nnz = 3680, sparse = 7.
"""


def diag_gaps():
    n = 4096
    rows = [i for i in range(n) if i % 10 != 9]
    rowptr, col = [0], []
    for i in range(n):
        if i % 10 != 9:
            col.append(i)
        rowptr.append(len(col))
    zero_rows = [i for i in range(4088) if i % 10 == 9]
    rem_rows = [i for i in range(4088, 4096) if i % 10 != 9]
    crow, ccol = [0], []
    for i in range(n):
        if i in zero_rows:
            ccol.append(i << 1)
        if i in rem_rows:
            ccol.append((i << 1) | 1)
        crow.append(len(ccol))
    assert len(rows) == 3687
    return {
        "M": n, "N": n, "rowptr": rowptr, "col": col,
        "thumb_line": True,
        "contour_len": 1022, "contour_first": [0, 0], "contour_mid": [511, 511],
        "shape": {"x": 0, "y": 0, "a": 4088, "b": 4088, "gain": 8.0, "width": 408},
        "format": {"dia_shapes": 1, "num_negs": 1, "dense_nnz": 3680, "zero_elements": 408,
                   "sparse_nnz": 7, "row_start": 0, "row_stop": 4088, "xStart": [0],
                   "yStart": [0], "xStop": [4088], "nDiasPtr": [0, 1], "neg_offsets": [0],
                   "csrRow": crow, "csrCol": ccol},
        "note": PARTITION_NOTE,
    }


BANDS_NOTE = """
Common to all three: M = N = 8192. Thumbnail (morohutils.cu:191-200): pM =
8192/512 = 16, W = (int)((512.0/8192)*8192) = 512, pN = 16. Both band
offsets (+-2048) are multiples of 16, so a nonzero (i, i+o) lands on pixel
(i/16, i/16 + o/16): a one-pixel diagonal line.

band_plus_2048: A[i][i+2048] = 1 for i in [0, 6144) (6144 nonzeros).
  image: pixels (p, p+128), p in [0, 384): a line from (row 0, col 128) to
  (row 383, col 511). Morphology: no hole; the 30-long openings erase a
  one-pixel line; the eye(5) opening keeps every pixel (its ends read the
  border, 255, outside the image). One contour of 2*384-2 = 766 points
  (>= 80, kept): contour[0] = Point(x=128, y=0), contour[383] = (511, 383).
  partition_shapes (TaiChi.cu:368-371): x1 = 128 (the column), y1 = 0,
  a1 = 511-128 = 383, b1 = 383-0 = 383. cal_shape_size: gain = 8192/512 =
  16.0f, x = 2048, y = 0, a = b = 383*16 = 6128 (no clamp: 0+6128 <= N,
  2048+6128 = 8176 <= M). The square: rows [2048, 8176), cols [0, 6128).
  partition_dia_shape: sd = 6128 < 1e4 -> width = (int)(6128*0.1) = 612,
  candidate offsets [-611, 611]. The band's entries inside the square are
  rows [2048, 4080) at column r+2048, local offset (r+2048-0)-(r-2048) =
  4096: no candidate holds an entry, no dense diagonal, nnz = 0: the shape
  is dropped. S = 0, so dense_nnz = 0 and the format aliases the input CSR
  (TaiChi.cu:575-578); row_start = row_stop = 0 (the caller's zeroed struct,
  main.cu:40-41; :458-459 run only when S > 0). (Had x been the row, the
  square would be rows [0, 6128), cols [2048, 8176), the band at offset 0:
  S = 1.)

band_minus_2048: A[i][i-2048] = 1 for i in [2048, 8192) (6144 nonzeros).
  image: pixels (p, p-128), p in [128, 512): contour[0] = (x=0, y=128),
  contour[383] = (383, 511), 766 points. x1 = 0, y1 = 128, a1 = b1 = 383:
  x = 0, y = 2048, a = b = 6128; square rows [0, 6128), cols [2048, 8176).
  The band inside it: rows [4096, 6128) at column r-2048, local offset
  (r-2048-2048)-r = -4096, beyond 611: S = 0, aliased CSR as above.

band_pair_2048: both bands, except that the +2048 band has no entry in
  rows r % 7 == 3 (878 of its 6144 rows: r = 3+7t, t = 0..877 below 6144;
  5266 entries) -> nnz = 5266 + 6144 = 11410. Every 16-row block still holds
  an entry of each band, so the image is the two lines above. Two contours;
  the output order lists the later-discovered border first (the -2048 line,
  first met at row 128, before the +2048 line, met at row 0); this order is
  the restatement's reading of OpenCV and is not pinned against it.
  shape 0 (the -2048 line's corners): x = 0, y = 2048, side 6128, rows
    [0, 6128), cols [2048, 8176). The +2048 band lies on its local offset
    0: rows [0, 6128) minus the 875 rows r = 3+7t < 6128 (t = 0..874):
    5253 / 6128 = 0.857 >= 0.6: dense, offset 0; the 875 missing rows are
    its zero cells. The -2048 band sits at local offset -4096: not a
    candidate.
  shape 1 (the +2048 line's corners): x = 2048, y = 0, rows [2048, 8176),
    cols [0, 6128). The -2048 band lies on its local offset 0 in every row:
    6128 / 6128: dense, no zero cell. The +2048 band is at 4096.
  Format: dia_shapes 2, num_negs 2, xStart [0, 2048], yStart [2048, 0],
  xStop [6128, 8176], nDiasPtr [0, 1, 2], neg_offsets [0, 0], row_start 0,
  row_stop 8176, dense_nnz 5253 + 6128 = 11381, zero_elements 875,
  sparse_nnz 11410 - 11381 = 29: the +2048 band in rows [6128, 6144)
  except 6128, 6135, 6142 (13; 6128 = 7*875 + 3) and the -2048 band in
  rows [8176, 8192) (16). Remainder, per row zeros first then sparse
  (TaiChi.cu:585-596): row r = 3+7t < 6128 holds (r+2048) << 1; the 13
  rows hold ((r+2048) << 1) | 1; rows [8176, 8192) hold ((r-2048) << 1) | 1.
"""


def _band_case(n, bands):
    """CSR of the union of bands (offset, first row, end row, rows dropped
    by r % 7 == 3 or not), columns ascending within a row"""
    rows = [[] for _ in range(n)]
    for o, r0, r1, gaps in bands:
        for r in range(r0, r1):
            if not (gaps and r % 7 == 3):
                rows[r].append(r + o)
    rowptr, col = [0], []
    for r in range(n):
        col.extend(sorted(rows[r]))
        rowptr.append(len(col))
    return rowptr, col


def band_partitions():
    n = 8192
    plus = (2048, 0, 6144, False)
    minus = (-2048, 2048, 8192, False)
    plus_gaps = (2048, 0, 6144, True)
    empty = {"dia_shapes": 0, "num_negs": 0, "dense_nnz": 0, "zero_elements": 0, "sparse_nnz": 6144,
             "row_start": 0, "row_stop": 0, "xStart": [], "yStart": [], "xStop": [], "nDiasPtr": [],
             "neg_offsets": [], "aliased": True}
    cases = {}
    rp, col = _band_case(n, [plus])
    assert len(col) == 6144
    cases["band_plus_2048"] = {
        "M": n, "N": n, "bands": [list(plus)], "contours": [{"len": 766, "first": [128, 0], "mid": [511, 383]}],
        "shapes": [{"x": 2048, "y": 0, "a": 6128, "b": 6128, "gain": 16.0, "width": 612}],
        "format": dict(empty)}
    rp, col = _band_case(n, [minus])
    assert len(col) == 6144
    cases["band_minus_2048"] = {
        "M": n, "N": n, "bands": [list(minus)], "contours": [{"len": 766, "first": [0, 128], "mid": [383, 511]}],
        "shapes": [{"x": 0, "y": 2048, "a": 6128, "b": 6128, "gain": 16.0, "width": 612}],
        "format": dict(empty)}
    # the pair: remainder entries (row, encoded column), rows ascending
    zero_rows = [3 + 7 * t for t in range(875)]
    assert zero_rows[-1] < 6128 <= 3 + 7 * 875
    rem = [[r, (r + 2048) << 1] for r in zero_rows]
    rem += [[r, ((r + 2048) << 1) | 1] for r in range(6128, 6144) if r % 7 != 3]
    rem += [[r, ((r - 2048) << 1) | 1] for r in range(8176, 8192)]
    assert len(rem) == 875 + 13 + 16
    rp, col = _band_case(n, [plus_gaps, minus])
    assert len(col) == 11410
    cases["band_pair_2048"] = {
        "M": n, "N": n, "bands": [list(plus_gaps), list(minus)],
        "contours": [{"len": 766, "first": [0, 128], "mid": [383, 511]},
                     {"len": 766, "first": [128, 0], "mid": [511, 383]}],
        "shapes": [{"x": 0, "y": 2048, "a": 6128, "b": 6128, "gain": 16.0, "width": 612},
                   {"x": 2048, "y": 0, "a": 6128, "b": 6128, "gain": 16.0, "width": 612}],
        "format": {"dia_shapes": 2, "num_negs": 2, "dense_nnz": 11381, "zero_elements": 875,
                   "sparse_nnz": 29, "row_start": 0, "row_stop": 8176, "xStart": [0, 2048],
                   "yStart": [2048, 0], "xStop": [6128, 8176], "nDiasPtr": [0, 1, 2], "neg_offsets": [0, 0],
                   "aliased": False, "remainder": rem}}
    return {"cases": cases, "note": BANDS_NOTE}


def main():
    for name, text in MTX.items():
        with open(os.path.join(HERE, name), "w") as f:
            f.write(text)
    with open(os.path.join(HERE, "expected.json"), "w") as f:
        json.dump({"mtx": EXPECTED}, f, indent=1, sort_keys=True)
    with open(os.path.join(HERE, "diag_gaps_4096.json"), "w") as f:
        json.dump(diag_gaps(), f)
    with open(os.path.join(HERE, "band_partitions.json"), "w") as f:
        json.dump(band_partitions(), f)


if __name__ == "__main__":
    main()
