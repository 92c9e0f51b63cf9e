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


def main():
    for name, text in MTX.items():
        with open(os.path.join(HERE, name), "w") as f:
            f.write(text)
    with open(os.path.join(HERE, "expected.json"), "w") as f:
        json.dump({"mtx": EXPECTED}, f, indent=1, sort_keys=True)
    with open(os.path.join(HERE, "diag_gaps_4096.json"), "w") as f:
        json.dump(diag_gaps(), f)


if __name__ == "__main__":
    main()
