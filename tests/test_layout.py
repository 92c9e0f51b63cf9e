"""Device layout sizes, host only (tc_matrix_footprint needs no device):
term-presence bitmaps replace the sign-0 entries of banded matrices with
many zero cells on their dense diagonals, and leave matrices whose dense
diagonals have few zero cells (R-MAT) as they were."""
import numpy as np

import taichi_amd as T

from helpers import host_format, rmat


def _entries_bytes(fmt):
    return 4 * int(fmt.csrRow[-1])


def test_banded_bitmaps_shrink_the_layout():
    for w in (16, 40, 64):
        n = 20000
        rp, col = T.gen_banded(n, w, 0.7, 3 + w)
        fmt = host_format(n, n, rp, col)
        R = int(fmt.csrRow[-1])
        zeros = int(np.count_nonzero((np.asarray(fmt.csrCol[:R]) & 1) == 0))
        assert zeros > 4 * n  # many cancelled zero cells a row
        # without bitmaps cols alone would be 4 * R bytes
        assert T.footprint(fmt) < _entries_bytes(fmt), w


def test_rmat_keeps_its_entries():
    M, N, rp, col = rmat(14)
    fmt = host_format(M, N, rp, col)
    assert T.footprint(fmt) >= _entries_bytes(fmt)
