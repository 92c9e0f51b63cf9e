"""Seeded synthetic workloads for the benchmark configurations.

The reference reads downloaded matrices (README.md:17-20); offline, these
generators build the same shapes of problem from the library's counter-based
generators (identical on every machine and thread count):

  diag5_random   n=16,384, ones on offsets -2..2 plus 16,384 uniform draws
  rmat           binary R-MAT, (a,b,c,d)=(.499,.001,.001,.499), self loops
                 kept, duplicates removed; n not a power of two uses scale
                 ceil(log2 n) with out-of-range endpoints rejected (not redrawn)
  stencil5       2-D 5-point stencil plus a fraction of uniform extra ones
  uneven_rmat    256 R-MAT matrices, seeds 100..355, n = round(2^U[12,18]),
                 edge factor uniform in {4..32}
  serving_mix    64 matrices: 32 banded (w uniform in {3..64} centred
                 diagonals, n uniform in [2^16, 2^20], density 0.7) and
                 32 R-MAT (n uniform in [2^16, 2^20], edge factor 16)

The per-matrix draws (n, edge factor, band width) use numpy's PCG64 seeded
as stated in each function.
"""
import math

import numpy as np

from . import gen_banded, gen_diagonals_random, gen_rmat, gen_stencil5

RMAT_ABC = (0.499, 0.001, 0.001)


def diag5_random(n=16384, draws=16384, seed=0):
    rp, col = gen_diagonals_random(n, [-2, -1, 0, 1, 2], draws, seed)
    return n, n, rp, col


def rmat(n, edge_factor=16, seed=1):
    scale = max(1, math.ceil(math.log2(n)))
    rp, col = gen_rmat(n, scale, edge_factor * n, *RMAT_ABC, seed)
    return n, n, rp, col


def stencil5(g=2048, extra=0.02, seed=2):
    rp, col = gen_stencil5(g, g, extra, seed)
    return g * g, g * g, rp, col


def uneven_rmat(count=256, first_seed=100, nmin_log2=12, nmax_log2=18):
    """[(n, n, rowptr, col, seed)] for seeds first_seed.. (n and edge factor
    drawn from PCG64(seed))."""
    out = []
    for s in range(first_seed, first_seed + count):
        rng = np.random.default_rng(s)
        n = int(round(2.0 ** rng.uniform(nmin_log2, nmax_log2)))
        ef = int(rng.integers(4, 33))
        _, _, rp, col = rmat(n, ef, s)
        out.append((n, n, rp, col, s))
    return out


def serving_mix(n_banded=32, n_rmat=32, seed=5, nmin=1 << 16, nmax=1 << 20):
    """[(kind, n, rowptr, col)]: banded matrices first, then R-MAT."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_banded):
        w = int(rng.integers(3, 65))
        n = int(rng.integers(nmin, nmax + 1))
        rp, col = gen_banded(n, w, 0.7, seed * 1000 + i)
        out.append(("banded", n, rp, col))
    for i in range(n_rmat):
        n = int(rng.integers(nmin, nmax + 1))
        _, _, rp, col = rmat(n, 16, seed * 1000 + 500 + i)
        out.append(("rmat", n, rp, col))
    return out
