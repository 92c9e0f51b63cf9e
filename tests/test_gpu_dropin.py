"""The drop-in path as the reference's driver runs it (main.cu:19-119):
readMtx, partition_shapes with csrVal, taichi_SpMV, the CPU CSR check.

  * the built taichi_spmv driver (the main.cu equivalent, through the
    TaiChi.h surface) runs on .mtx files written by tc_write_mtx and must
    report no wrong entry (x = 1, |y - y_ref| <= 1e-6, main.cu:53-86)
  * partition_shapes with csrVal, as main.cu:45 calls it, runs the device
    build and marks every claimed entry FLT_MAX (TaiChi.cu:866): the format
    and the markers equal the oracle's
  * a csrVal that already holds markers (entries the reference treats as
    claimed from the start, TaiChi.cu:730) takes the host build: the same
    format and markers as the oracle
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import taichi_amd as T
from taichi_amd import workloads as W

from helpers import band_csr, band_goldens, rmat

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(T.LIB_PATH), "taichi_spmv")
ARRAYS = ("xStart", "yStart", "xStop", "nDiasPtr", "neg_offsets", "csrRow", "csrCol")


def _cases():
    out = [("cfg1", *W.diag5_random()), ("rmat18", *rmat(18, seed=3))]
    g = band_goldens()["band_pair_2048"]
    out.append(("band_pair", g["M"], g["N"], *band_csr(g["M"], g["bands"])))
    rp, col = T.gen_banded(40000, 12, 0.75, 5)
    out.append(("banded40000_w12", 40000, 40000, rp, col))
    return out


@pytest.mark.parametrize("case", _cases()[:3], ids=lambda c: c[0])
def test_cli_driver_reports_no_wrong_entry(case, tmp_path, dev):
    name, M, N, rp, col = case
    path = str(tmp_path / f"{name}.mtx")
    T.write_mtx(path, M, N, rp, col)
    assert os.path.exists(CLI), "taichi_spmv is not built"
    r = subprocess.run([CLI, path], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Warning: 0 are wrong!" in r.stdout, r.stdout
    assert f"M={M} N={N} nnz={len(col)}" in r.stdout


def _same(fmt, of):
    for f in O.Format.FIELDS:
        assert getattr(fmt, f) == getattr(of, f), f
    for f in ARRAYS:
        assert np.array_equal(getattr(fmt, f), getattr(of, f)), f


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_partition_with_val_marks_like_the_oracle(case, dev):
    name, M, N, rp, col = case
    val = np.ones(len(col), np.float32)  # readMtx's pattern values
    fmt, v = T.partition_shapes(M, N, rp, col, val=val)
    of, ov = O.partition(M, N, rp, col, val=val)
    _same(fmt, of)
    assert np.array_equal(v, ov)
    assert int(np.sum(v == np.float32(np.finfo(np.float32).max))) == of.dense_nnz
    # the same format as the build without val
    f2, _ = T.partition_shapes(M, N, rp, col)
    _same(f2, of)


def test_partition_with_premarked_val_takes_the_host_build(dev):
    g = band_goldens()["band_pair_2048"]
    M, N = g["M"], g["N"]
    rp, col = band_csr(M, g["bands"])
    val = np.ones(len(col), np.float32)
    fmax = np.float32(np.finfo(np.float32).max)
    val[rp[100]] = fmax  # row 100's +2048 entry counts as claimed already
    fmt, v = T.partition_shapes(M, N, rp, col, val=val)
    of, ov = O.partition(M, N, rp, col, val=val)
    _same(fmt, of)
    assert np.array_equal(v, ov)
