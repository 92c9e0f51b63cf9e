"""a1 -- Matrix Market load: oracle and product readers against hand-derived
fixtures, and the banner/size parsing against the reference's own mmio.cpp
compiled into oracle/_ref (when the reference tree was present at build)."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle as O
import taichi_amd as T

GOLD = os.path.join(os.path.dirname(__file__), "golden")
EXP = json.load(open(os.path.join(GOLD, "expected.json")))["mtx"]
REF_SO = os.path.join(os.path.dirname(O.__file__), "_ref", "libmmio_ref.so")


@pytest.mark.parametrize("name", sorted(EXP))
@pytest.mark.parametrize("reader", ["oracle", "product"])
def test_reader_fixture(name, reader):
    exp = EXP[name]
    fn = O.read_mtx if reader == "oracle" else T.read_mtx
    rc, M, N, rp, col, val = fn(os.path.join(GOLD, name))
    assert rc == exp["rc"]
    if rc == 0:
        assert (M, N) == (exp["M"], exp["N"])
        assert rp.tolist() == exp["rowptr"]
        assert col.tolist() == exp["col"]
        assert np.array_equal(val, np.array(exp["val"], np.float32))


@pytest.mark.parametrize("reader", [O.read_mtx, T.read_mtx])
def test_missing_file(reader, tmp_path):
    assert reader(str(tmp_path / "nope.mtx"))[0] == -1


@pytest.mark.parametrize("name", sorted(EXP))
def test_oracle_banner_typecode(name):
    exp = EXP[name]
    rc, tc = O.banner(os.path.join(GOLD, name))
    if "banner_rc" in exp:
        assert rc == exp["banner_rc"]
    else:
        assert rc == 0 and tc == exp["typecode"]


def _ref():
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref not built (reference tree absent)")
    L = C.CDLL(REF_SO)
    L.ref_mm_banner.argtypes = [C.c_char_p, C.c_char_p]
    L.ref_mm_banner_size.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_int),
                                     C.POINTER(C.c_int), C.POINTER(C.c_ulong), C.POINTER(C.c_int)]
    return L


@pytest.mark.parametrize("name", sorted(EXP))
def test_banner_matches_reference_mmio(name):
    L = _ref()
    path = os.fsencode(os.path.join(GOLD, name))
    tc = C.create_string_buffer(4)
    rc_ref = L.ref_mm_banner(path, tc)
    rc_orc, tc_orc = O.banner(os.path.join(GOLD, name))
    assert rc_ref == rc_orc
    if rc_ref == 0:
        assert tc.raw.decode("latin1") == tc_orc


@pytest.mark.parametrize("name", [n for n, e in sorted(EXP.items())
                                  if e["rc"] in (0, -3) or n in ("truncated.mtx", "out_of_range.mtx")])
def test_size_line_matches_reference_mmio(name):
    """M and N of every coordinate fixture as mm_read_mtx_crd_size reads them."""
    L = _ref()
    tc = C.create_string_buffer(4)
    M, N, nz, rcs = C.c_int(), C.c_int(), C.c_ulong(), C.c_int()
    rc = L.ref_mm_banner_size(os.fsencode(os.path.join(GOLD, name)), tc, C.byref(M), C.byref(N),
                              C.byref(nz), C.byref(rcs))
    assert rc == 0 and rcs.value == 0
    exp = EXP[name]
    if exp["rc"] == 0:
        assert (M.value, N.value) == (exp["M"], exp["N"])
        # nz of a general file equals the listed entries
        if exp["typecode"][3] != "S":
            assert nz.value == len(exp["col"])


def test_roundtrip_write_read(tmp_path):
    n = 3000
    rp, col = T.gen_diagonals_random(n, [-1, 0, 3], 500, 11)
    p = str(tmp_path / "g.mtx")
    T.write_mtx(p, n, n, rp, col)
    for fn in (O.read_mtx, T.read_mtx):
        rc, M, N, rp2, col2, val2 = fn(p)
        assert rc == 0 and (M, N) == (n, n)
        assert np.array_equal(rp2, rp) and np.array_equal(col2, col)
        assert np.all(val2 == 1.0)
