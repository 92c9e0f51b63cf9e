"""The on-disk format cache round-trips every partitioned format exactly and
rejects damaged files; the parallel Matrix Market parse gives the serial
reader's CSR (the oracle's restatement of TaiChi.cu:120-278) on large files
of every value kind and layout. CPU only."""
import os

import numpy as np
import pytest

import oracle as O
import taichi_amd as T
from taichi_amd import workloads as W

from helpers import host_format


def _formats():
    out = []
    M, N, rp, col = W.diag5_random()
    out.append(("cfg1", M, N, host_format(M, N, rp, col)))
    M, N, rp, col = W.rmat(1 << 14, 16, 3)
    out.append(("rmat14", M, N, host_format(M, N, rp, col)))
    rp, col = T.gen_banded(20000, 17, 0.7, 4)
    out.append(("banded", 20000, 20000, host_format(20000, 20000, rp, col)))
    rp, col = T.gen_stencil5(120, 120, 0.02, 5)                 # S = 0: the CSR itself
    out.append(("stencil", 14400, 14400, host_format(14400, 14400, rp, col)))
    rp = np.zeros(601, np.int32)                                # M <= 512 rows? no: 600 empty rows
    out.append(("empty", 600, 600, host_format(600, 600, rp, np.zeros(0, np.int32))))
    return out


def _same(a, b):
    assert (a.M, a.N) == (b.M, b.N)
    for f in T.Format.FIELDS:
        assert getattr(a, f) == getattr(b, f), f
    for f in T.Format.ARRAYS:
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


@pytest.mark.parametrize("case", _formats(), ids=lambda c: c[0])
def test_round_trip(case, tmp_path):
    name, M, N, fmt = case
    path = tmp_path / f"{name}.tcfmt"
    T.save_format(path, fmt)
    back = T.load_format(path)
    _same(back, fmt)
    assert T.footprint(back) == T.footprint(fmt)


def test_damaged_files_are_rejected(tmp_path):
    M, N, rp, col = W.rmat(1 << 13, 16, 6)
    fmt = host_format(M, N, rp, col)
    good = tmp_path / "good.tcfmt"
    T.save_format(good, fmt)
    raw = good.read_bytes()
    cases = {
        "flipped": raw[:200] + bytes([raw[200] ^ 1]) + raw[201:],      # payload byte
        "truncated": raw[:-9],
        "extended": raw + b"\0\0\0\0",
        "magic": b"X" + raw[1:],
        "checksum": raw[:-1] + bytes([raw[-1] ^ 0x80]),
        "last_byte": raw[:-9] + bytes([raw[-9] ^ 0x01]) + raw[-8:],     # tail of the checksummed bytes
    }
    for what, data in cases.items():
        p = tmp_path / f"{what}.tcfmt"
        p.write_bytes(data)
        with pytest.raises(T.TaiChiError) as e:
            T.load_format(p)
        assert e.value.status == -101, what
    with pytest.raises(T.TaiChiError):
        T.load_format(tmp_path / "missing.tcfmt")


def _write(path, banner, M, N, entries, sep="\n"):
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate {banner}\n% generated\n{M} {N} {len(entries)}\n")
        f.write(sep.join(entries))
        f.write("\n")


def _big_entries(kind, n=400000, M=50000, N=60000, seed=0):
    rng = np.random.default_rng(seed)
    r = rng.integers(1, M + 1, n)
    c = rng.integers(1, N + 1, n)
    if kind == "pattern":
        return [f"{a} {b}" for a, b in zip(r, c)]
    if kind == "integer":
        v = rng.integers(-1000, 1000, n)
        return [f"{a} {b} {x}" for a, b, x in zip(r, c, v)]
    v = rng.standard_normal(n)
    return [f"{a}\t{b}  {x:.9g}" for a, b, x in zip(r, c, v)]


def _check_same_csr(path):
    want = O.read_mtx(str(path))
    got = T.read_mtx(str(path))
    assert got[0] == want[0]
    if want[0] == 0:
        for g, w in zip(got[1:], want[1:]):
            assert np.array_equal(np.asarray(g), np.asarray(w))


@pytest.mark.parametrize("kind,storage", [("pattern", "general"), ("integer", "general"), ("real", "general"),
                                          ("real", "symmetric")])
def test_parallel_parse_equals_serial(kind, storage, tmp_path):
    entries = _big_entries(kind)
    if storage == "symmetric":   # lower triangle only, square
        entries = [e for e in _big_entries(kind, M=60000, N=60000)
                   if int(e.split()[0]) >= int(e.split()[1])]
    path = tmp_path / "big.mtx"
    _write(path, f"{kind} {storage}", 60000 if storage == "symmetric" else 50000, 60000, entries)
    assert os.path.getsize(path) > (1 << 22)       # large enough for the parallel path
    _check_same_csr(path)


@pytest.mark.parametrize("layout", ["two_per_line", "crlf", "trailing_garbage", "short", "out_of_range"])
def test_irregular_layouts_match_the_serial_reader(layout, tmp_path):
    entries = _big_entries("pattern", n=600000)
    path = tmp_path / f"{layout}.mtx"
    if layout == "two_per_line":      # the reference reads tokens, not lines
        paired = [entries[i] + " " + entries[i + 1] for i in range(0, len(entries) - 1, 2)]
        with open(path, "w") as f:
            f.write(f"%%MatrixMarket matrix coordinate pattern general\n50000 60000 {len(entries)}\n")
            f.write("\n".join(paired) + "\n")
    elif layout == "crlf":
        _write(path, "pattern general", 50000, 60000, entries, sep="\r\n")
    elif layout == "trailing_garbage":  # entries past NZ are never read
        _write(path, "pattern general", 50000, 60000, entries)
        with open(path, "a") as f:
            f.write("this is not an entry\n")
    elif layout == "short":           # fewer entries than the size line says
        with open(path, "w") as f:
            f.write(f"%%MatrixMarket matrix coordinate pattern general\n50000 60000 {len(entries) + 5}\n")
            f.write("\n".join(entries) + "\n")
    else:
        entries[123456] = "50001 7"
        _write(path, "pattern general", 50000, 60000, entries)
    _check_same_csr(path)
