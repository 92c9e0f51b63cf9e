"""The C-ABI library loads on a machine without a GPU and exports every
function that include/taichi_amd.h and include/TaiChi.h declare (no compute
call is made). Also checks the generators' determinism, which the GPU tests
and the benchmark rely on."""
import ast
import ctypes as C
import os
import re
import subprocess

import numpy as np

import taichi_amd as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def _strip_comments(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def _declared(header, pattern):
    text = _strip_comments(open(os.path.join(INC, header)).read())
    return sorted(set(re.findall(pattern, text)))


def test_c_abi_exports_every_declaration():
    names = _declared("taichi_amd.h", r"\b(tc_\w+)\s*\(")
    assert len(names) > 40
    lib = C.CDLL(T.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_stream_calibration_library_loads():
    """The benchmark's streaming-ceiling kernels build into a library of their
    own, outside the SpMV interface."""
    lib = C.CDLL(os.path.join(ROOT, "taichi_amd", "libtc_stream.so"))
    assert hasattr(lib, "tcs_copy") and hasattr(lib, "tcs_read")
    assert not hasattr(lib, "tc_spmv")


def test_cpp_surface_exports_every_declaration():
    """TaiChi.h keeps the reference's C++ signatures (TaiChi.h:79-107); each
    declared function is defined in the library with C++ linkage."""
    text = _strip_comments(open(os.path.join(INC, "TaiChi.h")).read())
    decls = re.findall(r"^\s*(?:int|void)\s+(\w+)\s*\(([^)]*)\)\s*;", text, flags=re.M)
    assert {n for n, _ in decls} >= {"readMtx", "partition_shapes", "taichi_SpMV", "free_taichi_new",
                                      "cal_shape_size", "partition_dia_shape", "cal_start_end",
                                      "queryDevice", "free_taichi", "max", "min"}
    out = subprocess.run(["nm", "-DC", "--defined-only", T.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {}
    for line in out.splitlines():
        parts = line.split(" ", 2)
        if len(parts) == 3 and parts[1] == "T":
            m = re.match(r"(\w+)\((.*)\)$", parts[2])
            if m:
                exported.setdefault(m.group(1), []).append(m.group(2))
    for name, params in decls:
        assert name in exported, name
        # parameter count matches the declaration
        n_decl = 0 if not params.strip() else len(params.split(","))
        assert any((0 if not p.strip() else len(p.split(","))) == n_decl for p in exported[name]), name


def test_status_strings_and_version():
    lib = C.CDLL(T.LIB_PATH)
    lib.tc_status_string.restype = C.c_char_p
    for code in (0, -101, -102, -103, -104, -105, -106, -107):
        assert lib.tc_status_string(code)
    assert lib.tc_abi_version() > 0


def test_generators_are_deterministic():
    a = T.gen_rmat(1 << 14, 14, 16 << 14, 0.499, 0.001, 0.001, 7)
    b = T.gen_rmat(1 << 14, 14, 16 << 14, 0.499, 0.001, 0.001, 7)
    c = T.gen_rmat(1 << 14, 14, 16 << 14, 0.499, 0.001, 0.001, 8)
    assert all(np.array_equal(u, v) for u, v in zip(a, b))
    assert not np.array_equal(a[1], c[1])
    rp, col = a
    n = 1 << 14
    rows = np.repeat(np.arange(n), np.diff(rp))
    assert np.all(np.diff(rp) >= 0) and col.min() >= 0 and col.max() < n
    keys = rows.astype(np.int64) * n + col
    assert np.all(np.diff(keys) > 0)                  # sorted within rows, no duplicates
    assert np.array_equal(T.gen_vector(1000, 2, 3), T.gen_vector(1000, 2, 3))
    xi = T.gen_vector(10000, 1, 5)
    assert set(np.unique(xi).tolist()) <= set(range(8))
    xu = T.gen_vector(10000, 0, 5)
    assert xu.min() >= 0.0 and xu.max() < 1.0


def test_library_shares_torchs_hip_runtime():
    """Loading the library before torch must not leave two HIP runtimes in the
    process (torch would map its own beside /opt/rocm's, and the library's
    would then find no device)."""
    code = ("import sys; sys.path.insert(0, %r)\n"
            "import taichi_amd as T\nT.lib()\nimport torch\n"
            "print(sorted({l.split()[-1] for l in open('/proc/self/maps') if 'libamdhip64' in l}))"
            % ROOT)
    out = subprocess.run(["python3", "-c", code], capture_output=True, text=True, timeout=300, cwd="/tmp")
    assert out.returncode == 0, out.stderr
    maps = ast.literal_eval(out.stdout.strip().splitlines()[-1])
    assert len(maps) == 1, maps


def test_allocation_failure_is_a_status_not_a_crash():
    """An exception inside a C entry point (here the morphology's buffers for a
    2^30 x 2^30 image) comes back as TC_ERR_NOMEM, never out of the C ABI."""
    code = ("import sys; sys.path.insert(0, %r)\n"
            "import numpy as np, taichi_amd as T\n"
            "L = T.lib(); b = np.zeros(1, np.uint8)\n"
            "print(L.tc_morph_from_image(b, 1 << 30, 1 << 30, b), L.tc_last_error().decode())" % ROOT)
    out = subprocess.run(["python3", "-c", code], capture_output=True, text=True, timeout=300, cwd="/tmp")
    assert out.returncode == 0, out.stderr[-2000:]
    rc, msg = out.stdout.strip().splitlines()[-1].split(" ", 1)
    assert int(rc) == -103 and "tc_morph_from_image" in msg, out.stdout
