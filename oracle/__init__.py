"""ctypes wrapper over oracle/liboracle.so -- the CPU restatement of TaiChi's
binary-SpMV path.  TEST INFRASTRUCTURE: imported only by tests/, by
``__graft_entry__.smoke()`` and by bench.py's cpu_baseline leg, always as the
checker.  Product code (taichi_amd/) never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


class OrcTaiChiNew(C.Structure):
    """Field order of TaiChi_new (TaiChi.h:45-60) plus an 'aliased' flag."""
    _fields_ = [(n, C.c_int) for n in ("dia_shapes", "num_negs", "dense_nnz",
                                        "zero_elements", "sparse_nnz",
                                        "row_start", "row_stop")] + \
               [(n, C.POINTER(C.c_int)) for n in ("xStart", "yStart", "xStop",
                                                   "nDiasPtr", "neg_offsets",
                                                   "csrRow", "csrCol")] + \
               [("aliased", C.c_int)]


def build():
    subprocess.check_call(["make", "-B", "-s", "-C", _HERE, "all"])
    if os.path.isdir("/root/reference"):
        subprocess.check_call(["make", "-B", "-s", "-C", _HERE, "ref"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.POINTER
        L.orc_read_mtx.argtypes = [C.c_char_p, P(C.c_int), P(C.c_int), P(C.c_ulong),
                                   P(P(C.c_int)), P(P(C.c_int)), P(P(C.c_float))]
        L.orc_mm_read_banner_file.argtypes = [C.c_char_p, C.c_char_p]
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_thumb_dims.argtypes = [C.c_int, C.c_int] + [P(C.c_int)] * 4
        L.orc_thumbnail.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _u8p]
        L.orc_imfill.argtypes = [_u8p, _u8p, C.c_int, C.c_int]
        L.orc_erode.argtypes = [_u8p, _u8p, C.c_int, C.c_int, _i32p, _i32p, C.c_int]
        L.orc_dilate.argtypes = [_u8p, _u8p, C.c_int, C.c_int, _i32p, _i32p, C.c_int]
        L.orc_open.argtypes = [_u8p, _u8p, C.c_int, C.c_int, C.c_int]
        L.orc_find_contours.argtypes = [_u8p, C.c_int, C.c_int, P(P(C.c_int)),
                                        P(P(C.c_int)), P(P(C.c_int))]
        L.orc_sdf_from_image.argtypes = [_u8p, C.c_int, C.c_int, P(P(C.c_int))]
        L.orc_cal_shape_size.argtypes = [C.c_int] * 7 + [P(C.c_int)] * 4 + [P(C.c_float)]
        L.orc_cal_start_end.argtypes = [C.c_int] * 5 + [P(C.c_int)] * 2
        L.orc_width_for.argtypes = [C.c_int]
        L.orc_partition_shapes.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p,
                                           C.c_void_p, C.c_void_p, P(OrcTaiChiNew)]
        L.orc_free_taichi_new_arrays.argtypes = [P(OrcTaiChiNew)]
        L.orc_csr_spmv_f32.argtypes = [C.c_int, _i32p, _i32p, _f32p, _f32p]
        L.orc_csr_spmv_f32_mt.argtypes = [C.c_int, _i32p, _i32p, _f32p, _f32p, C.c_int]
        L.orc_csr_spmv_f64.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p]
        L.orc_csr_absrow_f64.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p]
        L.orc_csr_spmm_f64.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, C.c_int]
        L.orc_taichi_spmv_f32.argtypes = [C.c_int, P(OrcTaiChiNew), _f32p, _f32p]
        L.orc_taichi_spmv_f64.argtypes = [C.c_int, P(OrcTaiChiNew), _f64p, _f64p]
        _LIB = L
    return _LIB


def _take(ptr, n, dtype):
    """Copy n elements out of a malloc'd C array and free it."""
    L = lib()
    if n:
        out = np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)
    else:
        out = np.zeros(0, dtype)
    L.orc_free(C.cast(ptr, C.c_void_p))
    return out


def read_mtx(path):
    """readMtx (TaiChi.cu:120-278). Returns (rc, M, N, rowptr, col, val)."""
    L = lib()
    m, n, nnz = C.c_int(), C.c_int(), C.c_ulong()
    rp, cp, vp = C.POINTER(C.c_int)(), C.POINTER(C.c_int)(), C.POINTER(C.c_float)()
    rc = L.orc_read_mtx(os.fsencode(path), C.byref(m), C.byref(n), C.byref(nnz),
                        C.byref(rp), C.byref(cp), C.byref(vp))
    if rc != 0:
        return rc, 0, 0, None, None, None
    M = m.value
    return (0, M, n.value, _take(rp, M + 1, np.int32), _take(cp, nnz.value, np.int32),
            _take(vp, nnz.value, np.float32))


def banner(path):
    tc = C.create_string_buffer(4)
    rc = lib().orc_mm_read_banner_file(os.fsencode(path), tc)
    return rc, tc.raw.decode("latin1")


def thumb_dims(M, N):
    H, W, pM, pN = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    rc = lib().orc_thumb_dims(M, N, C.byref(H), C.byref(W), C.byref(pM), C.byref(pN))
    return rc, H.value, W.value, pM.value, pN.value


def thumbnail(M, N, rowptr, col):
    rc, H, W, _, _ = thumb_dims(M, N)
    if rc != 0:
        return None
    img = np.empty(H * W, np.uint8)
    lib().orc_thumbnail(M, N, np.ascontiguousarray(rowptr, np.int32),
                        np.ascontiguousarray(col, np.int32), img)
    return img.reshape(H, W)


def imfill(img):
    img = np.ascontiguousarray(img, np.uint8)
    out = np.empty_like(img)
    lib().orc_imfill(img.ravel(), out.ravel(), img.shape[0], img.shape[1])
    return out


def open_(img, which):
    img = np.ascontiguousarray(img, np.uint8)
    out = np.empty_like(img)
    lib().orc_open(img.ravel(), out.ravel(), img.shape[0], img.shape[1], which)
    return out


def erode(img, dx, dy):
    img = np.ascontiguousarray(img, np.uint8)
    out = np.empty_like(img)
    dx = np.ascontiguousarray(dx, np.int32)
    dy = np.ascontiguousarray(dy, np.int32)
    lib().orc_erode(img.ravel(), out.ravel(), img.shape[0], img.shape[1], dx, dy, len(dx))
    return out


def dilate(img, dx, dy):
    img = np.ascontiguousarray(img, np.uint8)
    out = np.empty_like(img)
    dx = np.ascontiguousarray(dx, np.int32)
    dy = np.ascontiguousarray(dy, np.int32)
    lib().orc_dilate(img.ravel(), out.ravel(), img.shape[0], img.shape[1], dx, dy, len(dx))
    return out


def find_contours(img):
    """Returns [(points (n,2) as (x,y), is_hole), ...] in OpenCV output order."""
    img = np.ascontiguousarray(img, np.uint8)
    P = C.POINTER(C.c_int)
    pts, ptr, hole = P(), P(), P()
    nc = lib().orc_find_contours(img.ravel(), img.shape[0], img.shape[1],
                                 C.byref(pts), C.byref(ptr), C.byref(hole))
    ptr_a = _take(ptr, nc + 1, np.int64)
    pts_a = _take(pts, int(2 * ptr_a[-1]), np.int64).reshape(-1, 2)
    hole_a = _take(hole, nc, np.int64)
    return [(pts_a[ptr_a[i]:ptr_a[i + 1]], bool(hole_a[i])) for i in range(nc)]


def sdf_from_image(img):
    """getSdf's morphology + contour corners: array (shapeNum, 4) of x1,y1,x2,y2."""
    img = np.ascontiguousarray(img, np.uint8)
    pts = C.POINTER(C.c_int)()
    k = lib().orc_sdf_from_image(img.ravel(), img.shape[0], img.shape[1], C.byref(pts))
    return _take(pts, 4 * k, np.int64).reshape(k, 4)


def cal_shape_size(M, N, mF, x1, y1, a1, b1):
    o = [C.c_int() for _ in range(4)]
    g = C.c_float()
    lib().orc_cal_shape_size(M, N, mF, x1, y1, a1, b1, *[C.byref(v) for v in o], C.byref(g))
    return tuple(v.value for v in o) + (g.value,)


def cal_start_end(M_sub, N_sub, width, off, n):
    s, e = C.c_int(), C.c_int()
    lib().orc_cal_start_end(M_sub, N_sub, width, off, n, C.byref(s), C.byref(e))
    return s.value, e.value


def width_for(sd):
    return lib().orc_width_for(sd)


class Format:
    """A TaiChi_new produced by the oracle, copied into numpy arrays."""

    FIELDS = ("dia_shapes", "num_negs", "dense_nnz", "zero_elements", "sparse_nnz",
              "row_start", "row_stop")

    def __init__(self, M, st):
        for f in self.FIELDS:
            setattr(self, f, getattr(st, f))
        S, D = st.dia_shapes, st.num_negs

        def arr(p, n):
            if n <= 0 or not p:
                return np.zeros(0, np.int32)
            return np.ctypeslib.as_array(p, shape=(n,)).copy()
        self.xStart = arr(st.xStart, S)
        self.yStart = arr(st.yStart, S)
        self.xStop = arr(st.xStop, S)
        self.nDiasPtr = arr(st.nDiasPtr, S + 1 if S else 0)
        self.neg_offsets = arr(st.neg_offsets, D)
        self.csrRow = arr(st.csrRow, M + 1)
        R = int(self.csrRow[-1]) if len(self.csrRow) else 0
        self.csrCol = arr(st.csrCol, R)
        self.aliased = bool(st.aliased)
        self.M = M


def partition(M, N, rowptr, col, val=None, img=None):
    """partition_shapes (TaiChi.cu:313-638), serial. Returns (Format, val')."""
    L = lib()
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    v = None if val is None else np.array(val, np.float32, copy=True)
    im = None if img is None else np.ascontiguousarray(img, np.uint8)
    st = OrcTaiChiNew()
    L.orc_partition_shapes(M, N, len(col), rowptr, col,
                           None if v is None else v.ctypes.data_as(C.c_void_p),
                           None if im is None else im.ctypes.data_as(C.c_void_p),
                           C.byref(st))
    fmt = Format(M, st)
    L.orc_free_taichi_new_arrays(C.byref(st))
    return fmt, v


def _as_struct(fmt):
    keep = []

    def p(a):
        a = np.ascontiguousarray(a, np.int32)
        keep.append(a)
        return a.ctypes.data_as(C.POINTER(C.c_int))
    st = OrcTaiChiNew()
    for f in Format.FIELDS:
        setattr(st, f, getattr(fmt, f))
    for f in ("xStart", "yStart", "xStop", "nDiasPtr", "neg_offsets", "csrRow", "csrCol"):
        setattr(st, f, p(getattr(fmt, f)))
    st.aliased = 1
    return st, keep


def taichi_spmv(fmt, x, dtype=np.float64):
    st, keep = _as_struct(fmt)
    x = np.ascontiguousarray(x, dtype)
    y = np.empty(fmt.M, dtype)
    fn = lib().orc_taichi_spmv_f64 if dtype == np.float64 else lib().orc_taichi_spmv_f32
    fn(fmt.M, C.byref(st), x, y)
    return y


def csr_spmv_f32(rowptr, col, x):
    M = len(rowptr) - 1
    y = np.empty(M, np.float32)
    lib().orc_csr_spmv_f32(M, np.ascontiguousarray(rowptr, np.int32),
                           np.ascontiguousarray(col, np.int32),
                           np.ascontiguousarray(x, np.float32), y)
    return y


def csr_spmv_f32_mt(rowptr, col, x, threads):
    """csr_spmv_f32 with rows split over `threads` OpenMP threads (same y)."""
    M = len(rowptr) - 1
    y = np.empty(M, np.float32)
    lib().orc_csr_spmv_f32_mt(M, np.ascontiguousarray(rowptr, np.int32),
                              np.ascontiguousarray(col, np.int32),
                              np.ascontiguousarray(x, np.float32), y, int(threads))
    return y


def csr_spmv_f64(rowptr, col, x):
    M = len(rowptr) - 1
    y = np.empty(M, np.float64)
    lib().orc_csr_spmv_f64(M, np.ascontiguousarray(rowptr, np.int32),
                           np.ascontiguousarray(col, np.int32),
                           np.ascontiguousarray(x, np.float64), y)
    return y


def csr_absrow_f64(rowptr, col, x):
    M = len(rowptr) - 1
    y = np.empty(M, np.float64)
    lib().orc_csr_absrow_f64(M, np.ascontiguousarray(rowptr, np.int32),
                             np.ascontiguousarray(col, np.int32),
                             np.ascontiguousarray(x, np.float64), y)
    return y


def csr_spmm_f64(rowptr, col, X):
    X = np.ascontiguousarray(X, np.float64)
    M = len(rowptr) - 1
    k = X.shape[1]
    Y = np.empty((M, k), np.float64)
    lib().orc_csr_spmm_f64(M, np.ascontiguousarray(rowptr, np.int32),
                           np.ascontiguousarray(col, np.int32), X.ravel(), Y.ravel(), k)
    return Y


def rel_err(y, y64, absrow):
    """max_i |y_i - y64_i| / max(|y64_i|, sum_j |x_j| over row i)."""
    den = np.maximum(np.abs(y64), absrow)
    den = np.where(den == 0, 1.0, den)
    return float(np.max(np.abs(np.asarray(y, np.float64) - y64) / den)) if len(y64) else 0.0
