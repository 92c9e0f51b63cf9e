"""taichi_amd -- MI355X-native TaiChi binary SpMV (Python host surface).

Python mirror of the reference's C++ operator surface (TaiChi.h:79-107):
``read_mtx`` (readMtx), ``partition_shapes`` (partition_shapes),
``taichi_spmv`` (taichi_SpMV), ``free_taichi_new`` is implicit (formats are
numpy-owned here).  Everything calls the C ABI of ``libtaichi_amd.so``
(include/taichi_amd.h) through ctypes; the compute runs in the library's HIP
kernels.  There is no Python or CPU fallback: if the library is missing,
``lib()`` raises.

Device buffers are torch tensors (plumbing only): ``Matrix.spmv`` takes
CUDA/HIP tensors and launches on the current torch stream unless a stream is
given.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtaichi_amd.so")
_LIB = None

F32, BF16, F16 = 0, 1, 2
TC_OK = 0


class TaiChiError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{msg} (status {status})")
        self.status = status


class TaiChiNew(C.Structure):
    """TaiChi_new with the reference's field order (TaiChi.h:45-60)."""
    _fields_ = [(n, C.c_int) for n in ("dia_shapes", "num_negs", "dense_nnz",
                                        "zero_elements", "sparse_nnz",
                                        "row_start", "row_stop")] + \
               [(n, C.POINTER(C.c_int)) for n in ("xStart", "yStart", "xStop",
                                                   "nDiasPtr", "neg_offsets",
                                                   "csrRow", "csrCol")]


class MatrixStats(C.Structure):
    _fields_ = [("M", C.c_int), ("N", C.c_int), ("dia_shapes", C.c_int),
                ("num_negs", C.c_int), ("long_rows", C.c_int), ("offset_bytes", C.c_int),
                ("nnz", C.c_longlong), ("remainder", C.c_longlong),
                ("format_bytes", C.c_longlong), ("device_bytes", C.c_longlong),
                ("holes", C.c_longlong), ("bitmap_shapes", C.c_int), ("near_window", C.c_int)]


class Config(C.Structure):
    _fields_ = [("rows_per_tile", C.c_int), ("lds_values", C.c_int), ("grid", C.c_int),
                ("flags", C.c_int)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_cap_bytes", C.c_longlong), ("max_batch", C.c_int), ("streams", C.c_int),
                ("max_inflight", C.c_int), ("launch", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("resident_bytes", C.c_longlong), ("matrices", C.c_int), ("pending", C.c_int),
                ("inflight", C.c_int), ("completed", C.c_longlong), ("batches", C.c_longlong),
                ("mean_batch", C.c_double)]


class LoadItem(C.Structure):
    _fields_ = [("matrix_id", C.c_int), ("X", C.c_void_p), ("ldx", C.c_int), ("dtype", C.c_int),
                ("k", C.c_int)]


class LoadSpec(C.Structure):
    _fields_ = [("items", C.POINTER(LoadItem)), ("nitems", C.c_int),
                ("ybufs", C.POINTER(C.c_void_p)), ("nybufs", C.c_int), ("ldy", C.c_int),
                ("rate", C.c_double), ("concurrency", C.c_int), ("count", C.c_longlong),
                ("seed", C.c_uint64)]


class LoadResult(C.Structure):
    _fields_ = [("completed", C.c_longlong), ("seconds", C.c_double), ("req_per_s", C.c_double),
                ("p50_us", C.c_double), ("p99_us", C.c_double), ("mean_us", C.c_double),
                ("max_us", C.c_double), ("mean_batch", C.c_double)]


class Request(C.Structure):
    _fields_ = [("A", C.c_void_p), ("X", C.c_void_p), ("ldx", C.c_int), ("dtype", C.c_int),
                ("Y", C.c_void_p), ("ldy", C.c_int), ("k", C.c_int)]


def build():
    """Compile libtaichi_amd.so and taichi_spmv for gfx950 in-tree (always a full build)."""
    jobs = str(min(16, os.cpu_count() or 4))
    # -B: every call compiles every source, also when the library is up to date
    subprocess.check_call(["make", "-B", "-s", "-C", os.path.join(_HERE, "csrc"), "-j", jobs, "all"])


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                           "(the HIP library has no fallback)")
    # torch first: its HIP runtime (soname libamdhip64.so.7) then satisfies
    # the library's dependency, so the library and torch's tensors share one
    # runtime. Loaded the other way round, torch maps a second HIP/HSA runtime
    # beside /opt/rocm's and the library's runtime finds no device.
    import torch  # noqa: F401
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    i32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    u8 = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
    f32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    vp = C.c_void_p
    sig = {
        "tc_status_string": (C.c_char_p, [C.c_int]),
        "tc_last_error": (C.c_char_p, []),
        "tc_abi_version": (C.c_int, []),
        "tc_device_count": (C.c_int, []),
        "tc_read_mtx": (C.c_int, [C.c_char_p, P(C.c_int), P(C.c_int), P(C.c_ulong),
                                  P(P(C.c_int)), P(P(C.c_int)), P(P(C.c_float))]),
        "tc_free_host": (None, [vp]),
        "tc_write_mtx": (C.c_int, [C.c_char_p, C.c_int, C.c_int, i32, vp]),
        "tc_thumb_dims": (C.c_int, [C.c_int, C.c_int] + [P(C.c_int)] * 4),
        "tc_thumbnail": (C.c_int, [C.c_int, C.c_int, vp, vp, vp, vp]),
        "tc_sdf_from_image": (C.c_int, [u8, C.c_int, C.c_int, vp, C.c_int, P(C.c_int)]),
        "tc_contours": (C.c_int, [u8, C.c_int, C.c_int, vp, C.c_int, P(C.c_int)]),
        "tc_cal_shape_size": (None, [C.c_int] * 7 + [P(C.c_int)] * 4 + [P(C.c_float)]),
        "tc_cal_start_end": (None, [C.c_int] * 5 + [P(C.c_int)] * 2),
        "tc_partition_from_image": (C.c_int, [C.c_int, C.c_int, C.c_int, i32, vp, vp, vp,
                                              P(TaiChiNew)]),
        "tc_partition_shapes": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, i32, vp, vp,
                                          P(TaiChiNew)]),
        "tc_partition_device": (C.c_int, [C.c_int, C.c_int, C.c_int, i32, vp, vp, vp, P(TaiChiNew),
                                          vp]),
        "tc_morph_from_image": (C.c_int, [u8, C.c_int, C.c_int, u8]),
        "tc_morph_work_bytes": (C.c_size_t, [C.c_int, C.c_int]),
        "tc_morph_device": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, vp]),
        "tc_taichi_new_alloc": (P(TaiChiNew), []),
        "tc_save_format": (C.c_int, [C.c_char_p, C.c_int, C.c_int, P(TaiChiNew)]),
        "tc_load_format": (C.c_int, [C.c_char_p, P(C.c_int), P(C.c_int), P(TaiChiNew)]),
        "tc_free_taichi_new": (C.c_int, [P(TaiChiNew)]),
        "tc_taichi_spmv": (C.c_int, [C.c_int, C.c_int, P(TaiChiNew), f32, f32, C.c_int, C.c_int,
                                     P(C.c_float), P(C.c_float)]),
        "tc_matrix_create": (C.c_int, [C.c_int, C.c_int, P(TaiChiNew), C.c_longlong, C.c_int,
                                       vp, P(vp)]),
        "tc_matrix_destroy": (C.c_int, [vp]),
        "tc_matrix_get_stats": (C.c_int, [vp, P(MatrixStats)]),
        "tc_spmv_bytes": (C.c_longlong, [vp, C.c_int, C.c_int]),
        "tc_spmv": (C.c_int, [vp, vp, C.c_int, C.c_int, vp, C.c_int, C.c_int, vp]),
        "tc_matrix_set_config": (C.c_int, [vp, C.c_int, C.c_int, P(Config)]),
        "tc_matrix_get_config": (C.c_int, [vp, C.c_int, C.c_int, P(Config)]),
        "tc_autotune": (C.c_int, [vp, vp, C.c_int, C.c_int, vp, C.c_int, C.c_int, vp,
                                  P(Config), P(C.c_float)]),
        "tc_spmv_batch": (C.c_int, [P(Request), C.c_int, vp]),
        "tc_spmv_batch_static": (C.c_int, [P(Request), C.c_int, vp]),
        "tc_graph_capture": (C.c_int, [vp, vp, C.c_int, C.c_int, vp, C.c_int, C.c_int, C.c_int,
                                       vp, P(vp)]),
        "tc_graph_launch": (C.c_int, [vp, vp]),
        "tc_pagerank_work_bytes": (C.c_size_t, [C.c_int]),
        "tc_pagerank": (C.c_int, [vp, vp, C.c_float, C.c_int, vp, vp, vp]),
        "tc_pagerank_capture": (C.c_int, [vp, vp, C.c_float, C.c_int, vp, vp, vp, P(vp)]),
        "tc_graph_destroy": (C.c_int, [vp]),
        "tc_matrix_footprint": (C.c_int, [C.c_int, C.c_int, P(TaiChiNew), C.c_int,
                                          P(C.c_longlong)]),
        "tc_server_create": (C.c_int, [P(ServerConfig), P(vp)]),
        "tc_server_destroy": (C.c_int, [vp]),
        "tc_server_add_matrix": (C.c_int, [vp, C.c_int, C.c_int, P(TaiChiNew), C.c_longlong,
                                           P(C.c_int)]),
        "tc_server_remove_matrix": (C.c_int, [vp, C.c_int]),
        "tc_server_submit": (C.c_int, [vp, C.c_int, vp, C.c_int, C.c_int, vp, C.c_int, C.c_int,
                                       P(C.c_ulonglong)]),
        "tc_server_wait": (C.c_int, [vp, C.c_ulonglong, P(C.c_double)]),
        "tc_server_get_stats": (C.c_int, [vp, P(ServerStats)]),
        "tc_server_run_load": (C.c_int, [vp, P(LoadSpec), P(LoadResult)]),
        "tc_server_run_plain": (C.c_int, [vp, P(LoadSpec), P(LoadResult)]),
        "tc_gen_diagonals_random": (C.c_int, [C.c_int, i32, C.c_int, C.c_longlong, C.c_uint64,
                                              P(P(C.c_int)), P(P(C.c_int)), P(C.c_longlong)]),
        "tc_gen_rmat": (C.c_int, [C.c_int, C.c_int, C.c_longlong, C.c_double, C.c_double,
                                  C.c_double, C.c_uint64, P(P(C.c_int)), P(P(C.c_int)),
                                  P(C.c_longlong)]),
        "tc_gen_stencil5": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_uint64,
                                      P(P(C.c_int)), P(P(C.c_int)), P(C.c_longlong)]),
        "tc_gen_banded": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_uint64,
                                    P(P(C.c_int)), P(P(C.c_int)), P(C.c_longlong)]),
        "tc_gen_vector": (C.c_int, [f32, C.c_longlong, C.c_int, C.c_uint64]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def check(rc, what=""):
    if rc != TC_OK:
        L = lib()
        msg = L.tc_last_error().decode(errors="replace") or L.tc_status_string(rc).decode()
        raise TaiChiError(rc, f"{what}: {msg}" if what else msg)


def _take(ptr, n, dtype=np.int32):
    L = lib()
    out = np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True) if n else np.zeros(0, dtype)
    L.tc_free_host(C.cast(ptr, C.c_void_p))
    return out


def _vp(a):
    return None if a is None or len(a) == 0 else a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------- a1: load
def read_mtx(path):
    """readMtx (TaiChi.cu:120-278): (rc, M, N, rowptr, col, val); rc in 0,-1..-4."""
    L = lib()
    m, n, nnz = C.c_int(), C.c_int(), C.c_ulong()
    rp, cp, vp_ = C.POINTER(C.c_int)(), C.POINTER(C.c_int)(), C.POINTER(C.c_float)()
    rc = L.tc_read_mtx(os.fsencode(path), C.byref(m), C.byref(n), C.byref(nnz),
                       C.byref(rp), C.byref(cp), C.byref(vp_))
    if rc != 0:
        return rc, 0, 0, None, None, None
    M, Z = m.value, nnz.value
    return (0, M, n.value, _take(rp, M + 1), _take(cp, Z),
            _take(vp_, Z, np.float32))


def write_mtx(path, M, N, rowptr, col):
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    check(lib().tc_write_mtx(os.fsencode(path), M, N, rowptr, _vp(col)), "write_mtx")


# ----------------------------------------------------------- generators
def _gen_out():
    return C.POINTER(C.c_int)(), C.POINTER(C.c_int)(), C.c_longlong()


def _gen_finish(n, rp, cp, nnz):
    return _take(rp, n + 1), _take(cp, nnz.value)


def gen_diagonals_random(n, offsets, ndraws, seed):
    offs = np.ascontiguousarray(offsets, np.int32)
    rp, cp, nnz = _gen_out()
    check(lib().tc_gen_diagonals_random(n, offs, len(offs), ndraws, seed, C.byref(rp),
                                        C.byref(cp), C.byref(nnz)), "gen_diagonals_random")
    return _gen_finish(n, rp, cp, nnz)


def gen_rmat(n, scale, nedges, a, b, c, seed):
    rp, cp, nnz = _gen_out()
    check(lib().tc_gen_rmat(n, scale, nedges, a, b, c, seed, C.byref(rp), C.byref(cp),
                            C.byref(nnz)), "gen_rmat")
    return _gen_finish(n, rp, cp, nnz)


def gen_stencil5(gx, gy, extra_frac, seed):
    rp, cp, nnz = _gen_out()
    check(lib().tc_gen_stencil5(gx, gy, extra_frac, seed, C.byref(rp), C.byref(cp),
                                C.byref(nnz)), "gen_stencil5")
    return _gen_finish(gx * gy, rp, cp, nnz)


def gen_banded(n, w, density, seed):
    rp, cp, nnz = _gen_out()
    check(lib().tc_gen_banded(n, w, density, seed, C.byref(rp), C.byref(cp), C.byref(nnz)),
          "gen_banded")
    return _gen_finish(n, rp, cp, nnz)


def gen_vector(n, kind, seed):
    """kind 0: U[0,1), 1: integers {0..7}, 2: N(0,1)."""
    out = np.empty(n, np.float32)
    check(lib().tc_gen_vector(out, n, kind, seed), "gen_vector")
    return out


# ------------------------------------------------------- a2-a6: partition
def thumb_dims(M, N):
    H, W, pM, pN = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    rc = lib().tc_thumb_dims(M, N, C.byref(H), C.byref(W), C.byref(pM), C.byref(pN))
    return rc, H.value, W.value, pM.value, pN.value


def sdf_from_image(img):
    img = np.ascontiguousarray(img, np.uint8)
    cap = 4096
    buf = np.empty(4 * cap, np.int32)
    k = C.c_int()
    check(lib().tc_sdf_from_image(img.ravel(), img.shape[0], img.shape[1], _vp(buf), cap,
                                  C.byref(k)), "sdf_from_image")
    return buf[:4 * k.value].reshape(k.value, 4).astype(np.int64)


def contours(img):
    """[(points (n,2) as (x,y), is_hole), ...] in OpenCV's output order."""
    img = np.ascontiguousarray(img, np.uint8)
    need = C.c_int()
    L = lib()
    rc = L.tc_contours(img.ravel(), img.shape[0], img.shape[1], None, 0, C.byref(need))
    buf = np.empty(max(need.value, 1), np.int32)
    check(L.tc_contours(img.ravel(), img.shape[0], img.shape[1], _vp(buf), len(buf),
                        C.byref(need)), "contours")
    del rc
    n, pos, out = int(buf[0]), 1, []
    for _ in range(n):
        ln, hole = int(buf[pos]), bool(buf[pos + 1])
        pts = buf[pos + 2:pos + 2 + 2 * ln].reshape(ln, 2).astype(np.int64)
        out.append((pts, hole))
        pos += 2 + 2 * ln
    return out


class Format:
    """A TaiChi_new (TaiChi.h:45-60) held in numpy arrays."""

    FIELDS = ("dia_shapes", "num_negs", "dense_nnz", "zero_elements", "sparse_nnz",
              "row_start", "row_stop")
    ARRAYS = ("xStart", "yStart", "xStop", "nDiasPtr", "neg_offsets", "csrRow", "csrCol")

    def __init__(self, M, N, st=None):
        self.M, self.N = M, N
        if st is None:
            return
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

    def struct(self):
        """ctypes TaiChi_new pointing at this object's arrays (keep self alive)."""
        st = TaiChiNew()
        for f in self.FIELDS:
            setattr(st, f, int(getattr(self, f)))
        self._keep = []
        for f in self.ARRAYS:
            a = np.ascontiguousarray(getattr(self, f), np.int32)
            self._keep.append(a)
            setattr(st, f, a.ctypes.data_as(C.POINTER(C.c_int)) if len(a) else
                    C.POINTER(C.c_int)())
        return st

    @property
    def remainder(self):
        return int(self.csrRow[-1]) if len(self.csrRow) else 0


def partition_from_image(M, N, rowptr, col, val=None, img=None):
    """partition_shapes given the thumbnail (host logic). Returns (Format, val')."""
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    v = None if val is None else np.array(val, np.float32, copy=True)
    im = None if img is None else np.ascontiguousarray(img, np.uint8).ravel()
    st = TaiChiNew()
    check(lib().tc_partition_from_image(M, N, len(col), rowptr, _vp(col), _vp(v), _vp(im),
                                        C.byref(st)), "partition_from_image")
    fmt = _own(M, N, st)
    return fmt, v


def partition_device(M, N, rowptr, col, rowptr_dev, col_dev, stream=None):
    """Device build of partition_shapes (tc_partition_device) for a binary CSR;
    rowptr_dev/col_dev are int32 device tensors holding rowptr/col. Raises
    TaiChiError (status -104) when an entry is duplicated on a dense diagonal."""
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    st = TaiChiNew()
    check(lib().tc_partition_device(M, N, len(col), rowptr, _vp(col), rowptr_dev.data_ptr(),
                                    col_dev.data_ptr() if len(col) else None, C.byref(st),
                                    _torch_stream(stream)), "partition_device")
    return _own(M, N, st)


def partition_shapes(M, N, rowptr, col, val=None, path="matrix.mtx"):
    """partition_shapes (TaiChi.cu:313-638): device build (GPU thumbnail,
    histogram, claim, zero rows and merge); with val, or duplicates on a dense
    diagonal, GPU thumbnail + host partition."""
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    v = None if val is None else np.array(val, np.float32, copy=True)
    st = TaiChiNew()
    check(lib().tc_partition_shapes(os.fsencode(path), M, N, len(col), rowptr, _vp(col), _vp(v),
                                    C.byref(st)), "partition_shapes")
    fmt = _own(M, N, st)
    return fmt, v


def _own(M, N, st):
    """Copy a library-produced TaiChi_new into numpy and free the C arrays."""
    fmt = Format(M, N, st)
    if st.dense_nnz > 0:
        # free_taichi_new frees the arrays and the struct: hand it a heap copy
        L = lib()
        heap = L.tc_taichi_new_alloc()
        C.memmove(heap, C.byref(st), C.sizeof(TaiChiNew))
        L.tc_free_taichi_new(heap)
    return fmt


def inverse_out_degree(N, col):
    """inv_deg[j] = 1 / (entries in column j), 0 for an empty column: the
    PageRank scaling of a binary A where A[i][j] = 1 means j links to i."""
    deg = np.bincount(np.asarray(col, np.int64), minlength=N).astype(np.float64)
    inv = np.zeros(N, np.float32)
    nz = deg > 0
    inv[nz] = (1.0 / deg[nz]).astype(np.float32)
    return inv


class PageRank:
    """PageRank power iteration on a resident square Matrix (tc_pagerank):
    inv_deg is a device float32 tensor (inverse_out_degree); run() enqueues
    the iterations, capture() records them once as a replayable graph."""

    def __init__(self, A, inv_deg, alpha=0.85):
        import torch
        self.A, self.inv_deg, self.alpha = A, inv_deg, float(alpha)
        n = inv_deg.shape[0]
        self.work = torch.empty(int(lib().tc_pagerank_work_bytes(n)), dtype=torch.uint8,
                                device=inv_deg.device)

    def run(self, r, iters, stream=None):
        check(lib().tc_pagerank(self.A.handle, self.inv_deg.data_ptr(), self.alpha, iters,
                                r.data_ptr(), self.work.data_ptr(), _torch_stream(stream)),
              "pagerank")

    def capture(self, r, iters, stream=None):
        h = C.c_void_p()
        check(lib().tc_pagerank_capture(self.A.handle, self.inv_deg.data_ptr(), self.alpha, iters,
                                        r.data_ptr(), self.work.data_ptr(), stream, C.byref(h)),
              "pagerank_capture")
        g = Graph.__new__(Graph)
        g.handle = h
        return g


def save_format(path, fmt):
    """Write a partitioned format to the on-disk cache (tc_save_format)."""
    st = fmt.struct()
    check(lib().tc_save_format(os.fsencode(path), fmt.M, fmt.N, C.byref(st)), "save_format")


def load_format(path):
    """Read a format written by save_format (validated); returns a Format."""
    L = lib()
    M, N = C.c_int(), C.c_int()
    st = TaiChiNew()
    check(L.tc_load_format(os.fsencode(path), C.byref(M), C.byref(N), C.byref(st)), "load_format")
    fmt = Format(M.value, N.value, st)
    heap = L.tc_taichi_new_alloc()
    C.memmove(heap, C.byref(st), C.sizeof(TaiChiNew))
    if st.dense_nnz == 0:  # free_taichi_new leaves the (here loaded) CSR alone
        for f in ("csrRow", "csrCol"):
            p = getattr(st, f)
            if p:
                L.tc_free_host(C.cast(p, C.c_void_p))
    L.tc_free_taichi_new(heap)
    return fmt


def morph_from_image(img):
    """getSdf's morphology on the host (tc_morph_from_image): the binary image
    whose contours give the shapes; img is a uint8 [H, W] array."""
    img = np.ascontiguousarray(img, np.uint8)
    thin = np.empty_like(img)
    check(lib().tc_morph_from_image(img.ravel(), img.shape[0], img.shape[1], thin.ravel()),
          "morph_from_image")
    return thin


def morph_device(img_dev, H, W, stream=None):
    """The same morphology on the device (tc_morph_device); img_dev is a
    uint8 device tensor of H*W. Returns the thin image as a device tensor."""
    import torch
    thin = torch.empty(H * W, dtype=torch.uint8, device=img_dev.device)
    work = torch.empty(int(lib().tc_morph_work_bytes(H, W)), dtype=torch.uint8, device=img_dev.device)
    check(lib().tc_morph_device(img_dev.data_ptr(), H, W, thin.data_ptr(), work.data_ptr(),
                                _torch_stream(stream)), "morph_device")
    return thin


def thumbnail_gpu(M, N, rowptr_dev, col_dev, img_dev, stream=0):
    """Kernel (a) on device tensors; img_dev is a uint8 tensor of H*W."""
    check(lib().tc_thumbnail(M, N, rowptr_dev.data_ptr(), col_dev.data_ptr(),
                             img_dev.data_ptr(), stream or None), "thumbnail")


# ------------------------------------------------------------- a7: SpMV
def _torch_stream(stream):
    if stream is not None:
        return stream
    import torch
    return torch.cuda.current_stream().cuda_stream


class Matrix:
    """A TaiChi format resident in HBM (tc_matrix)."""

    def __init__(self, fmt, nnz=None, long_threshold=0, stream=None):
        self.M, self.N = fmt.M, fmt.N
        st = fmt.struct()
        h = C.c_void_p()
        if nnz is None:
            nnz = int(fmt.dense_nnz) + int(fmt.sparse_nnz)
        check(lib().tc_matrix_create(self.M, self.N, C.byref(st), nnz, long_threshold,
                                     _torch_stream(stream), C.byref(h)), "matrix_create")
        self.handle = h
        self._fmt_keep = fmt

    def close(self):
        if getattr(self, "handle", None):
            lib().tc_matrix_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self):
        s = MatrixStats()
        check(lib().tc_matrix_get_stats(self.handle, C.byref(s)), "stats")
        return {f: getattr(s, f) for f, _ in MatrixStats._fields_}

    def bytes(self, k=1, dtype=F32):
        return int(lib().tc_spmv_bytes(self.handle, k, dtype))

    @staticmethod
    def _dtype_of(X):
        import torch
        return {torch.float32: F32, torch.bfloat16: BF16, torch.float16: F16}[X.dtype]

    def spmv(self, X, Y, stream=None):
        """Y[M,k] = A @ X[N,k] (or 1-D vectors); X float32/bfloat16/float16, Y float32."""
        k = 1 if X.dim() == 1 else X.shape[1]
        ldx = 1 if X.dim() == 1 else X.stride(0)
        ldy = 1 if Y.dim() == 1 else Y.stride(0)
        assert X.shape[0] >= self.N and Y.shape[0] >= self.M and Y.dtype.itemsize == 4
        check(lib().tc_spmv(self.handle, X.data_ptr(), ldx, self._dtype_of(X), Y.data_ptr(), ldy,
                            k, _torch_stream(stream)), "spmv")

    def set_config(self, k, dtype, rows_per_tile=0, lds_values=0, grid=0, flags=0):
        c = Config(rows_per_tile, lds_values, grid, flags)
        check(lib().tc_matrix_set_config(self.handle, k, dtype, C.byref(c)), "set_config")

    def get_config(self, k=1, dtype=F32):
        c = Config()
        check(lib().tc_matrix_get_config(self.handle, k, dtype, C.byref(c)), "get_config")
        return c.rows_per_tile, c.lds_values, c.grid, c.flags

    def autotune(self, X, Y, stream=None):
        k = 1 if X.dim() == 1 else X.shape[1]
        ldx = 1 if X.dim() == 1 else X.stride(0)
        ldy = 1 if Y.dim() == 1 else Y.stride(0)
        c, us = Config(), C.c_float()
        check(lib().tc_autotune(self.handle, X.data_ptr(), ldx, self._dtype_of(X), Y.data_ptr(),
                                ldy, k, _torch_stream(stream), C.byref(c), C.byref(us)),
              "autotune")
        return (c.rows_per_tile, c.lds_values, c.grid, c.flags), us.value


class Batch:
    """A fixed request set [(Matrix, X, Y), ...] (k <= 64 each) for repeated
    batched launches: the C request array is built once, and each launch
    is one tc_spmv_batch call (asynchronous on its stream). Holds the items,
    so their tensors stay alive as long as the batch."""

    def __init__(self, items):
        self.items = list(items)
        self.reqs = (Request * len(self.items))()
        for i, (A, X, Y) in enumerate(self.items):
            k = 1 if X.dim() == 1 else X.shape[1]
            self.reqs[i] = Request(A.handle, X.data_ptr(), 1 if X.dim() == 1 else X.stride(0),
                                   Matrix._dtype_of(X), Y.data_ptr(), 1 if Y.dim() == 1 else Y.stride(0), k)

    def launch(self, stream=None):
        check(lib().tc_spmv_batch(self.reqs, len(self.items), _torch_stream(stream)), "spmv_batch")

    def launch_static(self, stream=None):
        """The same launch through tc_spmv_batch_static (no allocation; the
        descriptor table is the library's fixed per-GPU array)."""
        check(lib().tc_spmv_batch_static(self.reqs, len(self.items), _torch_stream(stream)),
              "spmv_batch_static")


def spmv_batch(items, stream=None):
    """Batched SpMV over [(Matrix, X, Y), ...] (k <= 64 each): one launch per
    (column chunk width, x type) group; each Y equals Matrix.spmv alone."""
    Batch(items).launch(stream)


class Graph:
    """`iters` back-to-back SpMV launches captured in a hipGraph."""

    def __init__(self, A, X, Y, iters, stream=None):
        k = 1 if X.dim() == 1 else X.shape[1]
        h = C.c_void_p()
        check(lib().tc_graph_capture(A.handle, X.data_ptr(), 1 if X.dim() == 1 else X.stride(0),
                                     Matrix._dtype_of(X), Y.data_ptr(),
                                     1 if Y.dim() == 1 else Y.stride(0), k, iters, stream,
                                     C.byref(h)), "graph_capture")
        self.handle = h

    def launch(self, stream=None):
        check(lib().tc_graph_launch(self.handle, _torch_stream(stream)), "graph_launch")

    def close(self):
        if getattr(self, "handle", None):
            lib().tc_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def taichi_spmv(fmt, x, warmup=10, runs=500):
    """taichi_SpMV (TaiChi.cu:1331-1456): host x in, host y out, timings in ms."""
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros(fmt.M, np.float32)
    st = fmt.struct()
    t_ms, s_ms = C.c_float(), C.c_float()
    check(lib().tc_taichi_spmv(fmt.M, fmt.N, C.byref(st), x, y, warmup, runs, C.byref(t_ms),
                               C.byref(s_ms)), "taichi_spmv")
    return y, t_ms.value, s_ms.value


def footprint(fmt, long_threshold=0):
    """Device bytes tc_matrix_create would allocate for fmt (host only)."""
    st = fmt.struct()
    b = C.c_longlong()
    check(lib().tc_matrix_footprint(fmt.M, fmt.N, C.byref(st), long_threshold, C.byref(b)),
          "footprint")
    return b.value


class Server:
    """Continuous-batching SpMV server over resident matrices (tc_server)."""

    def __init__(self, hbm_cap_bytes=200 << 30, max_batch=64, streams=3, max_inflight=256, launch=0):
        """launch 0 or 1: each batch of requests runs as one coalesced
        batched launch group; 2: one by one. Same results either way."""
        h = C.c_void_p()
        cfg = ServerConfig(hbm_cap_bytes, max_batch, streams, max_inflight, launch)
        check(lib().tc_server_create(C.byref(cfg), C.byref(h)), "server_create")
        self.handle = h

    def add_matrix(self, fmt, nnz=None):
        st = fmt.struct()
        mid = C.c_int()
        if nnz is None:
            nnz = int(fmt.dense_nnz) + int(fmt.sparse_nnz)
        check(lib().tc_server_add_matrix(self.handle, fmt.M, fmt.N, C.byref(st), nnz,
                                         C.byref(mid)), "server_add_matrix")
        return mid.value

    def remove_matrix(self, mid):
        check(lib().tc_server_remove_matrix(self.handle, mid), "server_remove_matrix")

    def submit(self, mid, X, Y):
        k = 1 if X.dim() == 1 else X.shape[1]
        t = C.c_ulonglong()
        check(lib().tc_server_submit(self.handle, mid, X.data_ptr(),
                                     1 if X.dim() == 1 else X.stride(0), Matrix._dtype_of(X),
                                     Y.data_ptr(), 1 if Y.dim() == 1 else Y.stride(0), k,
                                     C.byref(t)), "server_submit")
        return t.value

    def wait(self, ticket):
        lat = C.c_double()
        check(lib().tc_server_wait(self.handle, ticket, C.byref(lat)), "server_wait")
        return lat.value

    def stats(self):
        s = ServerStats()
        check(lib().tc_server_get_stats(self.handle, C.byref(s)), "server_stats")
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def _spec(self, items, ybufs, ldy, rate, concurrency, count, seed):
        arr = (LoadItem * len(items))(*[LoadItem(*it) for it in items])
        yb = (C.c_void_p * len(ybufs))(*[y.data_ptr() for y in ybufs])
        spec = LoadSpec(arr, len(items), yb, len(ybufs), ldy, rate, concurrency, count, seed)
        return spec, (arr, yb)

    def run_load(self, items, ybufs, ldy, count, rate=0.0, concurrency=256, seed=7):
        """items: [(matrix_id, X_ptr, ldx, dtype, k)]; returns the result dict."""
        spec, keep = self._spec(items, ybufs, ldy, rate, concurrency, count, seed)
        res = LoadResult()
        check(lib().tc_server_run_load(self.handle, C.byref(spec), C.byref(res)), "run_load")
        del keep
        return {f: getattr(res, f) for f, _ in LoadResult._fields_}

    def run_plain(self, items, ybufs, ldy, count, seed=7):
        spec, keep = self._spec(items, ybufs, ldy, 0.0, 1, count, seed)
        res = LoadResult()
        check(lib().tc_server_run_plain(self.handle, C.byref(spec), C.byref(res)), "run_plain")
        del keep
        return {f: getattr(res, f) for f, _ in LoadResult._fields_}

    def close(self):
        if getattr(self, "handle", None):
            lib().tc_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
