// Host API of the SpMV (include/taichi_amd.h): resident matrices (upload,
// footprint, stats), launch configuration and the autotuner, tc_spmv with
// column chunking, batched and graph-captured launches, and the one-shot
// host-buffer taichi_SpMV. The kernels live in spmv_k1.hip, spmv_wide.hip,
// spmv_staged.hip and spmv_batch.hip.
#include <chrono>
#include <cstring>
#include <vector>

#include "batch.h"
#include "spmv_host.h"


namespace tcd {

int g_cu_count = 0;

int cu_count()
{
    if (g_cu_count == 0) {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            g_cu_count = prop.multiProcessorCount;
        else
            g_cu_count = 256;
    }
    return g_cu_count;
}

// k = 1: tiles of 512 rows and a 4096-value stage; k > 1: wide-rows tiles
// of one row per lane group (full occupancy): the best or near-best settings
// measured on R-MAT, banded and stencil matrices. Matrices whose remainder
// averages more than 8 entries a row, or whose remainder entries and
// diagonal terms together average more than 24 a row (wide bands), use
// 256-row tiles at k = 1: a tile's entries pass through the stage in about
// two rounds, and the tiles are short enough to spread the band's ends.
// Matrices below 2^18 rows also use 256-row tiles at k = 1: twice the
// workgroups for a launch that cannot fill the GPU (n = 16,384, 5 diagonals
// and 16,384 uniform entries: 3.5 against 4.1 us).
// Two rows per lane group (a config choice) measured up to 17 % slower than
// one on bands of 4 to 57 diagonals at k = 4..8 (57 diagonals of 923k rows,
// BF16 k = 6: 65 against 54 us), 2-3 % faster only for FP32 k = 6 on two.
tc_config default_config(int K, int esize, long long remainder, long long diag_terms, int M)
{
    tc_config c{};
    const bool band = M > 0 && remainder + diag_terms > 24ll * M;
    const bool heavy = M > 0 && (remainder > 8ll * M || band);
    const bool small = M < (1 << 18);
    c.rows_per_tile = K == 1      ? (heavy || small ? 256 : 512)
                      : wide_k(K) ? wide_rows(K, 1, esize)
                                  : (kWG * 2) / K;
    c.lds_values = 4096;
    // one workgroup per tile, except FP32 x at k = 64: a persistent grid of 8
    // workgroups per CU (2048 x 2048 stencil: 0.563 against 0.593 ms; with
    // BF16 x it is slower, 0.503 against 0.435 ms)
    c.grid = K == 64 && esize == 4 ? 8 * cu_count() : 0;
    // k = 1: one row per lane (quad rows measured no faster on R-MAT, and
    // slower on the 2048^2 stencil: 35.7 against 28.5 us FP32, 34.9 against
    // 31.0 us BF16); the autotuner tries both
    c.flags = K == 1 ? kQuadOff : 0;
    return c;
}

tc_config default_config(const tc_matrix *A, int K, int dtype)
{
    return default_config(K, elem_size(dtype), A->st.remainder, A->diag_terms, A->v.M);
}

size_t max_lds_bytes()
{
    static std::mutex mu;
    static std::map<int, size_t> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 65536;
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(dev);
    if (it != cache.end()) return it->second;
    hipDeviceProp_t prop;
    const size_t v = hipGetDeviceProperties(&prop, dev) == hipSuccess ? (size_t)prop.sharedMemPerBlock : 65536;
    cache[dev] = v;
    return v;
}

bool rows_per_tile_ok(int K, int r, int flags, int esize)
{
    if (K == 1) return r == 256 || r == 512 || r == 1024 || r == 2048 || r == 4096;
    if (wide_k(K) && !(flags & kStagedK)) return r == wide_rows(K, 1, esize) || r == wide_rows(K, 2, esize);
    for (int P : {1, 2, 4})
        if ((kWG * P) / K == r) return true;
    return false;
}

tc_config get_cfg(tc_matrix *A, int K, int dtype)
{
    std::lock_guard<std::mutex> g(A->mu);
    auto it = A->cfg.find({K, dtype});
    return it != A->cfg.end() ? it->second : default_config(A, K, dtype);
}


// one launch of K columns (K a supported count) with an explicit config
int launch_cfg(tc_matrix *A, const void *X, int ldx, int dtype, float *Y, int ldy, int K, const tc_config &c,
               hipStream_t s)
{
    const int capv = std::max(c.lds_values, 64 * K);
    if (K == 1 && ldx == 1 && ldy == 1 && quad_applies(A->v, X, Y, elem_size(dtype), c, capv)) {
        return launch_quad(A->v, X, Y, dtype, c, capv, s);
    }
    if (K == 1 && ((uintptr_t)Y & 15) == 0) {
        return launch_rows(A->v, X, ldx, Y, ldy, dtype, c, capv, s);
    }
    if (K > 1 && wide_applies(K, X, ldx, Y, ldy, elem_size(dtype), c)) {
        return launch_wide(A->v, X, ldx, Y, ldy, dtype, K, c, s);
    }
    // staged tile kernel: K=1 configs name the rows kernel's tile; use P = 2
    const int P = K == 1 ? 2 : p_for(K, c.rows_per_tile);
    return launch_staged(A->v, X, ldx, Y, ldy, dtype, K, P, capv, c.grid, s);
}

int check_spmv_args(const tc_matrix *A, const void *X, int ldx, int dtype, const float *Y, int ldy, int k)
{
    if (!A) return tc::fail(TC_ERR_INVALID, "tc_spmv: null matrix");
    if (dtype != TC_F32 && dtype != TC_BF16 && dtype != TC_F16) return tc::fail(TC_ERR_INVALID, "bad dtype %d", dtype);
    if (k < 1 || ldx < k || ldy < k) return tc::fail(TC_ERR_INVALID, "bad k=%d ldx=%d ldy=%d", k, ldx, ldy);
    if ((A->v.N > 0 && !X) || (A->v.M > 0 && !Y)) return tc::fail(TC_ERR_INVALID, "null X or Y");
    return TC_OK;
}

// Descriptor buffers of tc_spmv_batch: kDescSlots slots per device, each a
// pinned host buffer (the copy's source) and a device buffer (what the
// kernels read), used in turn. A slot is refilled once the event recorded
// after its batch's kernels has passed, so a call never waits for its
// stream; only a caller with more than kDescSlots batches in flight waits,
// for the oldest. (A stream-ordered allocation per call instead cost about
// 0.3 ms a call: the default pool maps its memory again after each release.)
constexpr int kDescSlots = 8;
struct DescSlot {
    void *host = nullptr;
    void *dev = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;
};
struct DescRing {
    std::mutex mu;
    DescSlot slot[kDescSlots];
    int next = 0;
};
DescRing &desc_ring(int dev)
{
    static std::mutex mu;
    static std::map<int, DescRing *> rings;  // one per device, kept until the process ends
    std::lock_guard<std::mutex> g(mu);
    DescRing *&r = rings[dev];
    if (!r) r = new DescRing();
    return *r;
}

}  // namespace tcd

using namespace tcd;

extern "C" {

int tc_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int tc_matrix_create(int M, int N, const TaiChi_new *T, long long nnz, int long_threshold, void *stream,
                     tc_matrix **out)
try {
    if (!T || !out || M < 0 || N < 0) return tc::fail(TC_ERR_INVALID, "tc_matrix_create: bad argument");
    *out = nullptr;
    Layout P;
    int rc = plan_layout(M, N, T, long_threshold, P);
    if (rc != TC_OK) return rc;

    // one device allocation, 256-byte aligned sub-arrays
    std::vector<LayoutPart> parts = layout_parts(P, P.S > 0 ? T->neg_offsets : nullptr);
    std::vector<size_t> off(parts.size());
    size_t total = 0, fmt_bytes = 0;
    for (size_t i = 0; i < parts.size(); i++) {
        off[i] = total;
        total += (parts[i].bytes + 255) & ~(size_t)255;
        fmt_bytes += parts[i].bytes;
    }
    tc_matrix *A = new tc_matrix();
    const hipError_t de = hipGetDevice(&A->device);
    if (de != hipSuccess) {
        delete A;
        return tc::fail(TC_ERR_NODEVICE, "tc_matrix_create: no HIP device (hipGetDevice: %s)", hipGetErrorString(de));
    }
    hipStream_t s = (hipStream_t)stream;
    if (total > 0) {
        hipError_t e = hipMalloc(&A->mem, total);
        if (e != hipSuccess) {
            delete A;
            return tc::fail(TC_ERR_NOMEM, "tc_matrix_create: hipMalloc(%zu): %s", total, hipGetErrorString(e));
        }
        for (size_t i = 0; i < parts.size(); i++)
            if (parts[i].bytes) {
                e = hipMemcpyAsync((char *)A->mem + off[i], parts[i].src, parts[i].bytes, hipMemcpyHostToDevice, s);
                if (e != hipSuccess) {
                    (void)hipFree(A->mem);
                    delete A;
                    return tc::fail(TC_ERR_HIP, "tc_matrix_create: upload: %s", hipGetErrorString(e));
                }
            }
        hipError_t se = hipStreamSynchronize(s);
        if (se != hipSuccess) {
            (void)hipFree(A->mem);
            delete A;
            return tc::fail(TC_ERR_HIP, "tc_matrix_create: sync: %s", hipGetErrorString(se));
        }
    }
    auto at = [&](int i) -> const void * { return parts[i].bytes ? (const char *)A->mem + off[i] : nullptr; };
    A->mem_bytes = total;
    MatView &v = A->v;
    v.M = M;
    v.N = N;
    v.S = P.S;
    v.L = P.L;
    v.ow = P.ow;
    v.ngroups = P.G;
    // the wide-rows tiles stage the x rows around each tile for the remainder
    // when most entries, and at least kNearMin a row, lie near their row:
    // staging 1.5 x rows a row measured slower at 12 and fewer (a band with
    // 37 near entries a row: 19-28 % faster at k = 4 and 8)
    v.near = P.near * 2 > P.R && P.near >= (long long)kNearMin * M;
    v.shapes = (const ShapeDev *)at(0);
    v.offs = (const int *)at(1);
    if (P.S > 0) v.sh0 = P.shapes[0];
    v.gbase = (const int *)at(2);
    v.roff = at(3);
    v.cols = (const int *)at(4);
    v.glong = (const int *)at(5);
    v.lrow = (const int *)at(6);
    v.lptr = (const int *)at(7);
    v.pres = (const uint32_t *)at(8);
    tc_matrix_stats &st = A->st;
    st.M = M;
    st.N = N;
    st.dia_shapes = P.S;
    st.num_negs = P.D;
    st.long_rows = P.L;
    st.offset_bytes = P.ow;
    st.nnz = nnz;
    st.remainder = P.R;
    st.format_bytes = (long long)fmt_bytes;
    st.device_bytes = (long long)total;
    st.holes = P.holes;
    int pm_shapes = 0;
    for (const ShapeDev &h : P.shapes) {
        pm_shapes += h.pm >= 0;
        A->diag_terms += (long long)(h.oe - h.ob) * (h.xe - h.xs);
    }
    st.bitmap_shapes = pm_shapes;
    st.near_window = v.near;
    *out = A;
    return TC_OK;
} catch (...) {
    return tc::exception_status(__func__);
}

int tc_matrix_footprint(int M, int N, const TaiChi_new *T, int long_threshold, long long *bytes)
try {
    if (!T || !bytes || M < 0 || N < 0 || (M > 0 && !T->csrRow))
        return tc::fail(TC_ERR_INVALID, "tc_matrix_footprint: bad argument");
    Layout P;
    int rc = plan_layout(M, N, T, long_threshold, P);
    if (rc != TC_OK) return rc;
    long long total = 0;
    for (const LayoutPart &p : layout_parts(P, nullptr)) total += ((long long)p.bytes + 255) & ~255ll;
    *bytes = total;
    return TC_OK;
} catch (...) {
    return tc::exception_status(__func__);
}

int tc_matrix_destroy(tc_matrix *A)
{
    if (!A) return TC_OK;
    if (A->mem) (void)hipFree(A->mem);
    delete A;
    return TC_OK;
}

int tc_matrix_get_stats(const tc_matrix *A, tc_matrix_stats *out)
try {
    if (!A || !out) return tc::fail(TC_ERR_INVALID, "tc_matrix_get_stats: null argument");
    *out = A->st;
    return TC_OK;
} catch (...) {
    return tc::exception_status(__func__);
}

long long tc_spmv_bytes(const tc_matrix *A, int k, int dtype)
{
    if (!A || k < 1) return -1;
    return A->st.format_bytes + (long long)elem_size(dtype) * A->st.N * k + 4ll * A->st.M * k;
}

int tc_spmv(tc_matrix *A, const void *X, int ldx, int dtype, float *Y, int ldy, int k, void *stream)
try {
    int rc = check_spmv_args(A, X, ldx, dtype, Y, ldy, k);
    if (rc != TC_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    const size_t es = elem_size(dtype);
    // columns in chunks of supported widths; each column's result does not
    // depend on the chunking
    for (int c0 = 0; c0 < k;) {
        const int K = pick_k(k - c0);
        const void *xc = (const char *)X + c0 * es;
        float *yc = Y + c0;
        const int W = masked_width(K);
        if (W) {
            const tc_config cw = get_cfg(A, W, dtype);
            // (a masked pass reads X's last row through the row above)
            if (!(cw.flags & kNoMask) && A->v.N >= 2 && wide_applies(W, xc, ldx, yc, ldy, (int)es, cw)) {
                rc = launch_wide(A->v, xc, ldx, yc, ldy, dtype, W, cw, s, K);
                if (rc != TC_OK) return rc;
                c0 += K;
                continue;
            }
        }
        const tc_config c = get_cfg(A, K, dtype);
        rc = launch_cfg(A, xc, ldx, dtype, yc, ldy, K, c, s);
        if (rc != TC_OK) return rc;
        c0 += K;
    }
    return TC_OK;
} catch (...) {
    return tc::exception_status(__func__);
}

int tc_matrix_set_config(tc_matrix *A, int k, int dtype, const tc_config *cfg)
try {
    if (!A || !cfg || k < 1) return tc::fail(TC_ERR_INVALID, "tc_matrix_set_config: bad argument");
    const int K = pick_k(k);
    tc_config c = *cfg;
    const tc_config d = default_config(A, K, dtype);
    if (c.rows_per_tile <= 0) c.rows_per_tile = d.rows_per_tile;
    if (c.lds_values <= 0) c.lds_values = d.lds_values;
    if (!rows_per_tile_ok(K, c.rows_per_tile, c.flags, elem_size(dtype)))
        return tc::fail(TC_ERR_INVALID, "rows_per_tile %d not available for k=%d flags=%d", c.rows_per_tile, K, c.flags);
    if (c.lds_values < 512 || c.lds_values > 8192)
        return tc::fail(TC_ERR_INVALID, "lds_values %d outside [512, 8192]", c.lds_values);
    if (c.grid < 0) return tc::fail(TC_ERR_INVALID, "grid < 0");
    std::lock_guard<std::mutex> g(A->mu);
    A->cfg[{K, dtype}] = c;
    return TC_OK;
} catch (...) {
    return tc::exception_status(__func__);
}

int tc_matrix_get_config(const tc_matrix *A, int k, int dtype, tc_config *cfg)
try {
    if (!A || !cfg || k < 1) return tc::fail(TC_ERR_INVALID, "tc_matrix_get_config: bad argument");
    *cfg = get_cfg(const_cast<tc_matrix *>(A), pick_k(k), dtype);
    return TC_OK;
} catch (...) {
    return tc::exception_status(__func__);
}

int tc_autotune(tc_matrix *A, const void *X, int ldx, int dtype, float *Y, int ldy, int k, void *stream,
                tc_config *best, float *best_us)
try {
    int rc = check_spmv_args(A, X, ldx, dtype, Y, ldy, k);
    if (rc != TC_OK) return rc;
    const int K = pick_k(k);
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    const int ncu = cu_count();
    std::vector<tc_config> cands;
    std::vector<int> tiles;
    // (rows_per_tile, flags) pairs: both k = 1 kernels; wide and staged for k > 1
    std::vector<std::pair<int, int>> tf;
    if (K == 1) {
        for (int r : {256, 512, 1024, 2048, 4096})
            for (int f : {0, kQuadOff}) tf.push_back({r, f});
    } else {
        if (wide_k(K))
            for (int RI : {1, 2}) tf.push_back({wide_rows(K, RI, elem_size(dtype)), 0});
        for (int P : {1, 2, 4}) tf.push_back({(kWG * P) / K, kStagedK});
    }
    for (auto [rpt, flags] : tf)
        for (int capv : {2048, 4096, 8192})
            for (int grid : {0, 4 * ncu, 8 * ncu}) {
                if (capv < 64 * K) continue;
                if (K == 1 && (flags & kQuadOff) == 0 && rpt != 512 && rpt != 1024) continue;
                if (K > 1 && !(flags & kStagedK) && capv != 4096) continue;  // the wide kernel stages nothing
                tc_config c{rpt, capv, grid, flags};
                cands.push_back(c);
            }
    tc_config bc = default_config(A, K, dtype);
    float bt = 1e30f;
    const int reps = 20;
    for (const tc_config &c : cands) {
        for (int w = 0; w < 3; w++)
            if ((rc = launch_cfg(A, X, ldx, dtype, Y, ldy, K, c, s)) != TC_OK) break;
        if (rc != TC_OK) break;
        HIPCHECK(hipEventRecord(e0, s));
        for (int i = 0; i < reps; i++) launch_cfg(A, X, ldx, dtype, Y, ldy, K, c, s);
        HIPCHECK(hipEventRecord(e1, s));
        HIPCHECK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
        const float us = ms * 1000.0f / reps;
        if (us < bt) {
            bt = us;
            bc = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != TC_OK) return rc;
    {
        std::lock_guard<std::mutex> g(A->mu);
        A->cfg[{K, dtype}] = bc;
    }
    if (best) *best = bc;
    if (best_us) *best_us = bt;
    return TC_OK;
} catch (...) {
    return tc::exception_status(__func__);
}

int tc_spmv_batch(const tc_request *reqs, int n, void *stream)
try {
    if (n < 0 || (n > 0 && !reqs)) return tc::fail(TC_ERR_INVALID, "tc_spmv_batch: bad argument");
    if (n == 0) return TC_OK;
    hipStream_t s = (hipStream_t)stream;
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    DescRing &R = desc_ring(dev);
    std::lock_guard<std::mutex> g(R.mu);
    DescSlot &sl = R.slot[R.next];
    R.next = (R.next + 1) % kDescSlots;
    if (sl.pending) {  // the copy that last read this slot
        HIPCHECK(hipEventSynchronize(sl.ev));
        sl.pending = false;
    }
    const size_t need = tc::batch_desc_bytes(n);
    if (sl.cap < need) {
        if (sl.host) (void)hipHostFree(sl.host);
        if (sl.dev) (void)hipFree(sl.dev);
        sl.host = sl.dev = nullptr;
        sl.cap = 0;
        if (hipHostMalloc(&sl.host, need, hipHostMallocDefault) != hipSuccess ||
            hipMalloc(&sl.dev, need) != hipSuccess)
            return tc::fail(TC_ERR_NOMEM, "tc_spmv_batch: %zu B of descriptor memory", need);
        sl.cap = need;
    }
    if (!sl.ev) HIPCHECK(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming));
    const int rc = tc::batch_enqueue(reqs, n, sl.host, sl.dev, s, need);
    HIPCHECK(hipEventRecord(sl.ev, s));
    sl.pending = true;
    return rc;
} catch (...) {
    return tc::exception_status(__func__);
}

int tc_graph_capture(tc_matrix *A, const void *X, int ldx, int dtype, float *Y, int ldy, int k, int iters,
                     void *stream, tc_graph **out)
try {
    int rc = check_spmv_args(A, X, ldx, dtype, Y, ldy, k);
    if (rc != TC_OK) return rc;
    if (!out || iters < 1) return tc::fail(TC_ERR_INVALID, "tc_graph_capture: bad argument");
    hipStream_t cs = (hipStream_t)stream;
    bool own = false;
    if (!cs) {
        HIPCHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        own = true;
    }
    tc_graph *g = new tc_graph();
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        for (int i = 0; i < iters && rc == TC_OK; i++) rc = tc_spmv(A, X, ldx, dtype, Y, ldy, k, cs);
        hipError_t e2 = hipStreamEndCapture(cs, &g->graph);
        if (e == hipSuccess) e = e2;
    }
    if (e == hipSuccess && rc == TC_OK) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (own) (void)hipStreamDestroy(cs);
    if (e != hipSuccess || rc != TC_OK) {
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g;
        return rc != TC_OK ? rc : tc::fail(TC_ERR_HIP, "tc_graph_capture: %s", hipGetErrorString(e));
    }
    *out = g;
    return TC_OK;
} catch (...) {
    return tc::exception_status(__func__);
}

int tc_graph_launch(tc_graph *g, void *stream)
try {
    if (!g) return tc::fail(TC_ERR_INVALID, "tc_graph_launch: null graph");
    HIPCHECK(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return TC_OK;
} catch (...) {
    return tc::exception_status(__func__);
}

int tc_graph_destroy(tc_graph *g)
{
    if (!g) return TC_OK;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return TC_OK;
}

int tc_taichi_spmv(int M, int N, const TaiChi_new *T, const float *x_host, float *y_host, int warmup, int runs,
                   float *transfer_ms, float *spmv_ms)
try {
    if (!T || (N > 0 && !x_host) || (M > 0 && !y_host) || warmup < 0 || runs < 0)
        return tc::fail(TC_ERR_INVALID, "tc_taichi_spmv: bad argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return tc::fail(TC_ERR_NODEVICE, "no HIP device");
    auto t0 = std::chrono::steady_clock::now();
    tc_matrix *A = nullptr;
    long long nnz = (long long)T->dense_nnz + T->sparse_nnz;
    int rc = tc_matrix_create(M, N, T, nnz, 0, nullptr, &A);
    if (rc != TC_OK) return rc;
    auto t1 = std::chrono::steady_clock::now();
    if (transfer_ms) *transfer_ms = std::chrono::duration<float, std::milli>(t1 - t0).count();
    float *dx = nullptr, *dy = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        if (dx) (void)hipFree(dx);
        if (dy) (void)hipFree(dy);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        tc_matrix_destroy(A);
    };
#define STEP(expr)                                                                                          \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess) {                                                                             \
            cleanup();                                                                                      \
            return tc::fail(TC_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));                            \
        }                                                                                                   \
    } while (0)
    STEP(hipMalloc((void **)&dx, (size_t)std::max(N, 1) * sizeof(float)));
    STEP(hipMalloc((void **)&dy, (size_t)std::max(M, 1) * sizeof(float)));
    if (N > 0) STEP(hipMemcpy(dx, x_host, (size_t)N * sizeof(float), hipMemcpyHostToDevice));
    STEP(hipEventCreate(&e0));
    STEP(hipEventCreate(&e1));
    for (int i = 0; i < std::max(warmup, 1) && rc == TC_OK; i++) rc = tc_spmv(A, dx, 1, TC_F32, dy, 1, 1, nullptr);
    STEP(hipEventRecord(e0, nullptr));
    for (int i = 0; i < runs && rc == TC_OK; i++) rc = tc_spmv(A, dx, 1, TC_F32, dy, 1, 1, nullptr);
    STEP(hipEventRecord(e1, nullptr));
    STEP(hipEventSynchronize(e1));
    if (rc != TC_OK) {
        cleanup();
        return rc;
    }
    float ms = 0;
    STEP(hipEventElapsedTime(&ms, e0, e1));
    if (spmv_ms) *spmv_ms = runs > 0 ? ms / runs : 0.0f;
    if (M > 0) STEP(hipMemcpy(y_host, dy, (size_t)M * sizeof(float), hipMemcpyDeviceToHost));
#undef STEP
    cleanup();
    return TC_OK;
} catch (...) {
    return tc::exception_status(__func__);
}

}  // extern "C"
