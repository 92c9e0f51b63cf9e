// Batched SpMV over many requests on any mix of resident matrices: the
// requests are cut into column chunks, grouped by (kernel, chunk width, x
// type), and each group is one launch over all its tiles.
#include "batch.h"
#include "spmv_host.h"

namespace tcd {

// ---- batched launches over many requests ----
// A batch is cut into column chunks of K in {8, 4, 2, 1} columns and the
// chunks are grouped by (K, x type); each group is one launch of a kernel
// specialised for it, so no launch carries the registers of the others. The
// tiles of a group are numbered consecutively, chunk by chunk; a workgroup
// finds its chunk by binary search over the group's tile starts (staged in
// LDS) and reads the chunk's descriptor with scalar loads. Column chunks
// change nothing in the result: the per-(row, column) order is independent
// of K.
constexpr int kBatchP = 2;       // K > 1: run_tile with P = 2 (512/K rows a tile)
// K = 1: rows_tile with 1024 rows a tile when the group holds at least
// kBatchBigRows rows (four thousand tiles: two full rounds of the GPU), else
// 512. On 256 uneven R-MAT matrices (16 M rows) 1024-row tiles take 70 us a
// batch, against 101 us with 512 rows and 85 us with 2048; a 2048- or
// 8192-value stage instead of 4096: 69 and 106 us. A lone request keeps
// 512-row tiles, twice the workgroups: at 1024 the served mix's p99 at 20k
// req/s went from 102-112 to 132-134 us.
constexpr int kBatchRPT = 2;
constexpr int kBatchRPTBig = 4;
constexpr long long kBatchBigRows = 4ll << 20;
constexpr int kBatchCapv = 4096; // staged values per pass
constexpr int kBatchLdsDesc = 512;

struct WorkDesc {
    MatView A;
    const void *X;
    float *Y;
    int ldx, ldy;
    int unit_begin;
    int fast;  // one shape with one offset: the rows tile's FAST path
    int kv;    // columns that exist (a masked wide chunk has kv < its width)
};

__device__ __forceinline__ WorkDesc load_desc(const WorkDesc *D, int d)
{
    constexpr int NW = sizeof(WorkDesc) / 4;
    union {
        WorkDesc w;
        int u[NW];
    } t;
    const int *p = (const int *)(D + d);
#pragma unroll
    for (int i = 0; i < NW; i++) t.u[i] = cload(p, i);
    return t.w;
}

// Runs f(desc, local_tile) for every tile of this workgroup's range of the
// group's `total` tiles. UB[i] is chunk i's first tile (ascending).
template <typename F>
__device__ __forceinline__ void batch_loop(const WorkDesc *__restrict__ D, const int *__restrict__ UB, int nd,
                                           int total, int *s_ub, F &&f)
{
    int ub, ue;
    block_units(total, &ub, &ue);
    if (ub >= ue) return;
    const bool in_lds = nd <= kBatchLdsDesc;
    if (in_lds) {
        for (int i = threadIdx.x; i < nd; i += kWG) s_ub[i] = UB[i];
        __syncthreads();
    }
    int lo = 0, hi = nd - 1;  // last chunk with UB <= ub
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((in_lds ? s_ub[mid] : cload(UB, mid)) <= ub) lo = mid;
        else hi = mid - 1;
    }
    int d = __builtin_amdgcn_readfirstlane(lo);
    for (int u = ub; u < ue; u++) {
        while (d + 1 < nd && (in_lds ? s_ub[d + 1] : cload(UB, d + 1)) <= u) d++;
        d = __builtin_amdgcn_readfirstlane(d);
        const WorkDesc w = load_desc(D, d);
        f(w, u - w.unit_begin);
        __syncthreads();
    }
}

template <typename XT, int RPT>
__global__ __launch_bounds__(kWG) void batch_rows_kernel(const WorkDesc *__restrict__ D, const int *__restrict__ UB,
                                                         int nd, int total, int capv)
{
    extern __shared__ float smem[];
    __shared__ int s_ub[kBatchLdsDesc];
    batch_loop(D, UB, nd, total, s_ub, [&](const WorkDesc &w, int t) {
        const XT *X = (const XT *)w.X;
        const bool lng = w.A.L > 0;
        if (w.A.ow == 1) {
            if (w.fast) lng ? rows_tile<XT, RPT, 1, true, true, false>(w.A, X, w.ldx, w.Y, w.ldy, t, smem, capv)
                            : rows_tile<XT, RPT, 1, true, false, false>(w.A, X, w.ldx, w.Y, w.ldy, t, smem, capv);
            else lng ? rows_tile<XT, RPT, 1, false, true, false>(w.A, X, w.ldx, w.Y, w.ldy, t, smem, capv)
                     : rows_tile<XT, RPT, 1, false, false, false>(w.A, X, w.ldx, w.Y, w.ldy, t, smem, capv);
        } else {
            if (w.fast) lng ? rows_tile<XT, RPT, 2, true, true, false>(w.A, X, w.ldx, w.Y, w.ldy, t, smem, capv)
                            : rows_tile<XT, RPT, 2, true, false, false>(w.A, X, w.ldx, w.Y, w.ldy, t, smem, capv);
            else lng ? rows_tile<XT, RPT, 2, false, true, false>(w.A, X, w.ldx, w.Y, w.ldy, t, smem, capv)
                     : rows_tile<XT, RPT, 2, false, false, false>(w.A, X, w.ldx, w.Y, w.ldy, t, smem, capv);
        }
    });
}

template <int K, typename XT>
__global__ __launch_bounds__(kWG) void batch_tile_kernel(const WorkDesc *__restrict__ D, const int *__restrict__ UB,
                                                         int nd, int total, int capv)
{
    constexpr int TRW = (kWG * kBatchP) / K;
    extern __shared__ float smem[];
    __shared__ int s_ub[kBatchLdsDesc];
    float *s_val = smem;
    acc_t *s_lsum = (acc_t *)(s_val + lsum_at(capv));
    int *s_start = (int *)(s_lsum + kWG * kBatchP);
    batch_loop(D, UB, nd, total, s_ub, [&](const WorkDesc &w, int t) {
        run_tile<K, XT, kBatchP>(w.A, (const XT *)w.X, w.ldx, w.Y, w.ldy, t * TRW, s_val, capv, s_start, s_lsum);
    });
}

// K > 1 chunks whose X and Y rows split into aligned vectors: wide-rows tiles
constexpr int kBatchRI = 1;  // one row per lane group: 39-75 VGPRs

// Waves a SIMD the compiler must fit the tiles into (scalar registers beyond
// 96 alone cost these kernels a wave): 8, or 7 at 8 columns; masked 8-column
// FP32 tiles keep the compiler's choice (a bound there spills to scratch).
constexpr int batch_wide_waves(int K, bool MASK, int esize) { return K < 8 ? 8 : MASK && esize == 4 ? 1 : 7; }

template <int K, typename XT, bool MASK>
__global__ __launch_bounds__(kWG, batch_wide_waves(K, MASK, (int)sizeof(XT))) void batch_wide_kernel(
    const WorkDesc *__restrict__ D, const int *__restrict__ UB, int nd, int total)
{
    extern __shared__ float smem[];
    __shared__ int s_ub[kBatchLdsDesc];
    batch_loop(D, UB, nd, total, s_ub, [&](const WorkDesc &w, int t) {
        if (w.A.L > 0) wide_tile<K, XT, kBatchRI, true, MASK>(w.A, (const XT *)w.X, w.ldx, w.Y, w.ldy, t, smem, w.kv);
        else wide_tile<K, XT, kBatchRI, false, MASK>(w.A, (const XT *)w.X, w.ldx, w.Y, w.ldy, t, smem, w.kv);
    });
}

}  // namespace tcd

namespace tc {

using namespace tcd;

// chunks of one request of k columns: 8s, then 4, 2, 1 as needed
int chunk_count(int k) { return k / 8 + ((k & 4) != 0) + ((k & 2) != 0) + (k & 1); }

constexpr int kMaxBatchK = 64;

size_t batch_desc_bytes(int max_requests)
{
    const size_t chunks = (size_t)max_requests * chunk_count(kMaxBatchK - 1);
    return chunks * sizeof(WorkDesc) + chunks * sizeof(int) + 16;
}

// Batch classes: (kernel, chunk width, x type). Kernel 0: rows tiles (K = 1);
// 1: wide-rows tiles (K > 1, aligned); 2: staged tiles (K > 1, otherwise);
// 3: masked wide-rows tiles (fewer columns than the width, aligned).
constexpr int kBatchClasses = 4 * 4 * 3;
inline int batch_class(int kern, int ki, int dt) { return (kern * 4 + ki) * 3 + dt; }

// big: a K = 1 group of at least kBatchBigRows rows
int batch_tile_rows(int kern, int K, int esize, bool big)
{
    if (kern == 0) return kWG * (big ? kBatchRPTBig : kBatchRPT);
    if (kern == 1 || kern == 3) return wide_rows(K, kBatchRI, esize);
    return (kWG * kBatchP) / K;
}

template <typename XT>
void launch_batch_class(int kern, int K, const WorkDesc *D, const int *UB, int nd, int units, bool big,
                        hipStream_t s)
{
    // one workgroup per tile: the dispatcher balances chunks of any size (a
    // grid of 8 workgroups per CU walking runs of tiles measured slower)
    const int grid = units;
    if (kern == 0) {
        const int trw = kWG * (big ? kBatchRPTBig : kBatchRPT);
        const size_t lds = ((size_t)lsum_at(kBatchCapv) + 2 * trw) * 4;  // s_val | long sums (2*TRW floats reserved)
        if (big) batch_rows_kernel<XT, kBatchRPTBig><<<grid, kWG, lds, s>>>(D, UB, nd, units, kBatchCapv);
        else batch_rows_kernel<XT, kBatchRPT><<<grid, kWG, lds, s>>>(D, UB, nd, units, kBatchCapv);
    } else if (kern == 1 || kern == 3) {
        const int trw = wide_rows(K, kBatchRI, (int)sizeof(XT));
        const size_t lds = wide_lds_bytes(K, trw, false, true);  // batched tiles run without the near window
        if (kern == 3) {
            if (K == 2) batch_wide_kernel<2, XT, true><<<grid, kWG, lds, s>>>(D, UB, nd, units);
            else if (K == 4) batch_wide_kernel<4, XT, true><<<grid, kWG, lds, s>>>(D, UB, nd, units);
            else batch_wide_kernel<8, XT, true><<<grid, kWG, lds, s>>>(D, UB, nd, units);
        } else {
            if (K == 2) batch_wide_kernel<2, XT, false><<<grid, kWG, lds, s>>>(D, UB, nd, units);
            else if (K == 4) batch_wide_kernel<4, XT, false><<<grid, kWG, lds, s>>>(D, UB, nd, units);
            else batch_wide_kernel<8, XT, false><<<grid, kWG, lds, s>>>(D, UB, nd, units);
        }
    } else {
        const size_t lds = lds_bytes(kBatchCapv, kBatchP, (kWG * kBatchP) / K);
        if (K == 2) batch_tile_kernel<2, XT><<<grid, kWG, lds, s>>>(D, UB, nd, units, kBatchCapv);
        else if (K == 4) batch_tile_kernel<4, XT><<<grid, kWG, lds, s>>>(D, UB, nd, units, kBatchCapv);
        else batch_tile_kernel<8, XT><<<grid, kWG, lds, s>>>(D, UB, nd, units, kBatchCapv);
    }
}

// chunk c0..c0+K of request r (kv of its columns exist): its batch class
bool chunk_aligned(const tc_request &r, int K, int c0)
{
    const int es = elem_size(r.dtype), cpl = wide_cpl(K, es);
    const uintptr_t xa = (uintptr_t)r.X + (uintptr_t)c0 * es, ya = (uintptr_t)(r.Y + c0);
    return r.ldx % cpl == 0 && r.ldy % cpl == 0 && xa % ((uintptr_t)cpl * es) == 0 &&
           ya % ((uintptr_t)cpl * 4) == 0;
}

int chunk_class(const tc_request &r, int ki, int K, int c0, int dt, int kv)
{
    if (kv < K) return batch_class(3, ki, dt);  // masked: chosen only when aligned
    if (K == 1) return batch_class(0, ki, dt);
    return batch_class(chunk_aligned(r, K, c0) ? 1 : 2, ki, dt);
}

// Fills the chunk descriptors of n requests into host_desc (at least
// batch_desc_bytes(n) bytes, valid until the copy on s has run), copies them
// to dev_desc and enqueues one launch per batch class present. Each
// request's y equals what tc_spmv computes for it alone.
int batch_enqueue(const tc_request *reqs, int n, void *host_desc, void *dev_desc, hipStream_t s, size_t cap)
{
    static const int kChunkK[4] = {1, 2, 4, 8};
    // chunks of a request: 8s; then the rest (1..7 columns) as one masked
    // chunk of the next width when its rows are aligned and the mask pays,
    // else 4, 2, 1 as its bits say. Calls f(ki, K, c0, kv) for each, widest
    // first.
    auto chunks = [](const tc_request &r, auto &&f) {
        const int k = r.k;
        int c0 = 0;
        while (k - c0 >= 8) {
            f(3, 8, c0, 8);
            c0 += 8;
        }
        const int rest = k - c0;
        // 3, 5, 6 or 7 columns run as one masked pass of width 4 or 8 (a
        // lane owns all 8 columns of a row, for every x type)
        const int W = rest > 0 ? masked_width(rest) : 0;
        if (W && r.A->v.N >= 2 && chunk_aligned(r, W, c0) && !(get_cfg(r.A, W, r.dtype).flags & kNoMask)) {
            f(W == 2 ? 1 : W == 4 ? 2 : 3, W, c0, rest);
            return;
        }
        for (int ki = 2; ki >= 0; ki--)
            if ((k - c0) & (1 << ki)) {
                f(ki, 1 << ki, c0, 1 << ki);
                c0 += 1 << ki;
            }
    };
    int cls_count[kBatchClasses] = {0};
    long long cls_rows[kBatchClasses] = {0};
    int nchunks = 0;
    for (int i = 0; i < n; i++) {
        const tc_request &r = reqs[i];
        int rc = check_spmv_args(r.A, r.X, r.ldx, r.dtype, r.Y, r.ldy, r.k);
        if (rc != TC_OK) return rc;
        if (r.k > kMaxBatchK) return tc::fail(TC_ERR_UNSUPPORTED, "batch: k=%d > %d", r.k, kMaxBatchK);
        const int dt = r.dtype == TC_F32 ? 0 : r.dtype == TC_BF16 ? 1 : 2;
        chunks(r, [&](int ki, int K, int c0, int kv) {
            const int cl = chunk_class(r, ki, K, c0, dt, kv);
            cls_count[cl]++;
            cls_rows[cl] += r.A->v.M;
            nchunks++;
        });
    }
    if (nchunks == 0) return TC_OK;
    {
        const size_t need = (((size_t)nchunks * sizeof(WorkDesc) + 15) & ~(size_t)15) + (size_t)nchunks * sizeof(int);
        if (need > cap)
            return tc::fail(TC_ERR_CAPACITY, "batch: %d column chunks need %zu descriptor bytes, %zu available",
                            nchunks, need, cap);
    }
    int cls_begin[kBatchClasses + 1] = {0};
    for (int c = 0; c < kBatchClasses; c++) cls_begin[c + 1] = cls_begin[c] + cls_count[c];
    WorkDesc *D = (WorkDesc *)host_desc;
    const size_t ub_off = ((size_t)nchunks * sizeof(WorkDesc) + 15) & ~(size_t)15;
    int *UB = (int *)((char *)host_desc + ub_off);
    int fill[kBatchClasses];
    long long units[kBatchClasses] = {0};
    for (int c = 0; c < kBatchClasses; c++) fill[c] = cls_begin[c];
    // chunks in request order within each class
    for (int i = 0; i < n; i++) {
        const tc_request &r = reqs[i];
        const int dt = r.dtype == TC_F32 ? 0 : r.dtype == TC_BF16 ? 1 : 2;
        const MatView &v = r.A->v;
        const bool fast = v.S == 1 && v.sh0.oe - v.sh0.ob == 1 && v.sh0.pm < 0;
        chunks(r, [&](int ki, int K, int c0, int kv) {
            const int cl = chunk_class(r, ki, K, c0, dt, kv);
            WorkDesc &w = D[fill[cl]];
            w.A = v;
            w.X = (const char *)r.X + (size_t)c0 * elem_size(r.dtype);
            w.Y = r.Y + c0;
            w.ldx = r.ldx;
            w.ldy = r.ldy;
            w.fast = fast ? 1 : 0;
            w.kv = kv;
            const int trw = batch_tile_rows(cl / 12, K, elem_size(r.dtype), cls_rows[cl] >= kBatchBigRows);
            w.unit_begin = (int)units[cl];
            UB[fill[cl]] = w.unit_begin;
            units[cl] += (v.M + trw - 1) / trw;
            fill[cl]++;
        });
    }
    for (int c = 0; c < kBatchClasses; c++)
        if (units[c] > INT32_MAX) return tc::fail(TC_ERR_UNSUPPORTED, "batch: too many tiles");
    HIPCHECK(hipMemcpyAsync(dev_desc, host_desc, ub_off + (size_t)nchunks * sizeof(int), hipMemcpyHostToDevice, s));
    const WorkDesc *dD = (const WorkDesc *)dev_desc;
    const int *dUB = (const int *)((const char *)dev_desc + ub_off);
    for (int c = 0; c < kBatchClasses; c++) {
        if (units[c] == 0) continue;
        const int kern = c / 12, K = kChunkK[(c / 3) % 4], dt = c % 3, nd = cls_count[c], b = cls_begin[c];
        const bool big = cls_rows[c] >= kBatchBigRows;
        if (dt == 0) launch_batch_class<float>(kern, K, dD + b, dUB + b, nd, (int)units[c], big, s);
        else if (dt == 1) launch_batch_class<bf16_t>(kern, K, dD + b, dUB + b, nd, (int)units[c], big, s);
        else launch_batch_class<f16_t>(kern, K, dD + b, dUB + b, nd, (int)units[c], big, s);
        HIPCHECK(hipGetLastError());
    }
    return TC_OK;
}

}  // namespace tc


// ---- tc_spmv_batch_static: the batched launch without any allocation ----
// The descriptor table lives in `scratch`, a fixed array of device memory
// that each GPU holds once for the whole process, and the descriptors are
// filled in a static host buffer; nothing is allocated on any call. In host
// code `scratch` names a host-side stand-in, not device memory: every host
// use goes through the address hipGetSymbolAddress gives for the current
// device (each GPU has its own copy at its own address), taken once per
// device at that device's first call.
constexpr size_t kStaticDescBytes = 1u << 20;  // descriptor bytes a call may use, per GPU

static __device__ __align__(256) char scratch[kStaticDescBytes];

namespace {

constexpr int kStaticMaxDevices = 64;

std::mutex g_static_mu;                               // guards the staging buffer and the addresses
void *g_static_dev[kStaticMaxDevices] = {};           // scratch's address on each device
alignas(64) char g_static_host[kStaticDescBytes];     // host staging of the descriptors

}  // namespace

extern "C" int tc_spmv_batch_static(const tc_request *reqs, int n, void *stream)
try {
    if (n < 0 || (n > 0 && !reqs)) return tc::fail(TC_ERR_INVALID, "tc_spmv_batch_static: bad argument");
    if (n == 0) return TC_OK;
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kStaticMaxDevices)
        return tc::fail(TC_ERR_UNSUPPORTED, "tc_spmv_batch_static: device %d beyond %d", dev, kStaticMaxDevices);
    std::lock_guard<std::mutex> g(g_static_mu);
    if (!g_static_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return tc::fail(TC_ERR_HIP, "tc_spmv_batch_static: hipGetSymbolAddress: %s", hipGetErrorString(e));
        g_static_dev[dev] = p;
    }
    // the staging buffer is pageable: the copy has read it when
    // hipMemcpyAsync returns, so the next call may refill it
    return tc::batch_enqueue(reqs, n, g_static_host, g_static_dev[dev], (hipStream_t)stream, kStaticDescBytes);
} catch (...) {
    return tc::exception_status(__func__);
}
