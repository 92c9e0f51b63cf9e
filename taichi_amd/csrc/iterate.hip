// Iterative driver (no reference counterpart; the step after SpMV): PageRank
// by power iteration over a resident binary matrix, with y -> x on the
// device and the whole loop optionally captured once as a hipGraph.
//
// A[i][j] = 1 means page j links to page i; inv_deg[j] = 1 / outdeg(j), or 0
// for a page without out-links (dangling). From r_0 = 1/n, each step is
//   z   = r (.) inv_deg                    (scaled ranks, the SpMV's x)
//   D   = sum of r over dangling pages     (their mass is spread uniformly)
//   y   = A z                              (tc_spmv)
//   r'  = (1 - alpha)/n + alpha (y + D/n)
// One fused kernel per step reads y, writes the next z (and r' on the last
// step) and one partial sum of the dangling mass per block; the next step's
// blocks each fold the previous partials in the same fixed order (two
// partial buffers alternate), so no separate reduction launch is needed. All
// sums have fixed orders, so the ranks are identical eagerly and replayed
// from the graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "spmv_host.h"

namespace {

constexpr int kWG = 256;
constexpr int kParts = 1024;  // blocks of the fused update (grid-stride), one partial each

// fixed-order sum of a workgroup's 256 values (tree over LDS)
__device__ float block_sum(float v, float *sh)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = kWG / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] += sh[t + s];
        __syncthreads();
    }
    const float r = sh[0];
    __syncthreads();
    return r;
}

// Step c of the iteration (c = 0: r = 1/n, no SpMV before it). Reads the
// partials of step c-1 from prev, writes its own to part; writes r only when
// write_r is set (the last step), z always. VEC: all four vectors are 16-byte
// aligned and each block takes chunks of kChunk elements, every lane issuing
// kU float4 loads of y and of inv_deg before it uses any (enough bytes in
// flight to stream HBM); the n % 4 tail is done by block 0. Scalar otherwise.
constexpr int kU = 4;
constexpr int kChunk = kWG * 4 * kU;

__device__ inline float rank_of(int first, float inv_n, float base, float alpha, float y, float spread)
{
    return first ? inv_n : base + alpha * (y + spread);
}

// Sum of the previous step's partials (at most kParts), the same value in
// every lane of every wave: lane l loads partials l, l + 64, ... all at once
// and adds them in that order, then an xor butterfly (both lanes of a pair add
// the same two values, so they stay identical).
__device__ inline float fold_partials(const float *__restrict__ prev, int nprev)
{
    const int lane = threadIdx.x & 63;
    float v[kParts / 64];
#pragma unroll
    for (int j = 0; j < kParts / 64; j++) {
        const int i = j * 64 + lane;
        v[j] = i < nprev ? prev[i] : 0.0f;
    }
    float d = 0.0f;
#pragma unroll
    for (int j = 0; j < kParts / 64; j++) d += v[j];
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) d += __shfl_xor(d, m, 64);
    return d;
}

__device__ inline void load_chunk(long long c, long long n4, int first, const float4 *__restrict__ y4,
                                  const float4 *__restrict__ w4, float4 *yv, float4 *wv)
{
#pragma unroll
    for (int u = 0; u < kU; u++) {
        const long long i = c + u * kWG + threadIdx.x;
        if (i < n4) {
            yv[u] = first ? make_float4(0.f, 0.f, 0.f, 0.f) : y4[i];
            wv[u] = w4[i];
        }
    }
}

template <bool VEC>
__global__ __launch_bounds__(kWG) void pr_update(int n, float alpha, int first, int write_r,
                                                 const float *__restrict__ y, const float *__restrict__ inv_deg,
                                                 const float *__restrict__ prev, int nprev, float *__restrict__ r,
                                                 float *__restrict__ z, float *__restrict__ part)
{
    __shared__ float sh[kWG];
    const float inv_n = 1.0f / (float)n;
    const float base = (1.0f - alpha) * inv_n;
    float d = 0.0f;
    if constexpr (VEC) {
        const long long n4 = n >> 2;
        const float4 *y4 = (const float4 *)y;
        const float4 *w4 = (const float4 *)inv_deg;
        const long long step = (long long)gridDim.x * (kChunk / 4);
        long long c = (long long)blockIdx.x * (kChunk / 4);
        float4 yv[kU], wv[kU];
        load_chunk(c, n4, first, y4, w4, yv, wv);  // in flight while the partials are folded
        const float spread = first ? 0.0f : fold_partials(prev, nprev) * inv_n;
        for (; c < n4; c += step) {
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const long long i = c + u * kWG + threadIdx.x;
                if (i < n4) {
                    float4 rv, zv;
                    rv.x = rank_of(first, inv_n, base, alpha, yv[u].x, spread);
                    rv.y = rank_of(first, inv_n, base, alpha, yv[u].y, spread);
                    rv.z = rank_of(first, inv_n, base, alpha, yv[u].z, spread);
                    rv.w = rank_of(first, inv_n, base, alpha, yv[u].w, spread);
                    zv.x = rv.x * wv[u].x;
                    zv.y = rv.y * wv[u].y;
                    zv.z = rv.z * wv[u].z;
                    zv.w = rv.w * wv[u].w;
                    if (write_r) ((float4 *)r)[i] = rv;
                    ((float4 *)z)[i] = zv;
                    if (wv[u].x == 0.0f) d += rv.x;
                    if (wv[u].y == 0.0f) d += rv.y;
                    if (wv[u].z == 0.0f) d += rv.z;
                    if (wv[u].w == 0.0f) d += rv.w;
                }
            }
            if (c + step < n4) load_chunk(c + step, n4, first, y4, w4, yv, wv);
        }
        if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {
            const int i = (int)(n4 << 2) + threadIdx.x;
            const float ri = rank_of(first, inv_n, base, alpha, first ? 0.0f : y[i], spread);
            const float w = inv_deg[i];
            if (write_r) r[i] = ri;
            z[i] = ri * w;
            if (w == 0.0f) d += ri;
        }
    } else {
        const float spread = first ? 0.0f : fold_partials(prev, nprev) * inv_n;
        for (int i = blockIdx.x * kWG + threadIdx.x; i < n; i += gridDim.x * kWG) {
            const float ri = rank_of(first, inv_n, base, alpha, first ? 0.0f : y[i], spread);
            const float w = inv_deg[i];
            if (write_r) r[i] = ri;
            z[i] = ri * w;
            if (w == 0.0f) d += ri;
        }
    }
    d = block_sum(d, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = d;
}

// work layout: y[n] | z[n] | part[2][kParts], each 256-byte aligned
size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

struct Work {
    float *y, *z, *part[2];
};

Work carve(void *work, int n)
{
    char *p = (char *)work;
    Work w;
    w.y = (float *)p;
    p += align256((size_t)n * 4);
    w.z = (float *)p;
    p += align256((size_t)n * 4);
    w.part[0] = (float *)p;
    p += align256((size_t)kParts * 4);
    w.part[1] = (float *)p;
    return w;
}

int grid_for(int n, bool vec)
{
    const int per = vec ? kChunk : kWG;
    return std::max(1, std::min(kParts, (n + per - 1) / per));
}

bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

// the first step's update (r = 1/n) then `iters` iterations, all on s
int enqueue_pagerank(tc_matrix *A, int n, const float *inv_deg, float alpha, int iters, float *r, const Work &w,
                     hipStream_t s)
{
    const bool vec = aligned16(inv_deg) && aligned16(r) && aligned16(w.y) && aligned16(w.z);
    const int G = grid_for(n, vec);
    auto update = vec ? pr_update<true> : pr_update<false>;
    update<<<G, kWG, 0, s>>>(n, alpha, 1, iters == 0, w.y, inv_deg, w.part[1], G, r, w.z, w.part[0]);
    for (int it = 1; it <= iters; it++) {
        int rc = tc_spmv(A, w.z, 1, TC_F32, w.y, 1, 1, s);
        if (rc != TC_OK) return rc;
        update<<<G, kWG, 0, s>>>(n, alpha, 0, it == iters, w.y, inv_deg, w.part[(it - 1) & 1], G, r, w.z,
                                 w.part[it & 1]);
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TC_OK : tc::fail(TC_ERR_HIP, "pagerank launch: %s", hipGetErrorString(e));
}

}  // namespace

extern "C" size_t tc_pagerank_work_bytes(int n)
{
    if (n < 0) return 0;
    return 2 * align256((size_t)n * 4) + 2 * align256((size_t)kParts * 4);
}

// checks shared by the eager and the captured form
static int pagerank_args(tc_matrix *A, const float *inv_deg, float alpha, int iters, float *r, void *work, int *n)
{
    tc_matrix_stats st;
    int rc = tc_matrix_get_stats(A, &st);
    if (rc != TC_OK) return rc;
    if (st.M != st.N) return tc::fail(TC_ERR_INVALID, "pagerank: matrix is %d x %d, not square", st.M, st.N);
    if (!inv_deg || !r || !work || iters < 0 || !(alpha >= 0.0f && alpha <= 1.0f))
        return tc::fail(TC_ERR_INVALID, "pagerank: bad argument");
    *n = st.M;
    return TC_OK;
}

extern "C" int tc_pagerank(tc_matrix *A, const float *inv_deg, float alpha, int iters, float *r, void *work,
                           void *stream)
try {
    int n = 0;
    int rc = pagerank_args(A, inv_deg, alpha, iters, r, work, &n);
    if (rc != TC_OK || n == 0) return rc;
    return enqueue_pagerank(A, n, inv_deg, alpha, iters, r, carve(work, n), (hipStream_t)stream);
} catch (...) {
    return tc::exception_status(__func__);
}

extern "C" int tc_pagerank_capture(tc_matrix *A, const float *inv_deg, float alpha, int iters, float *r,
                                   void *work, void *stream, tc_graph **out)
try {
    int n = 0;
    int rc = pagerank_args(A, inv_deg, alpha, iters, r, work, &n);
    if (rc != TC_OK) return rc;
    if (!out || n == 0) return tc::fail(TC_ERR_INVALID, "tc_pagerank_capture: bad argument");
    hipStream_t cs = (hipStream_t)stream;
    bool own = false;
    if (!cs) {  // the legacy stream cannot be captured
        if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess)
            return tc::fail(TC_ERR_HIP, "tc_pagerank_capture: stream create");
        own = true;
    }
    tc_graph *g = new tc_graph();
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        rc = enqueue_pagerank(A, n, inv_deg, alpha, iters, r, carve(work, n), cs);
        const hipError_t e2 = hipStreamEndCapture(cs, &g->graph);
        if (e == hipSuccess) e = e2;
    }
    if (e == hipSuccess && rc == TC_OK) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (own) (void)hipStreamDestroy(cs);
    if (e != hipSuccess || rc != TC_OK) {
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g;
        return rc != TC_OK ? rc : tc::fail(TC_ERR_HIP, "tc_pagerank_capture: %s", hipGetErrorString(e));
    }
    *out = g;
    return TC_OK;
} catch (...) {
    return tc::exception_status(__func__);
}
