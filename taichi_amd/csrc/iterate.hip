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
// One fused kernel writes r', the next z and per-block partial sums of the
// dangling mass; a one-block kernel folds the partials in a fixed order. All
// sums have fixed orders, so the ranks are identical eagerly and replayed
// from the graph.
#include <hip/hip_runtime.h>

#include <algorithm>

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

// r' from y (skipped on the first call: r = 1/n), then z and the partials
__global__ __launch_bounds__(kWG) void pr_update(int n, float alpha, int first, const float *__restrict__ y,
                                                 const float *__restrict__ inv_deg, const float *__restrict__ dang,
                                                 float *__restrict__ r, float *__restrict__ z,
                                                 float *__restrict__ part)
{
    __shared__ float sh[kWG];
    const float inv_n = 1.0f / (float)n;
    const float base = (1.0f - alpha) * inv_n;
    const float spread = first ? 0.0f : dang[0] * inv_n;
    float d = 0.0f;
    for (int i = blockIdx.x * kWG + threadIdx.x; i < n; i += gridDim.x * kWG) {
        const float ri = first ? inv_n : base + alpha * (y[i] + spread);
        const float w = inv_deg[i];
        r[i] = ri;
        z[i] = ri * w;
        if (w == 0.0f) d += ri;
    }
    d = block_sum(d, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = d;
}

__global__ __launch_bounds__(kWG) void pr_fold(const float *__restrict__ part, int nparts, float *__restrict__ dang)
{
    __shared__ float sh[kWG];
    float d = 0.0f;
    for (int i = threadIdx.x; i < nparts; i += kWG) d += part[i];
    d = block_sum(d, sh);
    if (threadIdx.x == 0) dang[0] = d;
}

// work layout: y[n] | z[n] | part[kParts] | dang[1], each 256-byte aligned
size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

struct Work {
    float *y, *z, *part, *dang;
};

Work carve(void *work, int n)
{
    char *p = (char *)work;
    Work w;
    w.y = (float *)p;
    p += align256((size_t)n * 4);
    w.z = (float *)p;
    p += align256((size_t)n * 4);
    w.part = (float *)p;
    p += align256((size_t)kParts * 4);
    w.dang = (float *)p;
    return w;
}

int grid_for(int n) { return std::max(1, std::min(kParts, (n + kWG - 1) / kWG)); }

// the first step's update (r = 1/n) then `iters` iterations, all on s
int enqueue_pagerank(tc_matrix *A, int n, const float *inv_deg, float alpha, int iters, float *r, const Work &w,
                     hipStream_t s)
{
    const int G = grid_for(n);
    pr_update<<<G, kWG, 0, s>>>(n, alpha, 1, w.y, inv_deg, w.dang, r, w.z, w.part);
    pr_fold<<<1, kWG, 0, s>>>(w.part, G, w.dang);
    for (int it = 0; it < iters; it++) {
        int rc = tc_spmv(A, w.z, 1, TC_F32, w.y, 1, 1, s);
        if (rc != TC_OK) return rc;
        pr_update<<<G, kWG, 0, s>>>(n, alpha, 0, w.y, inv_deg, w.dang, r, w.z, w.part);
        pr_fold<<<1, kWG, 0, s>>>(w.part, G, w.dang);
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TC_OK : tc::fail(TC_ERR_HIP, "pagerank launch: %s", hipGetErrorString(e));
}

}  // namespace

extern "C" size_t tc_pagerank_work_bytes(int n)
{
    if (n < 0) return 0;
    return 2 * align256((size_t)n * 4) + align256((size_t)kParts * 4) + 256;
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
{
    int n = 0;
    int rc = pagerank_args(A, inv_deg, alpha, iters, r, work, &n);
    if (rc != TC_OK || n == 0) return rc;
    return enqueue_pagerank(A, n, inv_deg, alpha, iters, r, carve(work, n), (hipStream_t)stream);
}

extern "C" int tc_pagerank_capture(tc_matrix *A, const float *inv_deg, float alpha, int iters, float *r,
                                   void *work, void *stream, tc_graph **out)
{
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
}
