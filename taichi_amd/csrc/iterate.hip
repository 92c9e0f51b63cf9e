// Iterative driver (no reference counterpart; the step after SpMV): PageRank
// by power iteration over a resident binary matrix, with y -> x on the
// device and the whole loop optionally captured once as a hipGraph.
//
// A[i][j] = 1 means page j links to page i; inv_deg[j] = 1 / outdeg(j), or 0
// for a page without out-links (dangling). From r_0 = 1/n, each step is
//   z   = r (.) inv_deg                    (scaled ranks, the SpMV's x)
//   D   = sum of r over dangling pages     (their mass is spread uniformly)
//   y   = A z
//   r'  = (1 - alpha)/n + alpha (y + D/n)
// A step is the k = 1 rows tiles of the SpMV (spmv_tiles.h) whose end turns
// each row's sum y_i straight into r'_i and z'_i = r'_i inv_deg_i (y is
// never stored; z alternates between two buffers) and sums r' over the
// dangling rows, one partial per block; then a one-block launch folds the
// partials in index order into D' for the next step. (The last block of the
// step doing the fold instead, found by an atomic count, measured 8x slower:
// thousands of device-scope atomics on one word serialise.) Every sum has a
// fixed order, so the ranks are identical eagerly and replayed from the
// graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "spmv_host.h"

namespace {

constexpr int kWG = 256;
constexpr int kInitBlocks = 1024;  // blocks of the first step (grid-stride)

// this wavefront's partial d to part[4 * blockIdx.x + wave]: an xor butterfly
// (both lanes of a pair add the same two values), no block barrier
__device__ void wave_partial(float d, float *part)
{
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) d += __shfl_xor(d, m, 64);
    if ((threadIdx.x & 63) == 0) part[blockIdx.x * (kWG / 64) + (threadIdx.x >> 6)] = d;
}

// D = the sum of part[0..n) in a fixed order (one block of kFold lanes):
// lane t adds t, t + kFold, ... (loads issued kFoldU at a time), then the
// wavefront butterflies and the wavefront sums in order
constexpr int kFold = 1024, kFoldU = 32;
__global__ __launch_bounds__(kFold) void pr_fold(const float *__restrict__ part, int n, float *__restrict__ D)
{
    __shared__ float sw[kFold / 64];
    float a = 0.0f;
    for (int b = 0; b < n; b += kFold * kFoldU) {
        float v[kFoldU];
#pragma unroll
        for (int u = 0; u < kFoldU; u++) {
            const int i = b + u * kFold + threadIdx.x;
            v[u] = i < n ? part[i] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kFoldU; u++) a += v[u];
    }
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) a += __shfl_xor(a, m, 64);
    if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.0f;
        for (int w = 0; w < kFold / 64; w++) t += sw[w];
        *D = t;
    }
}

// step 0: r = 1/n, z = r inv_deg, D = the dangling mass of r (r written only
// when no iteration follows)
__global__ __launch_bounds__(kWG) void pr_init(int n, int write_r, const float *__restrict__ inv_deg,
                                               float *__restrict__ r, float *__restrict__ z, float *part)
{
    const float inv_n = 1.0f / (float)n;
    float d = 0.0f;
    for (int i = blockIdx.x * kWG + threadIdx.x; i < n; i += gridDim.x * kWG) {
        const float w = inv_deg[i];
        if (write_r) r[i] = inv_n;
        z[i] = inv_n * w;
        if (w == 0.0f) d += inv_n;
    }
    wave_partial(d, part);
}

// what a row's sum becomes in a step: r'_i, z'_i and the dangling sum
struct PrEnd {
    static constexpr bool kStore = false;
    const float *inv_deg;
    float *z, *r;
    float base, alpha, spread;
    int write_r;
    float &d;
    __device__ void operator()(int i, tcd::acc_t sum) const
    {
        const float ri = base + alpha * ((float)sum + spread);
        const float w = inv_deg[i];
        if (write_r) r[i] = ri;
        z[i] = ri * w;
        if (w == 0.0f) d += ri;
    }
};

// one step: z' and r' from A zin with the previous step's D; one partial of
// the dangling mass per block
template <int RPT, int OW, bool FAST, bool LONG>
__global__ __launch_bounds__(kWG) void pr_step(tcd::MatView A, const float *__restrict__ zin, float *__restrict__ zout,
                                               float *__restrict__ r, const float *__restrict__ inv_deg, float alpha,
                                               int write_r, const float *Dprev, float *part, int ntiles, int capv)
{
    extern __shared__ float smem[];
    const float inv_n = 1.0f / (float)A.M;
    float d = 0.0f;
    const PrEnd end{inv_deg, zout, r, (1.0f - alpha) * inv_n, alpha, *Dprev * inv_n, write_r, d};
    tcd::rows_tile<float, RPT, OW, FAST, LONG, true, PrEnd>(A, zin, 1, nullptr, 1,
                                                            tcd::xcd_slot(blockIdx.x, ntiles), smem, capv, end);
    wave_partial(d, part);
}

template <int RPT, int OW, bool FAST, bool LONG>
void launch_step_k(const tcd::MatView &A, const float *zin, float *zout, float *r, const float *inv_deg,
                   float alpha, int write_r, const float *Dp, float *part, int ntiles, int capv, size_t lds,
                   hipStream_t s)
{
    pr_step<RPT, OW, FAST, LONG><<<ntiles, kWG, lds, s>>>(A, zin, zout, r, inv_deg, alpha, write_r, Dp, part, ntiles,
                                                          capv);
}

// one step and the fold of its partials into *Dn
template <int RPT>
int launch_step_t(const tcd::MatView &A, const float *zin, float *zout, float *r, const float *inv_deg,
                  float alpha, int write_r, const float *Dp, float *Dn, float *part, int capv, hipStream_t s)
{
    constexpr int TRW = tcd::kWG * RPT;
    const int ntiles = (A.M + TRW - 1) / TRW;
    const bool fast = A.S == 1 && A.sh0.oe - A.sh0.ob == 1 && A.sh0.pm < 0;
    const bool lng = A.L > 0;
    const size_t lds = ((size_t)tcd::lsum_at(capv) + (lng ? 2 * TRW : 0)) * 4;
    if (lds + kWG * 4 > tcd::max_lds_bytes())
        return tc::fail(TC_ERR_UNSUPPORTED, "pagerank: tiles of %d rows with %d LDS values need %zu B", TRW, capv, lds);
#define TC_STEP(OW, F, L) \
    launch_step_k<RPT, OW, F, L>(A, zin, zout, r, inv_deg, alpha, write_r, Dp, part, ntiles, capv, lds, s)
    if (A.ow == 1) {
        if (fast) lng ? TC_STEP(1, true, true) : TC_STEP(1, true, false);
        else lng ? TC_STEP(1, false, true) : TC_STEP(1, false, false);
    } else {
        if (fast) lng ? TC_STEP(2, true, true) : TC_STEP(2, true, false);
        else lng ? TC_STEP(2, false, true) : TC_STEP(2, false, false);
    }
#undef TC_STEP
    pr_fold<<<1, kFold, 0, s>>>(part, ntiles * (kWG / 64), Dn);
    return TC_OK;
}

// work layout: z[2][n] | part[nparts] | D[2], each 256-byte aligned
size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// one partial per wavefront of the first step's or a step's grid
size_t nparts_for(int n) { return (kWG / 64) * std::max((size_t)kInitBlocks, ((size_t)n + 255) / 256); }

struct Work {
    float *z[2], *part, *D;
};

Work carve(void *work, int n)
{
    char *p = (char *)work;
    Work w;
    w.z[0] = (float *)p;
    p += align256((size_t)n * 4);
    w.z[1] = (float *)p;
    p += align256((size_t)n * 4);
    w.part = (float *)p;
    p += align256(nparts_for(n) * 4);
    w.D = (float *)p;
    return w;
}

// the first step (r = 1/n) then `iters` iterations, all on s. The rows tile
// size follows the matrix's k = 1 configuration (256 rows, else 512) and its
// LDS stage.
int enqueue_pagerank(tc_matrix *A, int n, const float *inv_deg, float alpha, int iters, float *r, const Work &w,
                     hipStream_t s)
{
    const tc_config c = tcd::get_cfg(A, 1, TC_F32);
    const int capv = std::max(c.lds_values, 64);
    const bool one = c.rows_per_tile == 256;
    const int G = std::max(1, std::min(kInitBlocks, (n + kWG - 1) / kWG));
    pr_init<<<G, kWG, 0, s>>>(n, iters == 0, inv_deg, r, w.z[0], w.part);
    pr_fold<<<1, kFold, 0, s>>>(w.part, G * (kWG / 64), &w.D[0]);
    for (int it = 1; it <= iters; it++) {
        const float *zin = w.z[(it - 1) & 1];
        float *zout = w.z[it & 1];
        const float *Dp = &w.D[(it - 1) & 1];
        float *Dn = &w.D[it & 1];
        const int rc = one ? launch_step_t<1>(A->v, zin, zout, r, inv_deg, alpha, it == iters, Dp, Dn, w.part, capv, s)
                           : launch_step_t<2>(A->v, zin, zout, r, inv_deg, alpha, it == iters, Dp, Dn, w.part, capv, s);
        if (rc != TC_OK) return rc;
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TC_OK : tc::fail(TC_ERR_HIP, "pagerank launch: %s", hipGetErrorString(e));
}

}  // namespace

extern "C" size_t tc_pagerank_work_bytes(int n)
{
    if (n < 0) return 0;
    return 2 * align256((size_t)n * 4) + align256(nparts_for(n) * 4) + align256(2 * 4);
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
