// Hybrid diagonal + CSR SpMV for gfx950 (kernels (b) and (c), fused).
//
// Replaces taichi_SpMV's upload and CSR_coop_spmv_kernel{,_hybrid1,2,3}
// (TaiChi.cu:965-1456). y = A x with A in the TaiChi_new format: dense
// diagonal segments of square shapes plus a value-free, sign-encoded CSR
// remainder (csrCol = col<<1 | sign; sign 0 subtracts x, compensating a zero
// inside a dense diagonal).
//
// Device layout (built by tc_matrix_create from a host TaiChi_new):
//   shapes[S]   {xStart, xStop, yStart, off_begin, off_end, off_min, off_max}
//   offs[D]     dense diagonal offsets
//   gbase[G+1]  start of each 64-row group's short-row entries in cols
//   roff[M]     u8 or u16: end of row i's short entries, relative to gbase[i/64]
//   cols[R]     encoded entries: short rows in row order, then long rows
//   glong[G+1], lrow[L], lptr[L+1]   rows longer than the long-row threshold
// Compared with the reference's int rowptr (4 bytes per row) the per-row
// cost is 1 (or 2) bytes plus 4 bytes per 64 rows.
//
// One workgroup (4 wavefronts) computes a tile of TRW = 256*P/K rows for K
// right-hand sides:
//   1. row extents of the tile into LDS;
//   2. short-row entries of the tile streamed with coalesced loads, x gathered
//      and sign applied, values staged in LDS (chunks of `capv` floats); each
//      (row, column) pair is then summed serially by one lane in entry order;
//   3. each long row is summed by one wavefront (64-lane strided partial sums,
//      DPP/ds_swizzle butterfly, fixed order);
//   4. every diagonal shape overlapping the tile adds x[r - xS + yS + o] for its
//      offsets o inside the square's column window; the x window of the tile
//      is staged once in LDS when a shape has two or more offsets, so no
//      column index is read for the dense part;
//   5. y stored once, coalesced.
// The per-(row, column) summation order depends only on the format (short
// rows serial, long rows a fixed butterfly, then shapes and offsets in stored
// order), never on the tile size, the grid or the batch a request runs in, so
// a request gives bit-identical y alone, in a batch, or with another config.
//
// No MFMA: one add per stored index, bound by HBM bandwidth.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "batch.h"
#include "common.h"

namespace {

constexpr int kWG = 256;
constexpr int kWaves = kWG / 64;
constexpr int kStageU = 8;  // loads in flight per lane when staging streams
// Diagonal terms of a shape are taken in blocks of kDiagBlock offsets (by
// offset index from the shape's first): a block is summed with plain FP32
// adds in offset order and its sum enters the compensated accumulator once.
// A block of 4 adds at most 3 ulps-relative error to its terms; the
// compensated part keeps the cancellation against the remainder's zero
// entries exact.
constexpr int kDiagBlock = 4;

struct alignas(16) ShapeDev {
    int xs, xe, ys, ob, oe, omin, omax, pad;
};

// Device pointers and sizes of one resident matrix; passed by value.
struct MatView {
    int M, N, S, L;
    int ow;        // bytes per local row offset: 1 or 2
    int ngroups;
    ShapeDev sh0;  // copy of shapes[0]: arrives with the kernel arguments
    const ShapeDev *shapes;
    const int *offs;
    const int *gbase;
    const void *roff;
    const int *cols;
    const int *glong;
    const int *lrow;
    const int *lptr;
};

struct bf16_t {
    uint16_t bits;
};
typedef _Float16 f16_t;

template <typename XT>
__device__ __forceinline__ float load_x(const XT *__restrict__ X, size_t i);
template <>
__device__ __forceinline__ float load_x<float>(const float *__restrict__ X, size_t i)
{
    return X[i];
}
template <>
__device__ __forceinline__ float load_x<bf16_t>(const bf16_t *__restrict__ X, size_t i)
{
    return __uint_as_float((uint32_t)X[i].bits << 16);
}
template <>
__device__ __forceinline__ float load_x<f16_t>(const f16_t *__restrict__ X, size_t i)
{
    return (float)X[i];
}

// Read-only format metadata (group bases, shapes, offsets, long-row tables)
// is read through the constant address space: uniform indices then become
// scalar loads, which do not make the vector loads behind them wait.
template <typename T>
__device__ __forceinline__ T cload(const T *p, int i)
{
    return ((const __attribute__((address_space(4))) T *)p)[i];
}

__device__ __forceinline__ ShapeDev load_shape(const MatView &A, int s)
{
    const int *p = (const int *)(A.shapes + s);
    ShapeDev h;
    h.xs = cload(p, 0);
    h.xe = cload(p, 1);
    h.ys = cload(p, 2);
    h.ob = cload(p, 3);
    h.oe = cload(p, 4);
    h.omin = cload(p, 5);
    h.omax = cload(p, 6);
    h.pad = 0;
    return h;
}

// start of row i's short entries (i in [0, M]); row i ends at row_start(i+1)
__device__ __forceinline__ int row_start(const MatView &A, int i)
{
    int b = cload(A.gbase, i >> 6);
    if (i & 63) b += A.ow == 1 ? (int)((const uint8_t *)A.roff)[i - 1] : (int)((const uint16_t *)A.roff)[i - 1];
    return b;
}

// Sum over the 64 lanes of a wavefront in a fixed order: quad swaps and
// row mirrors by DPP, the 16-lane swap by ds_swizzle, halves by readlane.
__device__ __forceinline__ float wave_sum(float v)
{
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));   // quad_perm 1,0,3,2
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));   // quad_perm 2,3,0,1
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xF, 0xF, false));  // row_half_mirror
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x140, 0xF, 0xF, false));  // row_mirror
    v += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401F));                      // lane ^ 16
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0)) +
           __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
}

// Compensated accumulation: s + c carries the running sum with the exact
// rounding error of every add (Knuth's TwoSum, branch-free) kept in c, the
// same error term Neumaier's magnitude-ordered form computes. Dense diagonals
// add x over whole segments and the remainder subtracts it again at their
// zero cells, so plain FP32 sums lose accuracy to cancellation; the
// compensated sum keeps each y within a few ulps of the exact sum.
// Integer-valued inputs stay exact (c stays 0). Needs -ffp-contract=off.
__device__ __forceinline__ void cadd(float &s, float &c, float x)
{
    const float t = s + x;
    const float z = t - s;
    c += (s - (t - z)) + (x - z);
    s = t;
}

// Bijection of block ids so that each XCD (blocks are dispatched round-robin
// over the 8 XCDs) gets a contiguous range of slots, hence contiguous rows
// and x windows in its own L2.
__device__ __forceinline__ int xcd_slot(int b, int G)
{
    const int q = G >> 3, r = G & 7, x = b & 7, l = b >> 3;
    return x * q + min(x, r) + l;
}

// Contiguous range [*b, *e) of `total` work units for this block. With one
// block per unit it is just the XCD-remapped block id; otherwise a 32-bit
// split (callers keep total * G below 2^31).
__device__ __forceinline__ void block_units(int total, int *b, int *e)
{
    const int G = gridDim.x;
    const int v = xcd_slot(blockIdx.x, G);
    if (G == total) {
        *b = v;
        *e = v + 1;
    } else {
        *b = (int)((unsigned)v * (unsigned)total / (unsigned)G);
        *e = (int)((unsigned)(v + 1) * (unsigned)total / (unsigned)G);
    }
}

// Stage the signed x values of ne contiguous entries (K columns each) into
// s_val[j*K + c]. The loads of each batch (U items per lane) are all issued
// before the first gather waits on them; lanes past the end re-read a valid
// item and do not store.
template <int K, typename XT, int U = kStageU>
__device__ __forceinline__ void stage_entries(const int *__restrict__ cp, int ne, const XT *__restrict__ X, int ldx,
                                              float *s_val)
{
    const int tid = threadIdx.x;
    const int n = ne * K;
    for (int b = 0; b < n; b += kWG * U) {
        // slots past the end of the stream are skipped as a whole (the
        // condition is uniform across the workgroup)
        int v[U], cc[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            v[u] = 1;
            cc[u] = 0;
            if (b + u * kWG < n) {
                const int idx = min(b + u * kWG + tid, n - 1);
                const int j = idx / K;
                cc[u] = idx - j * K;
                v[u] = cp[j];
            }
        }
        float xv[U];
#pragma unroll
        for (int u = 0; u < U; u++) xv[u] = b + u * kWG < n ? load_x<XT>(X, (size_t)(v[u] >> 1) * ldx + cc[u]) : 0.0f;
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int idx = b + u * kWG + tid;
            if (idx < n) s_val[idx] = (v[u] & 1) ? xv[u] : -xv[u];
        }
    }
}

// Diagonal terms of one shape for the tile's pairs, in canonical order
// (offsets as stored), added to (acc, cmp). The x window of a shape with two
// or more offsets is staged in LDS once per tile (s_val, left free after).
template <int K, typename XT, int P>
__device__ __forceinline__ void diag_one_shape(const MatView &A, const XT *__restrict__ X, int ldx, int u0, int s,
                                               const ShapeDev &sh, int ra, int rb, const int (&prow)[P],
                                               const int (&pcol)[P], const bool (&pv)[P], float (&acc)[P],
                                               float (&cmp)[P], float *s_val, int capv)
{
    const int tid = threadIdx.x;
    const int shift = sh.ys - sh.xs;
    const int clo = sh.ys, chi = sh.ys + (sh.xe - sh.xs);
    const int w0 = max(clo, ra + shift + sh.omin);
    const int w1 = min(chi, rb - 1 + shift + sh.omax + 1);
    const bool stage = (sh.oe - sh.ob) >= 2 && w1 > w0 && (w1 - w0) * K <= capv;
    if (stage) {
        const int nw = (w1 - w0) * K;
        for (int b = 0; b < nw; b += kWG * kStageU) {
            float t[kStageU];
#pragma unroll
            for (int u = 0; u < kStageU; u++) {
                const int idx = min(b + u * kWG + tid, nw - 1);
                const int cc = idx / K, c = idx - cc * K;
                t[u] = load_x<XT>(X, (size_t)(w0 + cc) * ldx + c);
            }
#pragma unroll
            for (int u = 0; u < kStageU; u++) {
                const int idx = b + u * kWG + tid;
                if (idx < nw) s_val[idx] = t[u];
            }
        }
        __syncthreads();
    }
    // Offsets outer, pairs inner: each offset is one scalar load and the P
    // reads behind it are independent. Each pair sums the terms of a block
    // of kDiagBlock offsets in offset order with plain adds and adds the
    // block sum to (acc, cmp) compensated (the canonical order).
    bool act[P];
    int base[P];
    bool inner = true;  // all of this lane's pairs active, all columns inside the square
#pragma unroll
    for (int p = 0; p < P; p++) {
        const int r = u0 + prow[p];
        act[p] = pv[p] && r >= ra && r < rb;
        base[p] = r + shift;
        inner = inner && act[p] && base[p] + sh.omin >= clo && base[p] + sh.omax < chi;
    }
    if (stage) {
        int li[P];
#pragma unroll
        for (int p = 0; p < P; p++) li[p] = (base[p] - w0) * K + pcol[p];
        if (__builtin_amdgcn_ballot_w64(!inner) == 0) {
            for (int n0 = sh.ob; n0 < sh.oe; n0 += kDiagBlock) {
                const int n1 = min(n0 + kDiagBlock, sh.oe);
                float bs[P];
#pragma unroll
                for (int p = 0; p < P; p++) bs[p] = 0.0f;
                for (int n = n0; n < n1; n++) {
                    const int o = cload(A.offs, n) * K;
#pragma unroll
                    for (int p = 0; p < P; p++) bs[p] += s_val[li[p] + o];
                }
#pragma unroll
                for (int p = 0; p < P; p++) cadd(acc[p], cmp[p], bs[p]);
            }
        } else {
            for (int n0 = sh.ob; n0 < sh.oe; n0 += kDiagBlock) {
                const int n1 = min(n0 + kDiagBlock, sh.oe);
                float bs[P];
#pragma unroll
                for (int p = 0; p < P; p++) bs[p] = 0.0f;
                for (int n = n0; n < n1; n++) {
                    const int o = cload(A.offs, n);
#pragma unroll
                    for (int p = 0; p < P; p++) {
                        const int c = base[p] + o;
                        if (act[p] && c >= clo && c < chi) bs[p] += s_val[li[p] + o * K];
                    }
                }
#pragma unroll
                for (int p = 0; p < P; p++) cadd(acc[p], cmp[p], bs[p]);
            }
        }
        __syncthreads();
    } else {
        for (int n0 = sh.ob; n0 < sh.oe; n0 += kDiagBlock) {
            const int n1 = min(n0 + kDiagBlock, sh.oe);
            float bs[P];
#pragma unroll
            for (int p = 0; p < P; p++) bs[p] = 0.0f;
            for (int n = n0; n < n1; n++) {
                const int o = cload(A.offs, n);
#pragma unroll
                for (int p = 0; p < P; p++) {
                    const int c = base[p] + o;
                    if (act[p] && c >= clo && c < chi) bs[p] += load_x<XT>(X, (size_t)c * ldx + pcol[p]);
                }
            }
#pragma unroll
            for (int p = 0; p < P; p++) cadd(acc[p], cmp[p], bs[p]);
        }
    }
}

template <int K, typename XT, int P>
__device__ __forceinline__ void diag_phase(const MatView &A, const XT *__restrict__ X, int ldx, int u0, int rows,
                                           const int (&prow)[P], const int (&pcol)[P], const bool (&pv)[P],
                                           float (&acc)[P], float (&cmp)[P], float *s_val, int capv)
{
    for (int s = 0; s < A.S; s++) {
        const ShapeDev sh = s == 0 ? A.sh0 : load_shape(A, s);
        const int ra = max(u0, sh.xs), rb = min(u0 + rows, sh.xe);
        if (ra >= rb) continue;
        diag_one_shape<K, XT, P>(A, X, ldx, u0, s, sh, ra, rb, prow, pcol, pv, acc, cmp, s_val, capv);
    }
}

// Long rows of the tile: one wavefront each, partial sums strided over the
// 64 lanes and reduced by wave_sum; results land in s_lsum[local_row*K + c]
// and are added to acc. Uniform control flow; leaves s_lsum free.
template <int K, typename XT, int P>
__device__ __forceinline__ void long_phase(const MatView &A, const XT *__restrict__ X, int ldx, int u0, int rows,
                                           const int (&lq)[P], const bool (&pv)[P], float (&acc)[P],
                                           float (&cmp)[P], float *s_lsum, int nlsum)
{
    if (A.L == 0) return;
    const int lb = cload(A.glong, u0 >> 6), le = cload(A.glong, ((u0 + rows - 1) >> 6) + 1);
    if (lb >= le) return;
    const int tid = threadIdx.x;
    for (int q = tid; q < nlsum; q += kWG) s_lsum[q] = 0.0f;
    __syncthreads();
    const int wave = tid >> 6, lane = tid & 63;
    for (int l = lb + wave; l < le; l += kWaves) {
        const int row = cload(A.lrow, l);
        if (row < u0 || row >= u0 + rows) continue;
        const int j0 = cload(A.lptr, l), j1 = cload(A.lptr, l + 1);
        for (int c = 0; c < K; c++) {
            float part = 0.0f, pe = 0.0f;
            for (int j = j0 + lane; j < j1; j += 64) {
                const int v = A.cols[j];
                const float xv = load_x<XT>(X, (size_t)(v >> 1) * ldx + c);
                cadd(part, pe, (v & 1) ? xv : -xv);
            }
            part = wave_sum(part + pe);
            if (lane == 0) s_lsum[(row - u0) * K + c] = part;
        }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < P; p++)
        if (pv[p]) cadd(acc[p], cmp[p], s_lsum[lq[p]]);
    __syncthreads();
}

// Canonical per-(row, column) order, shared by every kernel below, with
// compensated accumulation (cadd) into (acc, cmp) and y = acc + cmp:
//   diagonal terms (shape order, offset order);
//   short-row entries in stored order, one lane, serially;
//   the long-row wavefront sum (long rows have no short entries).
template <int K, typename XT, int P>
__device__ __forceinline__ void run_tile(const MatView &A, const XT *__restrict__ X, int ldx, float *__restrict__ Y,
                                         int ldy, int u0, float *s_val, int capv, int *s_start, float *s_lsum)
{
    constexpr int TRW = (kWG * P) / K;
    constexpr int NP = TRW * K;
    const int tid = threadIdx.x;
    const int rows = min(TRW, A.M - u0);

    // 1. row extents
    for (int r = tid; r <= TRW; r += kWG) s_start[r] = row_start(A, min(u0 + r, A.M));

    float acc[P], cmp[P];
    int prow[P], pcol[P], lq[P];
    bool pv[P];
#pragma unroll
    for (int p = 0; p < P; p++) {
        const int q = tid + p * kWG;
        prow[p] = q / K;
        pcol[p] = q - prow[p] * K;
        pv[p] = q < NP && prow[p] < rows;
        lq[p] = q < NP ? q : 0;
        acc[p] = 0.0f;
        cmp[p] = 0.0f;
    }

    // 2. diagonal segments
    diag_phase<K, XT, P>(A, X, ldx, u0, rows, prow, pcol, pv, acc, cmp, s_val, capv);
    __syncthreads();
    const int ebase = s_start[0];
    const int E = s_start[TRW] - ebase;

    // 3. short rows: stream entries, stage signed x values, serial row sums
    const int cape = capv / K;
    for (int e0 = 0; e0 < E; e0 += cape) {
        const int ne = min(cape, E - e0);
        stage_entries<K, XT>(A.cols + ebase + e0, ne, X, ldx, s_val);
        __syncthreads();
#pragma unroll
        for (int p = 0; p < P; p++) {
            if (!pv[p]) continue;
            const int lo = max(s_start[prow[p]] - ebase, e0);
            const int hi = min(s_start[prow[p] + 1] - ebase, e0 + ne);
            float a = acc[p], e = cmp[p];
            for (int j = lo; j < hi; j++) cadd(a, e, s_val[(j - e0) * K + pcol[p]]);
            acc[p] = a;
            cmp[p] = e;
        }
        __syncthreads();
    }

    // 4. long rows: one wavefront per row
    long_phase<K, XT, P>(A, X, ldx, u0, rows, lq, pv, acc, cmp, s_lsum, NP);

    // 5. store
#pragma unroll
    for (int p = 0; p < P; p++)
        if (pv[p]) Y[(size_t)(u0 + prow[p]) * ldy + pcol[p]] = acc[p] + cmp[p];
}

// ---- single right-hand side: lane-strided row kernel ----
// A workgroup handles a tile of 256*RPT rows; lane l of wavefront w owns the
// rows u0 + 256*j + 64*w + l (j < RPT), so every row-indexed access -- local
// offsets, diagonal x, y -- is one coalesced dword per lane, a wavefront's 64
// rows form exactly one offset group (their base is one scalar load), and a
// row's first entry is its left neighbour's end. The tile's entry range is
// two scalar loads (tiles are 64-row aligned). All loads of a phase are
// issued before the first use. Same canonical order as run_tile. Column c of
// X is X[c * ldx], row r of Y is Y[r * ldy] (1 and 1 for a plain vector).
// LDS: s_val[capv] | s_lsum[TRW] (LONG) | s_win[TRW] (FAST).
template <typename XT, int RPT, int OW, bool FAST, bool LONG>
__device__ __forceinline__ void rows_tile(const MatView &A, const XT *__restrict__ X, int ldx, float *__restrict__ Y,
                                          int ldy, int t, float *smem, int capv)
{
    constexpr int TRW = kWG * RPT;
    float *s_val = smem;
    float *s_lsum = smem + capv;
    float *s_win = smem + capv + (LONG ? TRW : 0);
    const int tid = threadIdx.x, lane = tid & 63;
    const uint8_t *ro8 = (const uint8_t *)A.roff;
    const uint16_t *ro16 = (const uint16_t *)A.roff;
    {
        const int u0 = t * TRW;
        const int rows = min(TRW, A.M - u0);
        const int ebase = cload(A.gbase, u0 >> 6);
        const int E = cload(A.gbase, u0 + TRW >= A.M ? A.ngroups : (u0 + TRW) >> 6) - ebase;

        // Issue order: x window (FAST), row offsets, the first batch of the
        // column stream; the adds follow in the canonical order.
        constexpr int U = 4;
        int prow[RPT], pcol[RPT], lq[RPT], endv[RPT], gbv[RPT];
        bool pv[RPT];
        float acc[RPT], cmp[RPT];
#pragma unroll
        for (int j = 0; j < RPT; j++) {
            prow[j] = j * kWG + tid;
            pcol[j] = 0;
            lq[j] = prow[j];
            acc[j] = 0.0f;
            cmp[j] = 0.0f;
            pv[j] = u0 + prow[j] < A.M;
        }
        // FAST (one shape, one offset): x[w0 + i], i < TRW, holds this tile's
        // diagonal terms; it is loaded once, kept in LDS (s_win) and also
        // serves the remainder gathers that fall inside it.
        const int w0 = FAST ? u0 + A.sh0.ys - A.sh0.xs + A.sh0.omin : 0;
        float xw[RPT];
        if constexpr (FAST) {
#pragma unroll
            for (int j = 0; j < RPT; j++) xw[j] = load_x<XT>(X, (size_t)min(max(w0 + prow[j], 0), A.N - 1) * ldx);
        }
#pragma unroll
        for (int j = 0; j < RPT; j++) {
            const int rc = pv[j] ? u0 + prow[j] : A.M - 1;
            // the wavefront's 64 rows are one group: its base is a scalar load
            const int g = __builtin_amdgcn_readfirstlane(min((u0 + j * kWG + (tid & ~63)) >> 6, (A.M - 1) >> 6));
            gbv[j] = cload(A.gbase, g);
            endv[j] = OW == 1 ? (int)ro8[rc] : (int)ro16[rc];
        }
        const int ne0 = min(capv, E);
        int cv[U];
#pragma unroll
        for (int u = 0; u < U; u++) cv[u] = u * kWG < ne0 ? A.cols[ebase + min(u * kWG + tid, ne0 - 1)] : 1;

        if constexpr (FAST) {
            const ShapeDev &sh = A.sh0;
#pragma unroll
            for (int j = 0; j < RPT; j++) {
                s_win[prow[j]] = xw[j];
                const int r = u0 + prow[j], c = w0 + prow[j];
                if (pv[j] && r >= sh.xs && r < sh.xe && c >= sh.ys && c < sh.ys + (sh.xe - sh.xs))
                    cadd(acc[j], cmp[j], xw[j]);
            }
            __syncthreads();
        } else {
            diag_phase<1, XT, RPT>(A, X, ldx, u0, rows, prow, pcol, pv, acc, cmp, s_val, capv);
        }

        // a row starts where its left neighbour ends (DPP wave_shr:1); lane 0
        // starts the group
        int rs[RPT], re[RPT];
#pragma unroll
        for (int j = 0; j < RPT; j++) {
            const int beg = __builtin_amdgcn_update_dpp(0, endv[j], 0x138, 0xF, 0xF, false);
            rs[j] = gbv[j] - ebase + (lane == 0 ? 0 : beg);
            re[j] = pv[j] ? gbv[j] - ebase + endv[j] : rs[j];
        }

        for (int e0 = 0; e0 < E; e0 += capv) {
            const int ne = min(capv, E - e0);
            for (int b = 0; b < ne; b += kWG * U) {
                int vv[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    vv[u] = 1;
                    if (e0 == 0 && b == 0) vv[u] = cv[u];
                    else if (b + u * kWG < ne) vv[u] = A.cols[ebase + e0 + min(b + u * kWG + tid, ne - 1)];
                }
                float xv[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    xv[u] = 0.0f;
                    if (b + u * kWG < ne) {
                        const int col = vv[u] >> 1;
                        if (FAST && (unsigned)(col - w0) < (unsigned)TRW) xv[u] = s_win[col - w0];
                        else xv[u] = load_x<XT>(X, (size_t)col * ldx);
                    }
                }
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int idx = b + u * kWG + tid;
                    if (idx < ne) s_val[idx] = (vv[u] & 1) ? xv[u] : -xv[u];
                }
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < RPT; j++) {
                const int lo = max(rs[j], e0), hi = min(re[j], e0 + ne);
                float a = acc[j], e = cmp[j];
                for (int k = lo; k < hi; k++) cadd(a, e, s_val[k - e0]);
                acc[j] = a;
                cmp[j] = e;
            }
            __syncthreads();
        }

        if constexpr (LONG) long_phase<1, XT, RPT>(A, X, ldx, u0, rows, lq, pv, acc, cmp, s_lsum, TRW);

#pragma unroll
        for (int j = 0; j < RPT; j++) {
            if (!pv[j]) continue;
            Y[(size_t)(u0 + prow[j]) * ldy] = acc[j] + cmp[j];
        }
    }
}

template <typename XT, int RPT, int OW, bool FAST, bool LONG>
__global__ __launch_bounds__(kWG) void spmv_rows_kernel(MatView A, const XT *__restrict__ X, float *__restrict__ Y,
                                                        int ntiles, int capv)
{
    extern __shared__ float smem[];
    int tb, te;
    block_units(ntiles, &tb, &te);
    for (int t = tb; t < te; t++) rows_tile<XT, RPT, OW, FAST, LONG>(A, X, 1, Y, 1, t, smem, capv);
}

// ---- single right-hand side, quad rows: at most one shape with one offset ----
// Lane l of a workgroup owns the Q consecutive rows u0 + Q*l + i (i < Q) of a
// tile of 256*Q rows. Every row-indexed stream -- local offsets, the x window
// along the diagonal, y -- moves Q elements per lane in one load or store, and
// the tile's remainder entries are read as 16-byte-aligned int4 per lane
// (starting at the 4-entry boundary below the tile's first entry). Against
// one row per lane this cuts the vector memory instructions per row about
// threefold; on R-MAT the one-row-per-lane kernel is bound by their issue
// (texture addresser busy most of the kernel), not by bytes. Same canonical
// per-row order as every other kernel. Requires X and Y aligned to Q
// elements (checked on the host).
// LDS: s_val[capv] (also the long-row sums afterwards) | s_win[TRW] (DIAG).
template <typename XT, int Q>
__device__ __forceinline__ void load_q(const XT *__restrict__ p, float (&o)[Q])
{
    if constexpr (sizeof(XT) == 4) {
        if constexpr (Q == 4) {
            const float4 v = *(const float4 *)p;
            o[0] = v.x, o[1] = v.y, o[2] = v.z, o[3] = v.w;
        } else {
            const float2 v = *(const float2 *)p;
            o[0] = v.x, o[1] = v.y;
        }
    } else {
        uint32_t w[Q / 2];
        if constexpr (Q == 4) {
            const uint2 v = *(const uint2 *)p;
            w[0] = v.x, w[1] = v.y;
        } else {
            w[0] = *(const uint32_t *)p;
        }
#pragma unroll
        for (int i = 0; i < Q; i++) {
            const uint16_t h = (uint16_t)(w[i >> 1] >> (16 * (i & 1)));
            if constexpr (sizeof(XT) == 2 && __is_same(XT, bf16_t)) o[i] = __uint_as_float((uint32_t)h << 16);
            else o[i] = (float)__builtin_bit_cast(_Float16, h);
        }
    }
}

template <int OW, int Q>
__device__ __forceinline__ void load_ends(const void *roff, int r0, int (&e)[Q])
{
    if constexpr (OW == 1) {
        const uint32_t w = Q == 4 ? *(const uint32_t *)((const uint8_t *)roff + r0)
                                  : (uint32_t) * (const uint16_t *)((const uint8_t *)roff + r0);
#pragma unroll
        for (int i = 0; i < Q; i++) e[i] = (int)((w >> (8 * i)) & 0xFF);
    } else {
        uint32_t w[Q / 2];
        if constexpr (Q == 4) {
            const uint2 v = *(const uint2 *)((const uint16_t *)roff + r0);
            w[0] = v.x, w[1] = v.y;
        } else {
            w[0] = *(const uint32_t *)((const uint16_t *)roff + r0);
        }
#pragma unroll
        for (int i = 0; i < Q; i++) e[i] = (int)((w[i >> 1] >> (16 * (i & 1))) & 0xFFFF);
    }
}

template <typename XT, int Q, int OW, bool DIAG, bool LONG>
__device__ __forceinline__ void quad_tile(const MatView &A, const XT *__restrict__ X, float *__restrict__ Y, int t,
                                          float *smem, int capv)
{
    constexpr int TRW = kWG * Q;
    float *s_val = smem;
    float *s_win = smem + capv;
    const int tid = threadIdx.x, lane = tid & 63;
    const int u0 = t * TRW;
    const int r0 = u0 + Q * tid;  // this lane's first row
    const bool any = r0 < A.M;
    const bool full = r0 + Q <= A.M;

    // the tile's entries [ebase, eend), streamed from the 4-aligned lb
    const int ebase = cload(A.gbase, u0 >> 6);
    const int eend = cload(A.gbase, u0 + TRW >= A.M ? A.ngroups : (u0 + TRW) >> 6);
    const int lb = ebase & ~3;
    const int EA = eend - lb;

    // issue order: row ends, the x window, the first int4 of the stream
    int endv[Q];
#pragma unroll
    for (int i = 0; i < Q; i++) endv[i] = 0;
    if (any) load_ends<OW, Q>(A.roff, r0, endv);
    const int G = A.ngroups;
    const int gw = __builtin_amdgcn_readfirstlane((u0 + Q * 64 * (tid >> 6)) >> 6);  // the wavefront's first group
    int gb;
    if constexpr (Q == 4) {
        const int g0 = cload(A.gbase, min(gw, G)), g1 = cload(A.gbase, min(gw + 1, G));
        const int g2 = cload(A.gbase, min(gw + 2, G)), g3 = cload(A.gbase, min(gw + 3, G));
        gb = lane < 16 ? g0 : lane < 32 ? g1 : lane < 48 ? g2 : g3;
    } else {
        const int g0 = cload(A.gbase, min(gw, G)), g1 = cload(A.gbase, min(gw + 1, G));
        gb = lane < 32 ? g0 : g1;
    }
    const int w0 = DIAG ? u0 + A.sh0.ys - A.sh0.xs + A.sh0.omin : 0;  // x index of local row 0's diagonal term
    float xw[Q];
    if constexpr (DIAG) {
        const int c0 = w0 + Q * tid;
        if ((w0 & (Q - 1)) == 0 && c0 >= 0 && c0 + Q <= A.N) {
            load_q<XT, Q>(X + c0, xw);
        } else {
#pragma unroll
            for (int i = 0; i < Q; i++) xw[i] = load_x<XT>(X, (size_t)min(max(c0 + i, 0), A.N - 1));
        }
    }
    int4 cv = make_int4(1, 1, 1, 1);
    if (4 * tid < min(EA, capv)) cv = *(const int4 *)(A.cols + lb + 4 * tid);

    float acc[Q], cmp[Q];
#pragma unroll
    for (int i = 0; i < Q; i++) acc[i] = 0.0f, cmp[i] = 0.0f;
    if constexpr (DIAG) {
        const ShapeDev &sh = A.sh0;
#pragma unroll
        for (int i = 0; i < Q; i++) {
            const int r = r0 + i, c = w0 + Q * tid + i;
            if (r < A.M && r >= sh.xs && r < sh.xe && c >= sh.ys && c < sh.ys + (sh.xe - sh.xs))
                cadd(acc[i], cmp[i], xw[i]);
        }
        if constexpr (Q == 4) *(float4 *)(s_win + Q * tid) = make_float4(xw[0], xw[1], xw[2], xw[3]);
        else *(float2 *)(s_win + Q * tid) = make_float2(xw[0], xw[1]);
        __syncthreads();
    }

    // row i of this lane: entries [rs[i], re[i]) counted from lb; a row
    // starts where its left neighbour ends (DPP wave_shr:1), a 64-row group
    // at 0
    const int prev = __builtin_amdgcn_update_dpp(0, endv[Q - 1], 0x138, 0xF, 0xF, false);
    int rs[Q], re[Q];
    rs[0] = gb - lb + (((Q * lane) & 63) == 0 ? 0 : prev);
#pragma unroll
    for (int i = 0; i < Q; i++) {
        re[i] = r0 + i < A.M ? gb - lb + endv[i] : rs[i];
        if (i + 1 < Q) rs[i + 1] = re[i];
    }

    for (int e0 = 0; e0 < EA; e0 += capv) {
        const int ne = min(capv, EA - e0);
        for (int b = 0; b < ne; b += kWG * 4) {
            const int q = b + 4 * tid;
            int4 v = make_int4(1, 1, 1, 1);
            if (e0 == 0 && b == 0) v = cv;
            else if (q < ne) v = *(const int4 *)(A.cols + lb + e0 + q);
            const int vv[4] = {v.x, v.y, v.z, v.w};
            float xv[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int col = vv[k] >> 1;
                xv[k] = 0.0f;
                if (DIAG && (unsigned)(col - w0) < (unsigned)TRW) xv[k] = s_win[col - w0];
                else if (q + k < ne) xv[k] = load_x<XT>(X, (size_t)col);
            }
            if (q < ne)
                *(float4 *)(s_val + q) = make_float4((vv[0] & 1) ? xv[0] : -xv[0], (vv[1] & 1) ? xv[1] : -xv[1],
                                                     (vv[2] & 1) ? xv[2] : -xv[2], (vv[3] & 1) ? xv[3] : -xv[3]);
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < Q; i++) {
            const int lo = max(rs[i], e0), hi = min(re[i], e0 + ne);
            float a = acc[i], e = cmp[i];
            for (int k = lo; k < hi; k++) cadd(a, e, s_val[k - e0]);
            acc[i] = a;
            cmp[i] = e;
        }
        __syncthreads();
    }

    if constexpr (LONG) {
        int prow[Q], lq[Q];
        bool pv[Q];
#pragma unroll
        for (int i = 0; i < Q; i++) prow[i] = Q * tid + i, lq[i] = prow[i], pv[i] = r0 + i < A.M;
        long_phase<1, XT, Q>(A, X, 1, u0, min(TRW, A.M - u0), lq, pv, acc, cmp, s_val, TRW);
    }

    if (full) {
        if constexpr (Q == 4)
            *(float4 *)(Y + r0) = make_float4(acc[0] + cmp[0], acc[1] + cmp[1], acc[2] + cmp[2], acc[3] + cmp[3]);
        else *(float2 *)(Y + r0) = make_float2(acc[0] + cmp[0], acc[1] + cmp[1]);
    } else {
#pragma unroll
        for (int i = 0; i < Q; i++)
            if (r0 + i < A.M) Y[r0 + i] = acc[i] + cmp[i];
    }
}

template <typename XT, int Q, int OW, bool DIAG, bool LONG>
__global__ __launch_bounds__(kWG) void spmv_quad_kernel(MatView A, const XT *__restrict__ X, float *__restrict__ Y,
                                                        int ntiles, int capv)
{
    extern __shared__ float smem[];
    int tb, te;
    block_units(ntiles, &tb, &te);
    for (int t = tb; t < te; t++) {
        quad_tile<XT, Q, OW, DIAG, LONG>(A, X, Y, t, smem, capv);
        if (DIAG && t + 1 < te) __syncthreads();  // s_win is rewritten by the next tile
    }
}

// ---- k > 1: wide rows ----
// A row belongs to a group of LPR = K/CPL lanes; lane q of the group owns the
// CPL = min(K, 4) consecutive columns q*CPL.., so each x row segment the row
// touches (a diagonal term or a remainder entry) is one CPL-wide vector load
// per lane and y leaves as one vector store per lane. Consecutive groups own
// consecutive rows (y is written contiguously when ldy = K). Rows are
// independent: no LDS staging; only long rows (one wavefront each, the fixed
// butterfly) meet in LDS. Same canonical per-(row, column) order as every
// other kernel. Requires X, Y, ldx and ldy aligned to CPL elements.
// LDS: the x window [(TRW + kWideHalo) * K], reused for the long-row sums.
constexpr int kWideHalo = 64;  // window rows beyond the tile's that LDS holds (offset spread)

template <int K, typename XT, int RI, bool LONG>
__device__ __forceinline__ void wide_tile(const MatView &A, const XT *__restrict__ X, int ldx, float *__restrict__ Y,
                                          int ldy, int t, float *smem)
{
    constexpr int CPL = K < 4 ? K : 4;
    constexpr int LPR = K / CPL;
    constexpr int GPW = kWG / LPR;  // row groups per workgroup
    constexpr int TRW = GPW * RI;   // rows per tile
    constexpr int P = CPL * RI;     // (row, column) pairs per lane
    constexpr int U = 4;            // remainder entries in flight per lane
    const int tid = threadIdx.x;
    const int g = tid / LPR, c0 = (tid % LPR) * CPL;
    const uint8_t *ro8 = (const uint8_t *)A.roff;
    const uint16_t *ro16 = (const uint16_t *)A.roff;
    {
        const int u0 = t * TRW;
        float acc[P], cmp[P];
        int lq[P];
        bool pv[P];
        int rs[RI], re[RI];
#pragma unroll
        for (int i = 0; i < RI; i++) {
            const int r = u0 + GPW * i + g;
            const bool valid = r < A.M;
            const int rr = valid ? r : A.M - 1;
            const int gb = A.gbase[rr >> 6];
            const int prev = (rr & 63) ? (A.ow == 1 ? (int)ro8[rr - 1] : (int)ro16[rr - 1]) : 0;
            const int end = A.ow == 1 ? (int)ro8[rr] : (int)ro16[rr];
            rs[i] = gb + prev;
            re[i] = valid ? gb + end : rs[i];
#pragma unroll
            for (int m = 0; m < CPL; m++) {
                acc[i * CPL + m] = 0.0f;
                cmp[i * CPL + m] = 0.0f;
                lq[i * CPL + m] = (GPW * i + g) * K + c0 + m;
                pv[i * CPL + m] = valid;
            }
        }

        // diagonal terms: shapes in order, offsets in order; the rows' loads
        // of one offset are issued together. A shape with two or more
        // offsets first stages the tile's x window (rows [w0, w1), K columns)
        // in LDS, so each x row segment is loaded once per tile, not once
        // per offset.
        for (int s = 0; s < A.S; s++) {
            const ShapeDev sh = s == 0 ? A.sh0 : load_shape(A, s);
            if (u0 + TRW <= sh.xs || u0 >= sh.xe) continue;
            const int side = sh.xe - sh.xs;
            const int shift = sh.ys - sh.xs;
            const int ra = max(u0, sh.xs), rb = min(min(u0 + TRW, sh.xe), A.M);
            const int w0 = max(sh.ys, ra + shift + sh.omin);
            const int w1 = min(sh.ys + side, rb - 1 + shift + sh.omax + 1);
            if (sh.oe - sh.ob >= 2 && w1 > w0 && w1 - w0 <= TRW + kWideHalo) {
                const int nv = (w1 - w0) * LPR;  // CPL-wide vectors of the window
                for (int b = 0; b < nv; b += kWG * 4) {
                    float tv[4][CPL];
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const int v = min(b + u * kWG + tid, nv - 1);
                        load_q<XT, CPL>(X + (size_t)(w0 + v / LPR) * ldx + (v % LPR) * CPL, tv[u]);
                    }
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const int v = b + u * kWG + tid;
                        if (v < nv) {
                            float *d = smem + (v / LPR) * K + (v % LPR) * CPL;
                            if constexpr (CPL == 4) *(float4 *)d = make_float4(tv[u][0], tv[u][1], tv[u][2], tv[u][3]);
                            else *(float2 *)d = make_float2(tv[u][0], tv[u][1]);
                        }
                    }
                }
                __syncthreads();
                for (int n0 = sh.ob; n0 < sh.oe; n0 += kDiagBlock) {
                    const int n1 = min(n0 + kDiagBlock, sh.oe);
                    float bs[P];
#pragma unroll
                    for (int p = 0; p < P; p++) bs[p] = 0.0f;
                    for (int n = n0; n < n1; n++) {
                        const int o = cload(A.offs, n);
#pragma unroll
                        for (int i = 0; i < RI; i++) {
                            const int r = u0 + GPW * i + g;
                            const int c = r + shift + o;
                            if (r < A.M && r >= sh.xs && r < sh.xe && c >= sh.ys && c < sh.ys + side) {
                                const float *sp = smem + (c - w0) * K + c0;
                                if constexpr (CPL == 4) {
                                    const float4 q = *(const float4 *)sp;
                                    bs[i * 4] += q.x, bs[i * 4 + 1] += q.y, bs[i * 4 + 2] += q.z, bs[i * 4 + 3] += q.w;
                                } else {
                                    const float2 q = *(const float2 *)sp;
                                    bs[i * 2] += q.x, bs[i * 2 + 1] += q.y;
                                }
                            }
                        }
                    }
#pragma unroll
                    for (int p = 0; p < P; p++) cadd(acc[p], cmp[p], bs[p]);
                }
                __syncthreads();
                continue;
            }
            for (int n0 = sh.ob; n0 < sh.oe; n0 += kDiagBlock) {
                const int n1 = min(n0 + kDiagBlock, sh.oe);
                float bs[P];
#pragma unroll
                for (int p = 0; p < P; p++) bs[p] = 0.0f;
                for (int n = n0; n < n1; n++) {
                    const int o = cload(A.offs, n);
                    float xv[RI][CPL];
                    bool in[RI];
#pragma unroll
                    for (int i = 0; i < RI; i++) {
                        const int r = u0 + GPW * i + g;
                        const int c = r - sh.xs + sh.ys + o;
                        in[i] = r < A.M && r >= sh.xs && r < sh.xe && c >= sh.ys && c < sh.ys + side;
                        if (in[i]) load_q<XT, CPL>(X + (size_t)c * ldx + c0, xv[i]);
                    }
#pragma unroll
                    for (int i = 0; i < RI; i++)
                        if (in[i])
#pragma unroll
                            for (int m = 0; m < CPL; m++) bs[i * CPL + m] += xv[i][m];
                }
#pragma unroll
                for (int p = 0; p < P; p++) cadd(acc[p], cmp[p], bs[p]);
            }
        }

        // remainder entries, U in flight, added in stored order
#pragma unroll
        for (int i = 0; i < RI; i++) {
            for (int j0 = rs[i]; j0 < re[i]; j0 += U) {
                int v[U];
                float xv[U][CPL];
#pragma unroll
                for (int u = 0; u < U; u++) v[u] = j0 + u < re[i] ? A.cols[j0 + u] : 1;
#pragma unroll
                for (int u = 0; u < U; u++)
                    if (j0 + u < re[i]) load_q<XT, CPL>(X + (size_t)(v[u] >> 1) * ldx + c0, xv[u]);
#pragma unroll
                for (int u = 0; u < U; u++)
                    if (j0 + u < re[i])
#pragma unroll
                        for (int m = 0; m < CPL; m++)
                            cadd(acc[i * CPL + m], cmp[i * CPL + m], (v[u] & 1) ? xv[u][m] : -xv[u][m]);
            }
        }

        if constexpr (LONG) long_phase<K, XT, P>(A, X, ldx, u0, min(TRW, A.M - u0), lq, pv, acc, cmp, smem, TRW * K);

#pragma unroll
        for (int i = 0; i < RI; i++) {
            const int r = u0 + GPW * i + g;
            if (r >= A.M) continue;
            float *yp = Y + (size_t)r * ldy + c0;
            if constexpr (CPL == 4)
                *(float4 *)yp = make_float4(acc[i * 4] + cmp[i * 4], acc[i * 4 + 1] + cmp[i * 4 + 1],
                                            acc[i * 4 + 2] + cmp[i * 4 + 2], acc[i * 4 + 3] + cmp[i * 4 + 3]);
            else *(float2 *)yp = make_float2(acc[i * 2] + cmp[i * 2], acc[i * 2 + 1] + cmp[i * 2 + 1]);
        }
    }
}

template <int K, typename XT, int RI, bool LONG>
__global__ __launch_bounds__(kWG) void spmv_wide_kernel(MatView A, const XT *__restrict__ X, int ldx,
                                                        float *__restrict__ Y, int ldy, int ntiles)
{
    extern __shared__ float smem[];
    int tb, te;
    block_units(ntiles, &tb, &te);
    for (int t = tb; t < te; t++) wide_tile<K, XT, RI, LONG>(A, X, ldx, Y, ldy, t, smem);
}

// LDS: s_val[capv] | s_lsum[256*P] | s_start[TRW+1]
__host__ __device__ constexpr size_t lds_bytes(int capv, int P, int trw)
{
    return ((size_t)capv + (size_t)kWG * P + (size_t)trw + 1) * 4;
}

template <int K, typename XT, int P>
__global__ __launch_bounds__(kWG) void spmv_kernel(MatView A, const XT *__restrict__ X, int ldx, float *__restrict__ Y,
                                                   int ldy, int nunits, int capv)
{
    extern __shared__ float smem[];
    constexpr int TRW = (kWG * P) / K;
    float *s_val = smem;
    float *s_lsum = s_val + capv;
    int *s_start = (int *)(s_lsum + kWG * P);
    int ub, ue;
    block_units(nunits, &ub, &ue);
    for (int u = ub; u < ue; u++) {
        run_tile<K, XT, P>(A, X, ldx, Y, ldy, u * TRW, s_val, capv, s_start, s_lsum);
        __syncthreads();
    }
}

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
constexpr int kBatchRPT = 2;     // K = 1: rows_tile with 512 rows a tile
constexpr int kBatchCapv = 4096; // staged values per pass
constexpr int kBatchLdsDesc = 512;

struct WorkDesc {
    MatView A;
    const void *X;
    float *Y;
    int ldx, ldy;
    int unit_begin;
    int fast;  // one shape with one offset: the rows tile's FAST path
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

template <typename XT>
__global__ __launch_bounds__(kWG) void batch_rows_kernel(const WorkDesc *__restrict__ D, const int *__restrict__ UB,
                                                         int nd, int total, int capv)
{
    extern __shared__ float smem[];
    __shared__ int s_ub[kBatchLdsDesc];
    batch_loop(D, UB, nd, total, s_ub, [&](const WorkDesc &w, int t) {
        const XT *X = (const XT *)w.X;
        const bool lng = w.A.L > 0;
        if (w.A.ow == 1) {
            if (w.fast) lng ? rows_tile<XT, kBatchRPT, 1, true, true>(w.A, X, w.ldx, w.Y, w.ldy, t, smem, capv)
                            : rows_tile<XT, kBatchRPT, 1, true, false>(w.A, X, w.ldx, w.Y, w.ldy, t, smem, capv);
            else lng ? rows_tile<XT, kBatchRPT, 1, false, true>(w.A, X, w.ldx, w.Y, w.ldy, t, smem, capv)
                     : rows_tile<XT, kBatchRPT, 1, false, false>(w.A, X, w.ldx, w.Y, w.ldy, t, smem, capv);
        } else {
            if (w.fast) lng ? rows_tile<XT, kBatchRPT, 2, true, true>(w.A, X, w.ldx, w.Y, w.ldy, t, smem, capv)
                            : rows_tile<XT, kBatchRPT, 2, true, false>(w.A, X, w.ldx, w.Y, w.ldy, t, smem, capv);
            else lng ? rows_tile<XT, kBatchRPT, 2, false, true>(w.A, X, w.ldx, w.Y, w.ldy, t, smem, capv)
                     : rows_tile<XT, kBatchRPT, 2, false, false>(w.A, X, w.ldx, w.Y, w.ldy, t, smem, capv);
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
    float *s_lsum = s_val + capv;
    int *s_start = (int *)(s_lsum + kWG * kBatchP);
    batch_loop(D, UB, nd, total, s_ub, [&](const WorkDesc &w, int t) {
        run_tile<K, XT, kBatchP>(w.A, (const XT *)w.X, w.ldx, w.Y, w.ldy, t * TRW, s_val, capv, s_start, s_lsum);
    });
}

// K > 1 chunks whose X and Y rows split into aligned vectors: wide-rows tiles
constexpr int kBatchRI = 1;  // one row per lane group: 42-56 VGPRs, full occupancy

template <int K, typename XT>
__global__ __launch_bounds__(kWG) void batch_wide_kernel(const WorkDesc *__restrict__ D, const int *__restrict__ UB,
                                                         int nd, int total)
{
    extern __shared__ float smem[];
    __shared__ int s_ub[kBatchLdsDesc];
    batch_loop(D, UB, nd, total, s_ub, [&](const WorkDesc &w, int t) {
        if (w.A.L > 0) wide_tile<K, XT, kBatchRI, true>(w.A, (const XT *)w.X, w.ldx, w.Y, w.ldy, t, smem);
        else wide_tile<K, XT, kBatchRI, false>(w.A, (const XT *)w.X, w.ldx, w.Y, w.ldy, t, smem);
    });
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

#define HIPCHECK(expr)                                                                                      \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess) return tc::fail(TC_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));          \
    } while (0)

struct tc_matrix {
    int device = 0;
    MatView v{};
    void *mem = nullptr;
    size_t mem_bytes = 0;
    tc_matrix_stats st{};
    std::mutex mu;
    std::map<std::pair<int, int>, tc_config> cfg;
};

struct tc_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

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

int pow2_floor(int x)
{
    int p = 1;
    while (p * 2 <= x) p *= 2;
    return p;
}

// supported compile-time column counts, largest first
const int kKs[] = {64, 32, 16, 8, 7, 6, 5, 4, 3, 2, 1};

int pick_k(int k)
{
    for (int K : kKs)
        if (K <= k) return K;
    return 1;
}

// k = 1: tiles of 512 rows and a 4096-value stage; k > 1: wide-rows tiles
// of one row per lane group (full occupancy): the best or near-best settings
// measured on R-MAT, banded and stencil matrices. Matrices whose remainder
// averages more than 8 entries a row (banded with many zero cells) use
// 256-row tiles at k = 1, so a tile's entries pass through the stage in
// about two rounds.
// k > 1: the wide-rows kernel serves K in {2, 4, 8, 16, 32, 64} unless
// config flag bit 2 (kStagedK) asks for the staged tile kernel, which also
// serves every other K.
constexpr int kStagedK = 4;

bool wide_k(int K) { return K == 2 || K == 4 || K == 8 || K == 16 || K == 32 || K == 64; }

// rows of a wide-rows tile with RI rows per lane group
int wide_rows(int K, int RI) { return (kWG / (K / std::min(K, 4))) * RI; }

constexpr int kQuadOff = 2;  // config flag: use the one-row-per-lane kernel at k = 1

tc_config default_config(int K, long long remainder, int M)
{
    tc_config c{};
    const bool heavy = M > 0 && remainder > 8ll * M;
    c.rows_per_tile = K == 1 ? (heavy ? 256 : 512) : wide_k(K) ? wide_rows(K, 1) : (kWG * 2) / K;
    c.lds_values = 4096;
    c.grid = 0;
    c.flags = K == 1 ? kQuadOff : 0;  // quad rows measured no faster on R-MAT; the autotuner tries both
    return c;
}

tc_config default_config(const tc_matrix *A, int K) { return default_config(K, A->st.remainder, A->v.M); }

bool rows_per_tile_ok(int K, int r, int flags)
{
    if (K == 1) return r == 256 || r == 512 || r == 1024 || r == 2048 || r == 4096;
    if (wide_k(K) && !(flags & kStagedK)) return r == wide_rows(K, 1) || r == wide_rows(K, 2);
    for (int P : {1, 2, 4})
        if ((kWG * P) / K == r) return true;
    return false;
}

template <typename XT, int RPT, int OW, bool FAST, bool LONG>
void launch_rows_k(const MatView &A, const void *X, float *Y, int ntiles, int capv, int G, size_t lds, hipStream_t s)
{
    spmv_rows_kernel<XT, RPT, OW, FAST, LONG><<<G, kWG, lds, s>>>(A, (const XT *)X, Y, ntiles, capv);
}

template <typename XT, int RPT>
int launch_rows_t(const MatView &A, const void *X, float *Y, int capv, int grid, hipStream_t s)
{
    constexpr int TRW = kWG * RPT;
    const int ntiles = (A.M + TRW - 1) / TRW;
    if (ntiles == 0) return TC_OK;
    const int G = grid > 0 ? std::min(grid, ntiles) : ntiles;
    const bool fast = A.S == 1 && A.sh0.oe - A.sh0.ob == 1;
    const size_t lds = ((size_t)capv + (A.L > 0 ? TRW : 0) + (fast ? TRW : 0)) * 4;
    const bool lng = A.L > 0;
    // specialised on offset width, the one-shape-one-offset case and long
    // rows, so the common case carries no registers for the general paths
    if (A.ow == 1) {
        if (fast) lng ? launch_rows_k<XT, RPT, 1, true, true>(A, X, Y, ntiles, capv, G, lds, s)
                      : launch_rows_k<XT, RPT, 1, true, false>(A, X, Y, ntiles, capv, G, lds, s);
        else lng ? launch_rows_k<XT, RPT, 1, false, true>(A, X, Y, ntiles, capv, G, lds, s)
                 : launch_rows_k<XT, RPT, 1, false, false>(A, X, Y, ntiles, capv, G, lds, s);
    } else {
        if (fast) lng ? launch_rows_k<XT, RPT, 2, true, true>(A, X, Y, ntiles, capv, G, lds, s)
                      : launch_rows_k<XT, RPT, 2, true, false>(A, X, Y, ntiles, capv, G, lds, s);
        else lng ? launch_rows_k<XT, RPT, 2, false, true>(A, X, Y, ntiles, capv, G, lds, s)
                 : launch_rows_k<XT, RPT, 2, false, false>(A, X, Y, ntiles, capv, G, lds, s);
    }
    HIPCHECK(hipGetLastError());
    return TC_OK;
}

template <typename XT>
int launch_rows(const MatView &A, const void *X, float *Y, const tc_config &c, int capv, hipStream_t s)
{
    switch (c.rows_per_tile) {
    case 256: return launch_rows_t<XT, 1>(A, X, Y, capv, c.grid, s);
    case 512: return launch_rows_t<XT, 2>(A, X, Y, capv, c.grid, s);
    case 1024: return launch_rows_t<XT, 4>(A, X, Y, capv, c.grid, s);
    case 4096: return launch_rows_t<XT, 16>(A, X, Y, capv, c.grid, s);
    default: return launch_rows_t<XT, 8>(A, X, Y, capv, c.grid, s);
    }
}

template <typename XT, int Q, int OW, bool DIAG, bool LONG>
void launch_quad_k(const MatView &A, const void *X, float *Y, int ntiles, int capv, int G, size_t lds, hipStream_t s)
{
    spmv_quad_kernel<XT, Q, OW, DIAG, LONG><<<G, kWG, lds, s>>>(A, (const XT *)X, Y, ntiles, capv);
}

template <typename XT, int Q>
int launch_quad_t(const MatView &A, const void *X, float *Y, int capv, int grid, hipStream_t s)
{
    constexpr int TRW = kWG * Q;
    const int ntiles = (A.M + TRW - 1) / TRW;
    if (ntiles == 0) return TC_OK;
    const int G = grid > 0 ? std::min(grid, ntiles) : ntiles;
    const bool diag = A.S == 1;
    const bool lng = A.L > 0;
    const size_t lds = ((size_t)capv + (diag ? TRW : 0)) * 4;
    if (A.ow == 1) {
        if (diag) lng ? launch_quad_k<XT, Q, 1, true, true>(A, X, Y, ntiles, capv, G, lds, s)
                      : launch_quad_k<XT, Q, 1, true, false>(A, X, Y, ntiles, capv, G, lds, s);
        else lng ? launch_quad_k<XT, Q, 1, false, true>(A, X, Y, ntiles, capv, G, lds, s)
                 : launch_quad_k<XT, Q, 1, false, false>(A, X, Y, ntiles, capv, G, lds, s);
    } else {
        if (diag) lng ? launch_quad_k<XT, Q, 2, true, true>(A, X, Y, ntiles, capv, G, lds, s)
                      : launch_quad_k<XT, Q, 2, true, false>(A, X, Y, ntiles, capv, G, lds, s);
        else lng ? launch_quad_k<XT, Q, 2, false, true>(A, X, Y, ntiles, capv, G, lds, s)
                 : launch_quad_k<XT, Q, 2, false, false>(A, X, Y, ntiles, capv, G, lds, s);
    }
    HIPCHECK(hipGetLastError());
    return TC_OK;
}

// Quad-row kernel for k = 1 when it applies: no shape or one shape with one
// offset, tiles of 512 or 1024 rows, X and Y aligned to the rows per lane,
// and not switched off by config flag bit 1 (kQuadOff).

bool quad_applies(const MatView &A, const void *X, const float *Y, int esize, const tc_config &c, int capv)
{
    if (c.flags & kQuadOff) return false;
    if (!(A.S == 0 || (A.S == 1 && A.sh0.oe - A.sh0.ob == 1))) return false;
    const int Q = c.rows_per_tile == 512 ? 2 : c.rows_per_tile == 1024 ? 4 : 0;
    if (Q == 0 || capv % 4 != 0 || capv < kWG * Q) return false;
    return ((uintptr_t)X % ((uintptr_t)Q * esize)) == 0 && ((uintptr_t)Y % ((uintptr_t)Q * 4)) == 0;
}

template <typename XT>
int launch_quad(const MatView &A, const void *X, float *Y, const tc_config &c, int capv, hipStream_t s)
{
    if (c.rows_per_tile == 1024) return launch_quad_t<XT, 4>(A, X, Y, capv, c.grid, s);
    return launch_quad_t<XT, 2>(A, X, Y, capv, c.grid, s);
}

// P from a rows_per_tile request for column count K (1, 2 or 4)
int p_for(int K, int rows_per_tile)
{
    int best = 2;
    for (int P : {1, 2, 4})
        if ((kWG * P) / K == rows_per_tile) best = P;
    return best;
}

template <int K, typename XT, int P>
int launch_t(const MatView &A, const void *X, int ldx, float *Y, int ldy, int capv, int grid, hipStream_t s)
{
    constexpr int TRW = (kWG * P) / K;
    const int nunits = (A.M + TRW - 1) / TRW;
    if (nunits == 0) return TC_OK;
    int G = grid > 0 ? std::min(grid, nunits) : nunits;
    const size_t lds = lds_bytes(capv, P, TRW);
    spmv_kernel<K, XT, P><<<G, kWG, lds, s>>>(A, (const XT *)X, ldx, Y, ldy, nunits, capv);
    HIPCHECK(hipGetLastError());
    return TC_OK;
}

template <int K, typename XT>
int launch_p(const MatView &A, const void *X, int ldx, float *Y, int ldy, int P, int capv, int grid, hipStream_t s)
{
    if (P == 1) return launch_t<K, XT, 1>(A, X, ldx, Y, ldy, capv, grid, s);
    if (P == 4) return launch_t<K, XT, 4>(A, X, ldx, Y, ldy, capv, grid, s);
    return launch_t<K, XT, 2>(A, X, ldx, Y, ldy, capv, grid, s);
}

template <typename XT>
int launch_k(const MatView &A, const void *X, int ldx, float *Y, int ldy, int K, int P, int capv, int grid,
             hipStream_t s)
{
    switch (K) {
    case 1: return launch_p<1, XT>(A, X, ldx, Y, ldy, P, capv, grid, s);
    case 2: return launch_p<2, XT>(A, X, ldx, Y, ldy, P, capv, grid, s);
    case 3: return launch_p<3, XT>(A, X, ldx, Y, ldy, P, capv, grid, s);
    case 4: return launch_p<4, XT>(A, X, ldx, Y, ldy, P, capv, grid, s);
    case 5: return launch_p<5, XT>(A, X, ldx, Y, ldy, P, capv, grid, s);
    case 6: return launch_p<6, XT>(A, X, ldx, Y, ldy, P, capv, grid, s);
    case 7: return launch_p<7, XT>(A, X, ldx, Y, ldy, P, capv, grid, s);
    case 8: return launch_p<8, XT>(A, X, ldx, Y, ldy, P, capv, grid, s);
    case 16: return launch_p<16, XT>(A, X, ldx, Y, ldy, P, capv, grid, s);
    case 32: return launch_p<32, XT>(A, X, ldx, Y, ldy, P, capv, grid, s);
    default: return launch_p<64, XT>(A, X, ldx, Y, ldy, P, capv, grid, s);
    }
}

int elem_size(int dtype) { return dtype == TC_F32 ? 4 : 2; }

tc_config get_cfg(tc_matrix *A, int K, int dtype)
{
    std::lock_guard<std::mutex> g(A->mu);
    auto it = A->cfg.find({K, dtype});
    return it != A->cfg.end() ? it->second : default_config(A, K);
}

template <int K, typename XT, int RI>
int launch_wide_t(const MatView &A, const void *X, int ldx, float *Y, int ldy, int grid, hipStream_t s)
{
    constexpr int TRW = (kWG / (K / (K < 4 ? K : 4))) * RI;
    const int ntiles = (A.M + TRW - 1) / TRW;
    if (ntiles == 0) return TC_OK;
    const int G = grid > 0 ? std::min(grid, ntiles) : ntiles;
    const size_t lds = (A.S > 0 || A.L > 0) ? (size_t)(TRW + kWideHalo) * K * 4 : 0;
    if (A.L > 0)
        spmv_wide_kernel<K, XT, RI, true><<<G, kWG, lds, s>>>(A, (const XT *)X, ldx, Y, ldy, ntiles);
    else
        spmv_wide_kernel<K, XT, RI, false><<<G, kWG, lds, s>>>(A, (const XT *)X, ldx, Y, ldy, ntiles);
    HIPCHECK(hipGetLastError());
    return TC_OK;
}

template <int K, typename XT>
int launch_wide_k(const MatView &A, const void *X, int ldx, float *Y, int ldy, const tc_config &c, hipStream_t s)
{
    if (c.rows_per_tile == wide_rows(K, 1)) return launch_wide_t<K, XT, 1>(A, X, ldx, Y, ldy, c.grid, s);
    return launch_wide_t<K, XT, 2>(A, X, ldx, Y, ldy, c.grid, s);
}

template <typename XT>
int launch_wide(const MatView &A, const void *X, int ldx, float *Y, int ldy, int K, const tc_config &c,
                hipStream_t s)
{
    switch (K) {
    case 2: return launch_wide_k<2, XT>(A, X, ldx, Y, ldy, c, s);
    case 4: return launch_wide_k<4, XT>(A, X, ldx, Y, ldy, c, s);
    case 8: return launch_wide_k<8, XT>(A, X, ldx, Y, ldy, c, s);
    case 16: return launch_wide_k<16, XT>(A, X, ldx, Y, ldy, c, s);
    case 32: return launch_wide_k<32, XT>(A, X, ldx, Y, ldy, c, s);
    default: return launch_wide_k<64, XT>(A, X, ldx, Y, ldy, c, s);
    }
}

// wide rows apply when X and Y rows split into aligned CPL-element vectors
bool wide_applies(int K, const void *X, int ldx, const float *Y, int ldy, int esize, const tc_config &c)
{
    if (!wide_k(K) || (c.flags & kStagedK)) return false;
    if (c.rows_per_tile != wide_rows(K, 1) && c.rows_per_tile != wide_rows(K, 2)) return false;
    const int cpl = std::min(K, 4);
    return ldx % cpl == 0 && ldy % cpl == 0 && (uintptr_t)X % ((uintptr_t)cpl * esize) == 0 &&
           (uintptr_t)Y % ((uintptr_t)cpl * 4) == 0;
}

// one launch of K columns (K a supported count) with an explicit config
int launch_cfg(tc_matrix *A, const void *X, int ldx, int dtype, float *Y, int ldy, int K, const tc_config &c,
               hipStream_t s)
{
    const int capv = std::max(c.lds_values, 64 * K);
    if (K == 1 && ldx == 1 && ldy == 1 && quad_applies(A->v, X, Y, elem_size(dtype), c, capv)) {
        if (dtype == TC_BF16) return launch_quad<bf16_t>(A->v, X, Y, c, capv, s);
        if (dtype == TC_F16) return launch_quad<f16_t>(A->v, X, Y, c, capv, s);
        return launch_quad<float>(A->v, X, Y, c, capv, s);
    }
    if (K == 1 && ldx == 1 && ldy == 1 && ((uintptr_t)Y & 15) == 0) {
        if (dtype == TC_BF16) return launch_rows<bf16_t>(A->v, X, Y, c, capv, s);
        if (dtype == TC_F16) return launch_rows<f16_t>(A->v, X, Y, c, capv, s);
        return launch_rows<float>(A->v, X, Y, c, capv, s);
    }
    if (K > 1 && wide_applies(K, X, ldx, Y, ldy, elem_size(dtype), c)) {
        if (dtype == TC_BF16) return launch_wide<bf16_t>(A->v, X, ldx, Y, ldy, K, c, s);
        if (dtype == TC_F16) return launch_wide<f16_t>(A->v, X, ldx, Y, ldy, K, c, s);
        return launch_wide<float>(A->v, X, ldx, Y, ldy, K, c, s);
    }
    // staged tile kernel: K=1 configs name the rows kernel's tile; use P = 2
    const int P = K == 1 ? 2 : p_for(K, c.rows_per_tile);
    if (dtype == TC_BF16) return launch_k<bf16_t>(A->v, X, ldx, Y, ldy, K, P, capv, c.grid, s);
    if (dtype == TC_F16) return launch_k<f16_t>(A->v, X, ldx, Y, ldy, K, P, capv, c.grid, s);
    return launch_k<float>(A->v, X, ldx, Y, ldy, K, P, capv, c.grid, s);
}

int check_spmv_args(const tc_matrix *A, const void *X, int ldx, int dtype, const float *Y, int ldy, int k)
{
    if (!A) return tc::fail(TC_ERR_INVALID, "tc_spmv: null matrix");
    if (dtype != TC_F32 && dtype != TC_BF16 && dtype != TC_F16) return tc::fail(TC_ERR_INVALID, "bad dtype %d", dtype);
    if (k < 1 || ldx < k || ldy < k) return tc::fail(TC_ERR_INVALID, "bad k=%d ldx=%d ldy=%d", k, ldx, ldy);
    if ((A->v.N > 0 && !X) || (A->v.M > 0 && !Y)) return tc::fail(TC_ERR_INVALID, "null X or Y");
    return TC_OK;
}

}  // namespace

namespace tc {

// chunks of one request of k columns: 8s, then 4, 2, 1 as needed
int chunk_count(int k) { return k / 8 + ((k & 4) != 0) + ((k & 2) != 0) + (k & 1); }

constexpr int kMaxBatchK = 64;

size_t batch_desc_bytes(int max_requests)
{
    const size_t chunks = (size_t)max_requests * chunk_count(kMaxBatchK - 1);
    return chunks * sizeof(WorkDesc) + chunks * sizeof(int) + 16;
}

// Batch classes: (kernel, chunk width, x type). Kernel 0: rows tiles (K = 1);
// 1: wide-rows tiles (K > 1, aligned); 2: staged tiles (K > 1, otherwise).
constexpr int kBatchClasses = 3 * 4 * 3;
inline int batch_class(int kern, int ki, int dt) { return (kern * 4 + ki) * 3 + dt; }

int batch_tile_rows(int kern, int K)
{
    if (kern == 0) return kWG * kBatchRPT;
    if (kern == 1) return wide_rows(K, kBatchRI);
    return (kWG * kBatchP) / K;
}

template <typename XT>
void launch_batch_class(int kern, int K, const WorkDesc *D, const int *UB, int nd, int units, hipStream_t s)
{
    const int grid = units;  // one workgroup per tile: the dispatcher balances chunks of any size
    if (kern == 0) {
        constexpr int TRW = kWG * kBatchRPT;
        const size_t lds = ((size_t)kBatchCapv + 2 * TRW) * 4;
        batch_rows_kernel<XT><<<grid, kWG, lds, s>>>(D, UB, nd, units, kBatchCapv);
    } else if (kern == 1) {
        const size_t lds = (size_t)(wide_rows(K, kBatchRI) + kWideHalo) * K * 4;
        if (K == 2) batch_wide_kernel<2, XT><<<grid, kWG, lds, s>>>(D, UB, nd, units);
        else if (K == 4) batch_wide_kernel<4, XT><<<grid, kWG, lds, s>>>(D, UB, nd, units);
        else batch_wide_kernel<8, XT><<<grid, kWG, lds, s>>>(D, UB, nd, units);
    } else {
        const size_t lds = lds_bytes(kBatchCapv, kBatchP, (kWG * kBatchP) / K);
        if (K == 2) batch_tile_kernel<2, XT><<<grid, kWG, lds, s>>>(D, UB, nd, units, kBatchCapv);
        else if (K == 4) batch_tile_kernel<4, XT><<<grid, kWG, lds, s>>>(D, UB, nd, units, kBatchCapv);
        else batch_tile_kernel<8, XT><<<grid, kWG, lds, s>>>(D, UB, nd, units, kBatchCapv);
    }
}

// chunk c0..c0+K of request r: its batch class
int chunk_class(const tc_request &r, int ki, int K, int c0, int dt)
{
    if (K == 1) return batch_class(0, ki, dt);
    const int es = elem_size(r.dtype), cpl = std::min(K, 4);
    const uintptr_t xa = (uintptr_t)r.X + (uintptr_t)c0 * es, ya = (uintptr_t)(r.Y + c0);
    const bool aligned = r.ldx % cpl == 0 && r.ldy % cpl == 0 && xa % ((uintptr_t)cpl * es) == 0 &&
                         ya % ((uintptr_t)cpl * 4) == 0;
    return batch_class(aligned ? 1 : 2, ki, dt);
}

// Fills the chunk descriptors of n requests into host_desc (at least
// batch_desc_bytes(n) bytes, valid until the copy on s has run), copies them
// to dev_desc and enqueues one launch per batch class present. Each
// request's y equals what tc_spmv computes for it alone.
int batch_enqueue(const tc_request *reqs, int n, void *host_desc, void *dev_desc, hipStream_t s)
{
    static const int kChunkK[4] = {1, 2, 4, 8};
    // chunks of a request: 8s, then 4, 2, 1 as k's bits say; calls
    // f(ki, K, c0) for each, widest first
    auto chunks = [](int k, auto &&f) {
        int c0 = 0;
        while (k - c0 >= 8) {
            f(3, 8, c0);
            c0 += 8;
        }
        for (int ki = 2; ki >= 0; ki--)
            if ((k - c0) & (1 << ki)) {
                f(ki, 1 << ki, c0);
                c0 += 1 << ki;
            }
    };
    int cls_count[kBatchClasses] = {0};
    int nchunks = 0;
    for (int i = 0; i < n; i++) {
        const tc_request &r = reqs[i];
        int rc = check_spmv_args(r.A, r.X, r.ldx, r.dtype, r.Y, r.ldy, r.k);
        if (rc != TC_OK) return rc;
        if (r.k > kMaxBatchK) return tc::fail(TC_ERR_UNSUPPORTED, "batch: k=%d > %d", r.k, kMaxBatchK);
        const int dt = r.dtype == TC_F32 ? 0 : r.dtype == TC_BF16 ? 1 : 2;
        chunks(r.k, [&](int ki, int K, int c0) {
            cls_count[chunk_class(r, ki, K, c0, dt)]++;
            nchunks++;
        });
    }
    if (nchunks == 0) return TC_OK;
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
        const bool fast = v.S == 1 && v.sh0.oe - v.sh0.ob == 1;
        chunks(r.k, [&](int ki, int K, int c0) {
            const int cl = chunk_class(r, ki, K, c0, dt);
            WorkDesc &w = D[fill[cl]];
            w.A = v;
            w.X = (const char *)r.X + (size_t)c0 * elem_size(r.dtype);
            w.Y = r.Y + c0;
            w.ldx = r.ldx;
            w.ldy = r.ldy;
            w.fast = fast ? 1 : 0;
            const int trw = batch_tile_rows(cl / 12, K);
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
        if (dt == 0) launch_batch_class<float>(kern, K, dD + b, dUB + b, nd, (int)units[c], s);
        else if (dt == 1) launch_batch_class<bf16_t>(kern, K, dD + b, dUB + b, nd, (int)units[c], s);
        else launch_batch_class<f16_t>(kern, K, dD + b, dUB + b, nd, (int)units[c], s);
        HIPCHECK(hipGetLastError());
    }
    return TC_OK;
}

}  // namespace tc

extern "C" {

int tc_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int tc_matrix_create(int M, int N, const TaiChi_new *T, long long nnz, int long_threshold, void *stream,
                     tc_matrix **out)
{
    if (!T || !out || M < 0 || N < 0) return tc::fail(TC_ERR_INVALID, "tc_matrix_create: bad argument");
    *out = nullptr;
    if (M > 0 && !T->csrRow) return tc::fail(TC_ERR_INVALID, "tc_matrix_create: null csrRow");
    const int LT = long_threshold > 0 ? std::min(long_threshold, 1023) : tc::kDefaultLongRow;
    const int S = T->dia_shapes;
    const int R = M > 0 ? T->csrRow[M] : 0;
    if (S < 0 || R < 0 || (R > 0 && !T->csrCol)) return tc::fail(TC_ERR_INVALID, "tc_matrix_create: bad format");
    if (S > 0 && (!T->xStart || !T->xStop || !T->yStart || !T->nDiasPtr || (T->nDiasPtr[S] > 0 && !T->neg_offsets)))
        return tc::fail(TC_ERR_INVALID, "tc_matrix_create: null shape arrays");
    const bool encoded = T->dense_nnz > 0;
    if (!encoded && N >= (1 << 30)) return tc::fail(TC_ERR_UNSUPPORTED, "tc_matrix_create: N >= 2^30");
    const int D = S > 0 ? T->nDiasPtr[S] : 0;

    // shapes: validated so that every column the kernel forms lies in [0, N)
    std::vector<ShapeDev> shapes((size_t)S);
    for (int s = 0; s < S; s++) {
        ShapeDev &h = shapes[s];
        h.xs = T->xStart[s];
        h.xe = T->xStop[s];
        h.ys = T->yStart[s];
        h.ob = T->nDiasPtr[s];
        h.oe = T->nDiasPtr[s + 1];
        if (h.xs < 0 || h.xe > M || h.xs > h.xe || h.ys < 0 || (long long)h.ys + (h.xe - h.xs) > N || h.ob < 0 ||
            h.oe < h.ob || h.oe > D)
            return tc::fail(TC_ERR_INVALID, "tc_matrix_create: shape %d out of range", s);
        h.omin = 0;
        h.omax = 0;
        if (h.oe > h.ob) {
            h.omin = h.omax = T->neg_offsets[h.ob];
            for (int n = h.ob; n < h.oe; n++) {
                h.omin = std::min(h.omin, T->neg_offsets[n]);
                h.omax = std::max(h.omax, T->neg_offsets[n]);
            }
        }
        h.pad = 0;
    }
    // rows: lengths, long rows, group offsets
    const int G = (M + 63) / 64;
    std::vector<int> gbase((size_t)G + 1, 0), lrow, lptr;
    std::vector<uint16_t> roff16((size_t)M);
    long long short_total = 0;
    int maxoff = 0;
    for (int r = 0; r < M; r++) {
        const int b = T->csrRow[r], e = T->csrRow[r + 1];
        if (e < b) return tc::fail(TC_ERR_INVALID, "tc_matrix_create: csrRow not monotone at %d", r);
        if ((r & 63) == 0) gbase[r >> 6] = (int)short_total;
        if (e - b > LT) lrow.push_back(r);
        else short_total += e - b;
        const int o = (int)(short_total - gbase[r >> 6]);
        roff16[r] = (uint16_t)o;
        maxoff = std::max(maxoff, o);
    }
    gbase[G] = (int)short_total;
    const int L = (int)lrow.size();
    const int ow = maxoff <= 255 ? 1 : 2;
    std::vector<int> glong;
    if (L > 0) {
        glong.assign((size_t)G + 1, 0);
        for (int r : lrow) glong[(r >> 6) + 1]++;
        for (int g = 0; g < G; g++) glong[g + 1] += glong[g];
        lptr.resize((size_t)L + 1);
        lptr[0] = (int)short_total;
        for (int l = 0; l < L; l++) lptr[l + 1] = lptr[l] + (T->csrRow[lrow[l] + 1] - T->csrRow[lrow[l]]);
    }
    // encoded cols: short rows in row order, then long rows; checked against N
    std::vector<int> cols((size_t)R);
    {
        std::vector<long long> dst((size_t)M);
        long long sp = 0, lp = short_total;
        size_t li = 0;
        for (int r = 0; r < M; r++) {
            const int len = T->csrRow[r + 1] - T->csrRow[r];
            if (li < lrow.size() && lrow[li] == r) {
                dst[r] = lp;
                lp += len;
                li++;
            } else {
                dst[r] = sp;
                sp += len;
            }
        }
        int bad = 0;
        std::mutex bm;
        tc::parallel_for(M, [&](long long b, long long e) {
            int lbad = 0;
            for (long long r = b; r < e; r++) {
                long long o = dst[r];
                for (int j = T->csrRow[r]; j < T->csrRow[r + 1]; j++) {
                    const int raw = T->csrCol[j];
                    const int c = encoded ? raw >> 1 : raw;
                    if (raw < 0 || c >= N) {
                        lbad = 1;
                        cols[o++] = 1;
                        continue;
                    }
                    cols[o++] = encoded ? raw : ((raw << 1) | 1);
                }
            }
            if (lbad) {
                std::lock_guard<std::mutex> g(bm);
                bad = 1;
            }
        });
        if (bad) return tc::fail(TC_ERR_INVALID, "tc_matrix_create: column index out of range");
    }

    // one device allocation, 256-byte aligned sub-arrays
    struct Part {
        const void *src;
        size_t bytes;
        size_t off;
    };
    std::vector<uint8_t> roff8;
    if (ow == 1) {
        roff8.resize((size_t)M);
        for (int r = 0; r < M; r++) roff8[r] = (uint8_t)roff16[r];
    }
    Part parts[] = {
        {shapes.data(), shapes.size() * sizeof(ShapeDev), 0},
        {S > 0 ? T->neg_offsets : nullptr, (size_t)D * sizeof(int), 0},
        {gbase.data(), gbase.size() * sizeof(int), 0},
        {ow == 1 ? (const void *)roff8.data() : (const void *)roff16.data(), (size_t)M * ow, 0},
        {cols.data(), cols.size() * sizeof(int), 0},
        {glong.data(), glong.size() * sizeof(int), 0},
        {lrow.data(), lrow.size() * sizeof(int), 0},
        {lptr.data(), lptr.size() * sizeof(int), 0},
    };
    size_t total = 0, fmt_bytes = 0;
    for (Part &p : parts) {
        p.off = total;
        total += (p.bytes + 255) & ~(size_t)255;
        fmt_bytes += p.bytes;
    }
    tc_matrix *A = new tc_matrix();
    if (hipGetDevice(&A->device) != hipSuccess) {
        delete A;
        return tc::fail(TC_ERR_NODEVICE, "tc_matrix_create: no HIP device");
    }
    hipStream_t s = (hipStream_t)stream;
    if (total > 0) {
        hipError_t e = hipMalloc(&A->mem, total);
        if (e != hipSuccess) {
            delete A;
            return tc::fail(TC_ERR_NOMEM, "tc_matrix_create: hipMalloc(%zu): %s", total, hipGetErrorString(e));
        }
        for (Part &p : parts)
            if (p.bytes) {
                e = hipMemcpyAsync((char *)A->mem + p.off, p.src, p.bytes, hipMemcpyHostToDevice, s);
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
    auto at = [&](int i) -> const void * { return parts[i].bytes ? (const char *)A->mem + parts[i].off : nullptr; };
    A->mem_bytes = total;
    MatView &v = A->v;
    v.M = M;
    v.N = N;
    v.S = S;
    v.L = L;
    v.ow = ow;
    v.ngroups = G;
    v.shapes = (const ShapeDev *)at(0);
    v.offs = (const int *)at(1);
    if (S > 0) v.sh0 = shapes[0];
    v.gbase = (const int *)at(2);
    v.roff = at(3);
    v.cols = (const int *)at(4);
    v.glong = (const int *)at(5);
    v.lrow = (const int *)at(6);
    v.lptr = (const int *)at(7);
    tc_matrix_stats &st = A->st;
    st.M = M;
    st.N = N;
    st.dia_shapes = S;
    st.num_negs = D;
    st.long_rows = L;
    st.offset_bytes = ow;
    st.nnz = nnz;
    st.remainder = R;
    st.format_bytes = (long long)fmt_bytes;
    st.device_bytes = (long long)total;
    *out = A;
    return TC_OK;
}

int tc_matrix_footprint(int M, int N, const TaiChi_new *T, int long_threshold, long long *bytes)
{
    if (!T || !bytes || M < 0 || N < 0 || (M > 0 && !T->csrRow))
        return tc::fail(TC_ERR_INVALID, "tc_matrix_footprint: bad argument");
    const int LT = long_threshold > 0 ? std::min(long_threshold, 1023) : tc::kDefaultLongRow;
    const int S = std::max(T->dia_shapes, 0);
    const long long D = S > 0 ? T->nDiasPtr[S] : 0;
    const long long R = M > 0 ? T->csrRow[M] : 0;
    const long long G = (M + 63) / 64;
    long long L = 0, short_total = 0, gstart = 0;
    int maxoff = 0;
    for (int r = 0; r < M; r++) {
        const int len = T->csrRow[r + 1] - T->csrRow[r];
        if ((r & 63) == 0) gstart = short_total;
        if (len > LT) L++;
        else short_total += len;
        maxoff = std::max(maxoff, (int)(short_total - gstart));
    }
    const int ow = maxoff <= 255 ? 1 : 2;
    const long long sizes[] = {S * (long long)sizeof(ShapeDev), 4 * D, 4 * (G + 1), (long long)M * ow, 4 * R,
                               L > 0 ? 4 * (G + 1) : 0, 4 * L, L > 0 ? 4 * (L + 1) : 0};
    long long total = 0;
    for (long long b : sizes) total += (b + 255) & ~255ll;
    *bytes = total;
    return TC_OK;
}

int tc_matrix_destroy(tc_matrix *A)
{
    if (!A) return TC_OK;
    if (A->mem) (void)hipFree(A->mem);
    delete A;
    return TC_OK;
}

int tc_matrix_get_stats(const tc_matrix *A, tc_matrix_stats *out)
{
    if (!A || !out) return tc::fail(TC_ERR_INVALID, "tc_matrix_get_stats: null argument");
    *out = A->st;
    return TC_OK;
}

long long tc_spmv_bytes(const tc_matrix *A, int k, int dtype)
{
    if (!A || k < 1) return -1;
    return A->st.format_bytes + (long long)elem_size(dtype) * A->st.N * k + 4ll * A->st.M * k;
}

int tc_spmv(tc_matrix *A, const void *X, int ldx, int dtype, float *Y, int ldy, int k, void *stream)
{
    int rc = check_spmv_args(A, X, ldx, dtype, Y, ldy, k);
    if (rc != TC_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    const size_t es = elem_size(dtype);
    // columns in chunks of supported widths; each column's result does not
    // depend on the chunking
    for (int c0 = 0; c0 < k;) {
        const int K = pick_k(k - c0);
        const tc_config c = get_cfg(A, K, dtype);
        rc = launch_cfg(A, (const char *)X + c0 * es, ldx, dtype, Y + c0, ldy, K, c, s);
        if (rc != TC_OK) return rc;
        c0 += K;
    }
    return TC_OK;
}

int tc_matrix_set_config(tc_matrix *A, int k, int dtype, const tc_config *cfg)
{
    if (!A || !cfg || k < 1) return tc::fail(TC_ERR_INVALID, "tc_matrix_set_config: bad argument");
    const int K = pick_k(k);
    tc_config c = *cfg;
    const tc_config d = default_config(A, K);
    if (c.rows_per_tile <= 0) c.rows_per_tile = d.rows_per_tile;
    if (c.lds_values <= 0) c.lds_values = d.lds_values;
    if (!rows_per_tile_ok(K, c.rows_per_tile, c.flags))
        return tc::fail(TC_ERR_INVALID, "rows_per_tile %d not available for k=%d flags=%d", c.rows_per_tile, K, c.flags);
    if (c.lds_values < 512 || c.lds_values > 8192)
        return tc::fail(TC_ERR_INVALID, "lds_values %d outside [512, 8192]", c.lds_values);
    if (c.grid < 0) return tc::fail(TC_ERR_INVALID, "grid < 0");
    std::lock_guard<std::mutex> g(A->mu);
    A->cfg[{K, dtype}] = c;
    return TC_OK;
}

int tc_matrix_get_config(const tc_matrix *A, int k, int dtype, tc_config *cfg)
{
    if (!A || !cfg || k < 1) return tc::fail(TC_ERR_INVALID, "tc_matrix_get_config: bad argument");
    *cfg = get_cfg(const_cast<tc_matrix *>(A), pick_k(k), dtype);
    return TC_OK;
}

int tc_autotune(tc_matrix *A, const void *X, int ldx, int dtype, float *Y, int ldy, int k, void *stream,
                tc_config *best, float *best_us)
{
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
            for (int RI : {1, 2}) tf.push_back({wide_rows(K, RI), 0});
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
    tc_config bc = default_config(A, K);
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
}

int tc_spmv_batch(const tc_request *reqs, int n, void *stream)
{
    if (n < 0 || (n > 0 && !reqs)) return tc::fail(TC_ERR_INVALID, "tc_spmv_batch: bad argument");
    if (n == 0) return TC_OK;
    hipStream_t s = (hipStream_t)stream;
    std::vector<char> host(tc::batch_desc_bytes(n));
    void *dd = nullptr;
    HIPCHECK(hipMallocAsync(&dd, host.size(), s));
    int rc = tc::batch_enqueue(reqs, n, host.data(), dd, s);
    HIPCHECK(hipFreeAsync(dd, s));
    // the host descriptor array must outlive the async copy
    HIPCHECK(hipStreamSynchronize(s));
    return rc;
}

int tc_graph_capture(tc_matrix *A, const void *X, int ldx, int dtype, float *Y, int ldy, int k, int iters,
                     void *stream, tc_graph **out)
{
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
}

int tc_graph_launch(tc_graph *g, void *stream)
{
    if (!g) return tc::fail(TC_ERR_INVALID, "tc_graph_launch: null graph");
    HIPCHECK(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return TC_OK;
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
{
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
}

}  // extern "C"
