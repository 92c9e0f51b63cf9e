// Device side of the hybrid diagonal + CSR SpMV for gfx950: the device
// layout of a resident matrix and the per-tile routines every SpMV kernel
// (spmv_k1.hip, spmv_wide.hip, spmv_staged.hip, spmv_batch.hip) is built from.
//
// Replaces the device half of taichi_SpMV and CSR_coop_spmv_kernel{,_hybrid1,2,3}
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
//   pres[]      term-presence bitmaps of the shapes that have them (pm >= 0):
//               for shape s, word w (w < ceil(D_s/32)) of its local row i
//               is pres[pm + w*rows_s + i]; bit j says whether the term of
//               offset ob + j is present. A shape whose dense diagonals have
//               many zero cells gets one: the sign-0 entries that cancel those
//               cells leave cols and their bits are cleared, so the kernel
//               adds only the present terms (values stay the same: a zero
//               cell added and subtracted is a zero cell not added).
// Compared with the reference's int rowptr (4 bytes per row) the per-row
// cost is 1 (or 2) bytes plus 4 bytes per 64 rows.
//
// A workgroup (4 wavefronts) computes a tile of rows. Every kernel adds a
// (row, column) pair's terms in one canonical order fixed by the format:
//   1. diagonal terms, shape by shape, offsets in stored order; the tile's x
//      window is staged in LDS when a shape has two or more offsets, so no
//      column index is read for it;
//   2. remainder entries in stored order (streamed with coalesced loads, x
//      gathered and signed, staged in LDS, summed by one lane per row);
//   3. long rows: one wavefront each, 64-lane strided partial sums and a
//      fixed DPP/ds_swizzle butterfly.
// Every term is added in FP32 (accum) into a sum that starts at +0.0.
// The order never depends on the tile size, the grid, the column chunking or
// the batch a request runs in, so a request gives bit-identical y alone, in a
// batch, or with another configuration.
//
// No MFMA: one add per stored index, bound by HBM bandwidth.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace tcd {

constexpr int kWG = 256;
constexpr int kWaves = kWG / 64;
constexpr int kStageU = 8;  // loads in flight per lane when staging streams
constexpr int kNearRows = 64;  // row distance of a remainder entry counted as near its row
constexpr int kNearMin = 24;   // near entries a row (on average) for the wide tiles' near window

// Non-temporal stores of 1, 2 or 4 floats (y: written once, not read back by
// the kernel; streaming it past L2 leaves the cache to the x lines gathers
// reuse -- measured 5% faster on R-MAT 2^22)
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void st_nt(float *p, float a) { __builtin_nontemporal_store(a, p); }
__device__ __forceinline__ void st_nt2(float *p, float a, float b)
{
    const f32x2 v = {a, b};
    __builtin_nontemporal_store(v, (f32x2 *)p);
}
__device__ __forceinline__ void st_nt4(float *p, float a, float b, float c, float d)
{
    const f32x4 v = {a, b, c, d};
    __builtin_nontemporal_store(v, (f32x4 *)p);
}

// 1, 2 or 4 floats: non-temporal (NT) or plain stores
template <bool NT>
__device__ __forceinline__ void st1(float *p, float a)
{
    if constexpr (NT) st_nt(p, a);
    else *p = a;
}
template <bool NT>
__device__ __forceinline__ void st2(float *p, float a, float b)
{
    if constexpr (NT) st_nt2(p, a, b);
    else *(float2 *)p = make_float2(a, b);
}
template <bool NT>
__device__ __forceinline__ void st4(float *p, float a, float b, float c, float d)
{
    if constexpr (NT) st_nt4(p, a, b, c, d);
    else *(float4 *)p = make_float4(a, b, c, d);
}

// Offset (in floats) of an 8-byte-aligned array placed after capv floats of LDS
__host__ __device__ constexpr int lsum_at(int capv) { return (capv + 1) & ~1; }

// One shape: rows [xs, xe), offsets offs[ob..oe) (spread omin..omax), term
// of row r and offset o at column ys + (r - xs) + o; pm = word offset of its
// presence bitmap in pres, or -1 when every term is present; seq = 1 when
// the offsets are stored as omin, omin + 1, ..., omax (a band: offs[ob + j]
// = omin + j), else 0.
struct alignas(16) ShapeDev {
    int xs, xe, ys, ob, oe, omin, omax, pm;
    int seq, pad0, pad1, pad2;
};

// Device pointers and sizes of one resident matrix; passed by value.
struct MatView {
    int M, N, S, L;
    int ow;        // bytes per local row offset: 1 or 2
    int ngroups;
    int near;      // 1: most remainder entries lie within kNearRows of their row (bands)
    ShapeDev sh0;  // copy of shapes[0]: arrives with the kernel arguments
    const ShapeDev *shapes;
    const int *offs;
    const int *gbase;
    const void *roff;
    const int *cols;
    const int *glong;
    const int *lrow;
    const int *lptr;
    const uint32_t *pres;
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

// x[i] as loaded: all 32 bits (4-byte x) or the element's 16 bits in the low
// half (2-byte x). A gather that sits behind a branch loads bits and converts
// them (xbits_f) after the batch's last gather is issued: a conversion inside
// the branch would make each gather wait for the one before it.
template <typename XT>
__device__ __forceinline__ uint32_t load_xbits(const XT *__restrict__ X, size_t i)
{
    if constexpr (sizeof(XT) == 4) return __float_as_uint(((const float *)X)[i]);
    else return (uint32_t)((const uint16_t *)X)[i];
}
template <typename XT>
__device__ __forceinline__ float xbits_f(uint32_t b)
{
    if constexpr (sizeof(XT) == 4) return __uint_as_float(b);
    else if constexpr (__is_same(XT, bf16_t)) return __uint_as_float(b << 16);
    else return (float)__builtin_bit_cast(_Float16, (uint16_t)b);
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
    h.pm = cload(p, 7);
    h.seq = cload(p, 8);
    h.pad0 = h.pad1 = h.pad2 = 0;
    return h;
}

// Presence words of row r in shape sh (sh.pm >= 0): bit j of lo (hi) is set
// when the term of offset ob + j (ob + 32 + j) is present. Shapes with a
// bitmap have at most 64 offsets. Rows outside the shape read its first row
// (callers skip them).
__device__ __forceinline__ void pres_words(const MatView &A, const ShapeDev &sh, int r, uint32_t &lo, uint32_t &hi)
{
    const int rows = sh.xe - sh.xs;
    const int i = min(max(r - sh.xs, 0), rows - 1);
    lo = A.pres[sh.pm + i];
    hi = sh.oe - sh.ob > 32 ? A.pres[sh.pm + rows + i] : 0u;
}

// v if bit j of w is set, else +0.0, without a branch: the signed one-bit
// field is 0 or all ones. Adding +0.0 leaves a sum that started at +0.0
// unchanged (it can never be -0.0), so skipping a term and adding this zero
// give the same bits.
__device__ __forceinline__ float keep_if(float v, uint32_t w, int j)
{
    return __int_as_float(__float_as_int(v) & __builtin_amdgcn_sbfe((int)w, j, 1));
}

// LDS index idx if bit j of w is set, else idx mod 64: a run of 64 (+ the
// read's width) floats of +0.0 is kept at index 0, and reading it at idx mod
// 64 touches the same LDS banks as reading idx would, so the lanes of an
// absent term never collide with their neighbours' reads (a single shared
// zero slot measured 2-4 M bank-conflict cycles a launch on wide bands)
__device__ __forceinline__ int keep_idx(int idx, uint32_t w, int j)
{
    return idx & (__builtin_amdgcn_sbfe((int)w, j, 1) | 63);
}

// start of row i's short entries (i in [0, M]); row i ends at row_start(i+1)
__device__ __forceinline__ int row_start(const MatView &A, int i)
{
    int b = cload(A.gbase, i >> 6);
    if (i & 63) b += A.ow == 1 ? (int)((const uint8_t *)A.roff)[i - 1] : (int)((const uint16_t *)A.roff)[i - 1];
    return b;
}

// Accumulation: FP32, as the reference's kernels and CPU check sum
// (TaiChi.cu:1004-1024, main.cu:61-69). Every (row, column) pair adds its
// terms in the canonical order into a float that starts at +0.0, and y is
// that float. An FP32 sum of n terms of one sign-free kind is within
// (n-1) * 2^-24 * sum|terms| of the exact sum. What could spoil that is
// cancellation: a dense diagonal adds x over a whole segment and a sign-0
// remainder entry subtracts it again at a zero cell. The device layout
// leaves no inexact cancellation: a zero cell of a row whose only diagonal
// term is that cell cancels exactly (the term is the row's first, the
// sign-0 entry its second: x - x = +0.0), and every other zero cell is a
// cleared bit of a presence bitmap instead of a sign-0 entry (layout.hip).
// Integer-valued terms stay exact (below 2^24), so y is bit-identical to the
// reference's CPU sums for integer x.
typedef float acc_t;
__device__ __forceinline__ void accum(acc_t &s, float x) { s += x; }

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Sum over the 64 lanes of a wavefront in a fixed order: quad swaps and row
// mirrors by DPP, the 16-lane swap by ds_swizzle, halves by readlane. Every
// lane ends with the same value (each pairwise add sees the same two
// operands).
__device__ __forceinline__ float wave_sum(float v)
{
    v += dpp_f<0xB1>(v);   // quad_perm 1,0,3,2
    v += dpp_f<0x4E>(v);   // quad_perm 2,3,0,1
    v += dpp_f<0x141>(v);  // row_half_mirror
    v += dpp_f<0x140>(v);  // row_mirror
    v += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401F));  // lane ^ 16
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0)) +
           __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
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

// f(integral_constant<JB>) for JB = 0, DU, 2 DU, ... below min(nd, 64):
// a loop over groups of DU offsets unrolled at compile time
template <int DU, int JB = 0, class F>
__device__ __forceinline__ void offset_groups(int nd, F &&f)
{
    if constexpr (JB < 64) {
        if (JB >= nd) return;
        f(std::integral_constant<int, JB>{});
        offset_groups<DU, JB + DU>(nd, f);
    }
}

// acc += s[i * K] for i in [lo, hi), in order, with four LDS reads in flight
// (a lane's serial sum of its row's staged entries)
template <int K>
__device__ __forceinline__ void sum_run(acc_t &acc, const float *s, int lo, int hi)
{
    int i = lo;
    for (; i + 4 <= hi; i += 4) {
        const float a = s[i * K], b = s[(i + 1) * K], c = s[(i + 2) * K], d = s[(i + 3) * K];
        accum(acc, a);
        accum(acc, b);
        accum(acc, c);
        accum(acc, d);
    }
    for (; i < hi; i++) accum(acc, s[i * K]);
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
        uint32_t xb[U];
#pragma unroll
        for (int u = 0; u < U; u++) xb[u] = b + u * kWG < n ? load_xbits<XT>(X, (size_t)(v[u] >> 1) * ldx + cc[u]) : 0u;
        float xv[U];
#pragma unroll
        for (int u = 0; u < U; u++) xv[u] = xbits_f<XT>(xb[u]);
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int idx = b + u * kWG + tid;
            if (idx < n) s_val[idx] = (v[u] & 1) ? xv[u] : -xv[u];
        }
    }
}

// Diagonal terms of one shape for the tile's pairs, in canonical order
// (offsets as stored), added to acc. The x window of a shape with two
// or more offsets is staged in LDS once per tile (s_val, left free after).
// PM: the shape has a presence bitmap; absent terms are skipped.
template <int K, typename XT, int P, bool PM>
__device__ __forceinline__ void diag_one_shape_t(const MatView &A, const XT *__restrict__ X, int ldx, int u0, int s,
                                               const ShapeDev &sh, int ra, int rb, const int (&prow)[P],
                                               const int (&pcol)[P], const bool (&pv)[P], acc_t (&acc)[P],
                                               float *s_val, int capv)
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
    // reads behind it are independent; each pair adds its terms in offset
    // order (the canonical order).
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
    uint32_t plo[P], phi[P];
#pragma unroll
    for (int p = 0; p < P; p++) {
        plo[p] = phi[p] = 0;
        if (PM) pres_words(A, sh, u0 + prow[p], plo[p], phi[p]);
    }
    // with a bitmap, offsets run in segments of 32 (one presence word each)
    const int nseg = PM ? (sh.oe - sh.ob + 31) >> 5 : 1;
    if (stage) {
        int li[P];
#pragma unroll
        for (int p = 0; p < P; p++) li[p] = (base[p] - w0) * K + pcol[p];
        const int nd = sh.oe - sh.ob;
        if (__builtin_amdgcn_ballot_w64(!inner) == 0 && sh.seq && nd <= 64 && (w1 - w0 + 8) * K <= capv) {
            // contiguous offsets omin + j (a band): offset j of a pair reads
            // the window at its base + j K, a compile-time offset (the LDS
            // instruction's immediate), with no scalar load of offsets in
            // the loop; absent terms are masked (keep_if). Groups of DU
            // offsets, reads of a group in flight together; a group past
            // the last offset reads rows the staged window holds room for
            // (8 rows) and masks them (their bits are clear).
            constexpr int DU = P <= 2 ? 8 : 4;
            const uint32_t klo = nd >= 32 ? ~0u : (1u << nd) - 1u;
            const uint32_t khi = nd >= 64 ? ~0u : nd > 32 ? (1u << (nd - 32)) - 1u : 0u;
            uint32_t mlo[P], mhi[P];
            int lb[P];
#pragma unroll
            for (int p = 0; p < P; p++) {
                mlo[p] = (PM ? plo[p] : ~0u) & klo;
                mhi[p] = (PM ? phi[p] : ~0u) & khi;
                lb[p] = li[p] + sh.omin * K;
            }
            offset_groups<DU>(nd, [&](auto jbc) {
                constexpr int jb = decltype(jbc)::value;
                float v[DU][P];
#pragma unroll
                for (int u = 0; u < DU; u++)
#pragma unroll
                    for (int p = 0; p < P; p++) v[u][p] = s_val[lb[p] + (jb + u) * K];
#pragma unroll
                for (int u = 0; u < DU; u++)
#pragma unroll
                    for (int p = 0; p < P; p++)
                        accum(acc[p], keep_if(v[u][p], (jb + u) < 32 ? mlo[p] : mhi[p], (jb + u) & 31));
            });
        } else if (__builtin_amdgcn_ballot_w64(!inner) == 0) {
            // DU offsets per step: their scalar loads, then all DU * P LDS
            // reads, then the adds (each pair in offset order). (16 offsets
            // in flight at K = 1 measured no faster on 57-diagonal bands and
            // raised the rows kernel to 89 VGPRs.)
            constexpr int DU = 4;
            for (int seg = 0; seg < nseg; seg++) {
                const int nb = sh.ob + 32 * seg, ne = PM ? min(sh.oe, nb + 32) : sh.oe;
                uint32_t w[P];
#pragma unroll
                for (int p = 0; p < P; p++) w[p] = seg ? phi[p] : plo[p];
                int n = nb;
                for (; n + DU <= ne; n += DU) {
                    int o[DU];
#pragma unroll
                    for (int u = 0; u < DU; u++) o[u] = cload(A.offs, n + u) * K;
                    float v[DU][P];
#pragma unroll
                    for (int u = 0; u < DU; u++)
#pragma unroll
                        for (int p = 0; p < P; p++) v[u][p] = s_val[li[p] + o[u]];
#pragma unroll
                    for (int u = 0; u < DU; u++)
#pragma unroll
                        for (int p = 0; p < P; p++) accum(acc[p], PM ? keep_if(v[u][p], w[p], n - nb + u) : v[u][p]);
                }
                for (; n < ne; n++) {
                    const int o = cload(A.offs, n) * K;
#pragma unroll
                    for (int p = 0; p < P; p++) {
                        const float v = s_val[li[p] + o];
                        accum(acc[p], PM ? keep_if(v, w[p], n - nb) : v);
                    }
                }
            }
        } else {
            for (int n = sh.ob; n < sh.oe; n++) {
                const int o = cload(A.offs, n);
#pragma unroll
                for (int p = 0; p < P; p++) {
                    const int c = base[p] + o;
                    if (act[p] && c >= clo && c < chi) {
                        const float v = s_val[li[p] + o * K];
                        const int j = n - sh.ob;
                        accum(acc[p], PM ? keep_if(v, j < 32 ? plo[p] : phi[p], j & 31) : v);
                    }
                }
            }
        }
        __syncthreads();
    } else {
        for (int n = sh.ob; n < sh.oe; n++) {
            const int o = cload(A.offs, n);
#pragma unroll
            for (int p = 0; p < P; p++) {
                const int c = base[p] + o;
                if (act[p] && c >= clo && c < chi) {
                    const float v = load_x<XT>(X, (size_t)c * ldx + pcol[p]);
                    const int j = n - sh.ob;
                    accum(acc[p], PM ? keep_if(v, j < 32 ? plo[p] : phi[p], j & 31) : v);
                }
            }
        }
    }
}

template <int K, typename XT, int P>
__device__ __forceinline__ void diag_one_shape(const MatView &A, const XT *__restrict__ X, int ldx, int u0, int s,
                                               const ShapeDev &sh, int ra, int rb, const int (&prow)[P],
                                               const int (&pcol)[P], const bool (&pv)[P], acc_t (&acc)[P],
                                               float *s_val, int capv)
{
    if (sh.pm >= 0)
        diag_one_shape_t<K, XT, P, true>(A, X, ldx, u0, s, sh, ra, rb, prow, pcol, pv, acc, s_val, capv);
    else diag_one_shape_t<K, XT, P, false>(A, X, ldx, u0, s, sh, ra, rb, prow, pcol, pv, acc, s_val, capv);
}

template <int K, typename XT, int P>
__device__ __forceinline__ void diag_phase(const MatView &A, const XT *__restrict__ X, int ldx, int u0, int rows,
                                           const int (&prow)[P], const int (&pcol)[P], const bool (&pv)[P],
                                           acc_t (&acc)[P], float *s_val, int capv)
{
    for (int s = 0; s < A.S; s++) {
        const ShapeDev sh = s == 0 ? A.sh0 : load_shape(A, s);
        const int ra = max(u0, sh.xs), rb = min(u0 + rows, sh.xe);
        if (ra >= rb) continue;
        diag_one_shape<K, XT, P>(A, X, ldx, u0, s, sh, ra, rb, prow, pcol, pv, acc, s_val, capv);
    }
}

// Long rows of the tile: one wavefront each, FP32 partial sums strided over
// the 64 lanes and reduced by wave_sum_d; results land in s_lsum[local_row*K
// + c] and are added to acc. Only columns c < kc
// are read (kc < K for a masked chunk; the others stay 0). Uniform control
// flow; leaves s_lsum free.
template <int K, typename XT, int P>
__device__ __forceinline__ void long_phase(const MatView &A, const XT *__restrict__ X, int ldx, int u0, int rows,
                                           const int (&lq)[P], const bool (&pv)[P], acc_t (&acc)[P],
                                           acc_t *s_lsum, int nlsum, int kc = K)
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
        for (int c = 0; c < kc; c++) {
            acc_t part = 0.0f;
            for (int j = j0 + lane; j < j1; j += 64) {
                const int v = A.cols[j];
                const float xv = load_x<XT>(X, (size_t)(v >> 1) * ldx + c);
                accum(part, (v & 1) ? xv : -xv);
            }
            part = wave_sum(part);
            if (lane == 0) s_lsum[(row - u0) * K + c] = part;
        }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < P; p++)
        if (pv[p]) acc[p] += s_lsum[lq[p]];
    __syncthreads();
}

// Canonical per-(row, column) order, shared by every kernel below, with
// FP32 accumulation (accum) into acc and y = acc:
//   diagonal terms (shape order, offset order);
//   short-row entries in stored order, one lane, serially;
//   the long-row wavefront sum (long rows have no short entries).
template <int K, typename XT, int P>
__device__ __forceinline__ void run_tile(const MatView &A, const XT *__restrict__ X, int ldx, float *__restrict__ Y,
                                         int ldy, int u0, float *s_val, int capv, int *s_start, acc_t *s_lsum)
{
    constexpr int TRW = (kWG * P) / K;
    constexpr int NP = TRW * K;
    const int tid = threadIdx.x;
    const int rows = min(TRW, A.M - u0);

    // 1. row extents
    for (int r = tid; r <= TRW; r += kWG) s_start[r] = row_start(A, min(u0 + r, A.M));

    acc_t acc[P];
    int prow[P], pcol[P], lq[P];
    bool pv[P];
#pragma unroll
    for (int p = 0; p < P; p++) {
        const int q = tid + p * kWG;
        prow[p] = q / K;
        pcol[p] = q - prow[p] * K;
        pv[p] = q < NP && prow[p] < rows;
        lq[p] = q < NP ? q : 0;
        acc[p] = 0.0;
    }

    // 2. diagonal segments
    diag_phase<K, XT, P>(A, X, ldx, u0, rows, prow, pcol, pv, acc, s_val, capv);
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
            acc_t a = acc[p];
            sum_run<K>(a, s_val + pcol[p], lo - e0, hi - e0);
            acc[p] = a;
        }
        __syncthreads();
    }

    // 4. long rows: one wavefront per row
    long_phase<K, XT, P>(A, X, ldx, u0, rows, lq, pv, acc, s_lsum, NP);

    // 5. store
#pragma unroll
    for (int p = 0; p < P; p++)
        if (pv[p]) st_nt(&Y[(size_t)(u0 + prow[p]) * ldy + pcol[p]], (float)acc[p]);
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
// LDS: s_val[capv] | s_lsum[TRW] (LONG; 2*TRW floats reserved).
// PIPE (general path only): prefetch the next batch's column indices and sum
// rows with four LDS reads in flight; the batched kernel, whose registers
// serve every matrix kind in one launch, leaves it off.
// The default end of a rows tile: y[row] = (float)sum.
struct StoreY {
    static constexpr bool kStore = true;
    __device__ void operator()(int, acc_t) const {}
};

// EPI: what a finished row's sum becomes; StoreY writes it to Y, any
// other functor is called as epi(row, sum) instead (the iterative driver
// turns it into the next iterate there).
// more: another tile follows in this workgroup (its LDS writes then wait for
// this tile's last reads); without one the tile's last barrier is skipped.
template <typename XT, int RPT, int OW, bool FAST, bool LONG, bool PIPE = true, class EPI = StoreY>
__device__ __forceinline__ void rows_tile(const MatView &A, const XT *__restrict__ X, int ldx, float *__restrict__ Y,
                                          int ldy, int t, float *smem, int capv, const EPI &epi = EPI{},
                                          bool more = true)
{
    constexpr int TRW = kWG * RPT;
    float *s_val = smem;
    acc_t *s_lsum = smem + lsum_at(capv);
    const int tid = threadIdx.x, lane = tid & 63;
    const uint8_t *ro8 = (const uint8_t *)A.roff;
    const uint16_t *ro16 = (const uint16_t *)A.roff;
    {
        const int u0 = t * TRW;
        const int rows = min(TRW, A.M - u0);
        const int ebase = cload(A.gbase, u0 >> 6);
        const int E = cload(A.gbase, u0 + TRW >= A.M ? A.ngroups : (u0 + TRW) >> 6) - ebase;

        // Issue order: the diagonal's x (FAST), row offsets, the first batch
        // of the column stream; the adds follow in the canonical order.
        constexpr int U = 4;
        int prow[RPT], pcol[RPT], lq[RPT], endv[RPT], gbv[RPT];
        bool pv[RPT];
        acc_t acc[RPT];
#pragma unroll
        for (int j = 0; j < RPT; j++) {
            prow[j] = j * kWG + tid;
            pcol[j] = 0;
            lq[j] = prow[j];
            acc[j] = 0.0;
            pv[j] = u0 + prow[j] < A.M;
        }
        // FAST (one shape, one offset): x[w0 + i], i < TRW, holds this tile's
        // diagonal terms, one coalesced load per row. Remainder gathers that
        // fall near the diagonal find those lines in L1; staging them in LDS
        // for those gathers (a barrier and a select per gather) measured
        // slower on R-MAT 2^22 (15.8 against 15.6 us FP32, 14.4 against 13.5
        // BF16).
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
                const int r = u0 + prow[j], c = w0 + prow[j];
                if (pv[j] && r >= sh.xs && r < sh.xe && c >= sh.ys && c < sh.ys + (sh.xe - sh.xs))
                    accum(acc[j], xw[j]);
            }
        } else {
            diag_phase<1, XT, RPT>(A, X, ldx, u0, rows, prow, pcol, pv, acc, s_val, capv);
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

        // cv holds the column indices of the next batch (the first one was
        // loaded with the row offsets). Outside FAST a batch loads the indices
        // of the one after it before its own gathers, so the two are in flight
        // together and a tile of many entries pays one memory trip a batch,
        // not two.
        for (int e0 = 0; e0 < E; e0 += capv) {
            const int ne = min(capv, E - e0);
            for (int b = 0; b < ne; b += kWG * U) {
                int vv[U];
                if constexpr (FAST || !PIPE) {
                    // one shape of one offset (R-MAT-like): a tile's entries are
                    // rarely more than one batch; no prefetch
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        vv[u] = 1;
                        if (e0 == 0 && b == 0) vv[u] = cv[u];
                        else if (b + u * kWG < ne) vv[u] = A.cols[ebase + e0 + min(b + u * kWG + tid, ne - 1)];
                    }
                } else {
#pragma unroll
                    for (int u = 0; u < U; u++) vv[u] = cv[u];
                    int nb = b + kWG * U, e2 = e0, ne2 = ne;  // the next batch: later in this round, or the next round's first
                    if (nb >= ne) {
                        e2 = e0 + capv;
                        nb = 0;
                        ne2 = min(capv, E - e2);
                    }
                    if (e2 < E) {
#pragma unroll
                        for (int u = 0; u < U; u++)
                            cv[u] = nb + u * kWG < ne2 ? A.cols[ebase + e2 + min(nb + u * kWG + tid, ne2 - 1)] : 1;
                    }
                }
                float xv[U];
                {
                    // 2-byte x: bits behind the branch, converted after every
                    // gather is issued (a conversion inside it serialises them)
                    uint32_t xb[U];
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        xb[u] = 0u;
                        if (b + u * kWG < ne) xb[u] = load_xbits<XT>(X, (size_t)(vv[u] >> 1) * ldx);
                    }
#pragma unroll
                    for (int u = 0; u < U; u++) xv[u] = xbits_f<XT>(xb[u]);
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
                acc_t a = acc[j];
                if constexpr (FAST || !PIPE) {  // rows of a few entries: the plain loop
                    for (int k = lo; k < hi; k++) accum(a, s_val[k - e0]);
                } else {
                    sum_run<1>(a, s_val, lo - e0, hi - e0);
                }
                acc[j] = a;
            }
            // long_phase writes s_lsum only, apart from s_val
            if (more || e0 + capv < E) __syncthreads();
        }

        if constexpr (LONG) long_phase<1, XT, RPT>(A, X, ldx, u0, rows, lq, pv, acc, s_lsum, TRW);

#pragma unroll
        for (int j = 0; j < RPT; j++) {
            if (!pv[j]) continue;
            if constexpr (EPI::kStore) st_nt(&Y[(size_t)(u0 + prow[j]) * ldy], (float)acc[j]);
            else epi(u0 + prow[j], acc[j]);
        }
    }
}

// Q consecutive floats of LDS (Q = 2, 4 or 8: one or two vector accesses)
template <int Q>
__device__ __forceinline__ void lds_get(const float *s, float (&o)[Q])
{
    if constexpr (Q == 2) {
        const float2 v = *(const float2 *)s;
        o[0] = v.x, o[1] = v.y;
    } else {
#pragma unroll
        for (int h = 0; h < Q; h += 4) {
            const float4 v = *(const float4 *)(s + h);
            o[h] = v.x, o[h + 1] = v.y, o[h + 2] = v.z, o[h + 3] = v.w;
        }
    }
}
template <int Q>
__device__ __forceinline__ void lds_put(float *d, const float (&v)[Q])
{
    if constexpr (Q == 2) {
        *(float2 *)d = make_float2(v[0], v[1]);
    } else {
#pragma unroll
        for (int h = 0; h < Q; h += 4) *(float4 *)(d + h) = make_float4(v[h], v[h + 1], v[h + 2], v[h + 3]);
    }
}

// LDS byte address of a shared-memory pointer, and Q floats (Q = 2, 4, 8) read
// at an LDS byte address (constant parts of the address become the
// instruction's immediate offset)
typedef __attribute__((address_space(3))) float lds_float;
__device__ __forceinline__ uint32_t lds_addr(const float *p) { return (uint32_t)(uintptr_t)(const lds_float *)p; }
template <int Q>
__device__ __forceinline__ void lds_get_at(uint32_t a, float (&o)[Q])
{
    static_assert(Q == 2 || Q == 4 || Q == 8, "2, 4 or 8 floats");
    if constexpr (Q == 2) {
        const f32x2 v = *(const __attribute__((address_space(3))) f32x2 *)a;
        o[0] = v[0], o[1] = v[1];
    } else {
#pragma unroll
        for (int h = 0; h < Q; h += 4) {
            const f32x4 v = *(const __attribute__((address_space(3))) f32x4 *)(a + 4 * h);
            o[h] = v[0], o[h + 1] = v[1], o[h + 2] = v[2], o[h + 3] = v[3];
        }
    }
}

// x row segment of Q elements as floats: one vector load (Q * sizeof(XT)
// bytes: 8 or 16 for 4-byte x, 4, 8 or 16 for 2-byte x)
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
        if constexpr (Q == 8) {
            const uint4 v = *(const uint4 *)p;
            w[0] = v.x, w[1] = v.y, w[2] = v.z, w[3] = v.w;
        } else if constexpr (Q == 4) {
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

// A raw x row segment of Q elements as loaded: Q words for 4-byte x, Q/2
// words (two elements each) for 2-byte x. Gathers fill segments and the adds
// convert them (seg_at), so no conversion sits between one gather and the
// next and a lane's gathers stay in flight together.
template <typename XT, int Q>
struct XSeg {
    static constexpr int NW = sizeof(XT) == 4 ? Q : Q / 2;
    uint32_t w[NW];
};

// NW consecutive words (global or LDS): one 4-, 8- or 16-byte access, or
// two 16-byte ones
template <int NW>
__device__ __forceinline__ void ld_words(const void *p, uint32_t (&w)[NW])
{
    static_assert(NW == 1 || NW == 2 || NW == 4 || NW == 8, "x segments are 4, 8, 16 or 32 bytes");
    if constexpr (NW == 8) {
        const uint4 a = ((const uint4 *)p)[0], b = ((const uint4 *)p)[1];
        w[0] = a.x, w[1] = a.y, w[2] = a.z, w[3] = a.w, w[4] = b.x, w[5] = b.y, w[6] = b.z, w[7] = b.w;
    } else if constexpr (NW == 4) {
        const uint4 v = *(const uint4 *)p;
        w[0] = v.x, w[1] = v.y, w[2] = v.z, w[3] = v.w;
    } else if constexpr (NW == 2) {
        const uint2 v = *(const uint2 *)p;
        w[0] = v.x, w[1] = v.y;
    } else {
        w[0] = *(const uint32_t *)p;
    }
}
template <int NW>
__device__ __forceinline__ void st_words(void *p, const uint32_t (&w)[NW])
{
    static_assert(NW == 1 || NW == 2 || NW == 4 || NW == 8, "x segments are 4, 8, 16 or 32 bytes");
    if constexpr (NW == 8) {
        ((uint4 *)p)[0] = make_uint4(w[0], w[1], w[2], w[3]);
        ((uint4 *)p)[1] = make_uint4(w[4], w[5], w[6], w[7]);
    } else if constexpr (NW == 4) *(uint4 *)p = make_uint4(w[0], w[1], w[2], w[3]);
    else if constexpr (NW == 2) *(uint2 *)p = make_uint2(w[0], w[1]);
    else *(uint32_t *)p = w[0];
}

// element m (a constant after unrolling) of a segment, as a float
template <typename XT, int Q>
__device__ __forceinline__ float seg_at(const XSeg<XT, Q> &s, int m)
{
    if constexpr (sizeof(XT) == 4) return __uint_as_float(s.w[m]);
    else if constexpr (__is_same(XT, bf16_t))
        return __uint_as_float((m & 1) ? (s.w[m >> 1] & 0xffff0000u) : (s.w[m >> 1] << 16));
    else return (float)__builtin_bit_cast(_Float16, (uint16_t)(s.w[m >> 1] >> (16 * (m & 1))));
}

template <typename XT, int Q>
__device__ __forceinline__ void load_seg(const XT *__restrict__ p, XSeg<XT, Q> &s)
{
    ld_words(p, s.w);
}

// Segment of a masked chunk. The whole vector is loaded: it lies inside the
// row's ldx span (ldx is a multiple of Q and the lane's first column is below
// kv <= ldx), except in the last row of X, which may end at column kv --
// there (the returned tail: last, nv < Q) the vector is read from the row
// above, the same columns one ldx earlier (masked chunks need N >= 2), and
// seg_m_tail later puts the first nv elements in, read one by one (the
// others stay 0). Elsewhere the elements at or past nv are whatever X holds
// there: they feed only columns that are never stored. The vector load is
// unconditional (an address select between two offsets of p: no pad array
// whose address the loop would rebuild). Callers issue a whole batch of loads
// before the rare tail fix-ups: a fix-up writes the segment's registers, so
// it waits for the loads, and one inside the batch would serialise them.
template <typename XT, int Q>
__device__ __forceinline__ bool load_seg_m(const XT *__restrict__ X, int row, int ldx, int c0, int nv, int N,
                                           XSeg<XT, Q> &s)
{
    const bool tail = row == N - 1 && nv < Q;
    ld_words(X + (size_t)(row - (int)tail) * ldx + c0, s.w);
    return tail;
}
template <typename XT, int Q>
__device__ __forceinline__ void seg_m_tail(const XT *__restrict__ p, int nv, XSeg<XT, Q> &s)
{
    uint32_t w[XSeg<XT, Q>::NW];
#pragma unroll
    for (int i = 0; i < XSeg<XT, Q>::NW; i++) w[i] = 0u;
#pragma unroll
    for (int m = 0; m < Q; m++)
        if (m < nv) {
            if constexpr (sizeof(XT) == 4) w[m] = ((const uint32_t *)p)[m];
            else w[m >> 1] |= (uint32_t)((const uint16_t *)p)[m] << (16 * (m & 1));
        }
#pragma unroll
    for (int i = 0; i < XSeg<XT, Q>::NW; i++) s.w[i] = w[i];
}

// The first nv (0 < nv < Q) of Q values to yp (Q-aligned) in aligned pieces
// of 4, 2 and 1: at most three stores, nothing past column nv. Indices are
// constants in every branch, so f stays in registers.
template <int Q, bool NT>
__device__ __forceinline__ void store_first(float *yp, const float (&f)[Q], int nv)
{
    if constexpr (Q >= 8) {
        if (nv >= 4) {
            st4<NT>(yp, f[0], f[1], f[2], f[3]);
            if (nv == 5) st1<NT>(yp + 4, f[4]);
            else if (nv >= 6) st2<NT>(yp + 4, f[4], f[5]);
            if (nv == 7) st1<NT>(yp + 6, f[6]);
            return;
        }
    }
    if constexpr (Q >= 4) {
        if (nv >= 2) {
            st2<NT>(yp, f[0], f[1]);
            if (nv == 3) st1<NT>(yp + 2, f[2]);
            return;
        }
    }
    st1<NT>(yp, f[0]);
}

// Columns a lane owns in the wide-rows tiles: 4 (a 16-byte FP32 vector, an
// 8-byte BF16/FP16 one), or 8 for 2-byte x when K >= 8 (a 16-byte vector: half
// the lanes per row, half the vector memory instructions) and for FP32 x at
// K = 8 (two 16-byte vectors: one lane a row, so the per-offset address and
// presence mask of a dense diagonal are paid once a row, not twice; the
// diagonal loop of wide bands is bound by VALU issue)
__host__ __device__ constexpr int wide_cpl(int K, int esize)
{
    return (esize == 2 && K >= 8) || K == 8 ? 8 : (K < 4 ? K : 4);
}

// ---- k > 1: wide rows ----
// A row belongs to a group of LPR = K/CPL lanes; lane q of the group owns the
// CPL = wide_cpl(K, sizeof(XT)) consecutive columns q*CPL.., so each x row segment the row
// touches (a diagonal term or a remainder entry) is one CPL-wide vector load
// per lane and y leaves as one vector store per lane. Consecutive groups own
// consecutive rows (y is written contiguously when ldy = K). Rows are
// independent: no LDS staging; only long rows (one wavefront each, the fixed
// butterfly) meet in LDS. Same canonical per-(row, column) order as every
// other kernel. Requires X, Y, ldx and ldy aligned to CPL elements.
// MASK: only the first kv < K columns exist; a lane computes all its
// columns but stores only those below kv (load_seg_m keeps every load in
// bounds). A masked chunk runs k = 3, 5, 6, 7 as one pass of width 4 or 8
// instead of 2 + 1, 4 + 1, 4 + 2, 4 + 2 + 1 passes. Needs kv > K/2 (so a
// lane's vector stays inside its X row, or X's last row is handled apart)
// and N >= 2 (X's last row borrows the row above for its vector loads).
// LDS: a shape's x window (a run of kWideZero floats of +0.0, then up to
// TRW + kWideHalo rows of K columns; for K = 8 two planes of 4 columns,
// wide_plane floats apart), reused for the rows around the tile that a near
// remainder reads ((TRW + 2 kWideHalo) * K floats, K <= 8) and for the
// long-row sums (TRW * K floats); the launch sizes LDS by wide_lds_bytes.
constexpr int kWideHalo = 64;  // window rows beyond the tile's that LDS holds (offset spread)
// floats of +0.0 before a shape's x window (before each plane): absent terms
// read them at an index congruent mod 64 to the present term's, and the
// contiguous-offset loop reads them at that index + j * 4 for offset j < 64
constexpr int kWideZero = 320;

// distance in floats between the two planes of a K = 8 window (tiles of trw rows)
__host__ __device__ constexpr int wide_plane(int trw) { return (kWideZero + (trw + kWideHalo) * 4 + 63) & ~63; }

// LDS bytes of a wide-rows tile of trw rows and K columns: the shape window
// (or, near, the rows around the tile the remainder reads) and the long-row
// sums (lng)
__host__ __device__ constexpr size_t wide_lds_bytes(int K, int trw, bool near, bool lng)
{
    const size_t shape = K == 8 ? (size_t)2 * wide_plane(trw) * 4 : ((size_t)(trw + kWideHalo) * K + kWideZero) * 4;
    const size_t nearw = near && K <= 8 ? (size_t)(trw + 2 * kWideHalo) * K * 4 : 0;
    const size_t win = shape > nearw ? shape : nearw;
    const size_t lsum = lng ? (size_t)trw * K * sizeof(acc_t) : 0;
    return win > lsum ? win : lsum;
}

// Stage x rows [w0, w0 + nrows), all K columns, in LDS as floats smem[(row -
// w0) * K + col], CPL-wide vectors, four loads in flight per lane. Lanes past
// the window repeat its last vector (same address, same value), so neither
// the loads nor the LDS stores sit behind a branch.
template <int K, typename XT, int CPL, bool MASK, bool PLANAR = false>
__device__ __forceinline__ void stage_x_rows(const XT *__restrict__ X, int ldx, int N, int w0, int nrows, int kv,
                                             float *smem, int ps = 0)
{
    constexpr int LPR = K / CPL;
    const int nw = nrows * LPR;  // CPL-wide vectors of the window
    const int tid = threadIdx.x;
    for (int b = 0; b < nw; b += kWG * 4) {
        XSeg<XT, CPL> tv[4];
        bool tl[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int v = min(b + u * kWG + tid, nw - 1);
            const XT *xp = X + (size_t)(w0 + v / LPR) * ldx + (v % LPR) * CPL;
            tl[u] = false;
            if constexpr (MASK)
                tl[u] = load_seg_m<XT, CPL>(X, w0 + v / LPR, ldx, (v % LPR) * CPL, min(max(kv - (v % LPR) * CPL, 0), CPL), N, tv[u]);
            else load_seg<XT, CPL>(xp, tv[u]);
        }
        if constexpr (MASK) {
            // one wave-uniform test skips the fix-ups when no lane has a tail
            if (__builtin_amdgcn_ballot_w64(tl[0] || tl[1] || tl[2] || tl[3]) != 0) {
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (tl[u]) {
                        const int v = min(b + u * kWG + tid, nw - 1);
                        seg_m_tail<XT, CPL>(X + (size_t)(w0 + v / LPR) * ldx + (v % LPR) * CPL,
                                            min(max(kv - (v % LPR) * CPL, 0), CPL), tv[u]);
                    }
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int v = min(b + u * kWG + tid, nw - 1);
            float f[CPL];
#pragma unroll
            for (int m = 0; m < CPL; m++) f[m] = seg_at(tv[u], m);
            if constexpr (PLANAR) {
                float lo[4], hi[4];
#pragma unroll
                for (int m = 0; m < 4; m++) lo[m] = f[m], hi[m] = f[4 + m];
                lds_put<4>(smem + v * 4, lo);
                lds_put<4>(smem + ps + v * 4, hi);
            } else {
                lds_put<CPL>(smem + (v / LPR) * K + (v % LPR) * CPL, f);
            }
        }
    }
}

// Stage x rows [w0, w0 + nrows), all K columns, in LDS as raw segments (the
// near window): row i at word i * K * sizeof(XT) / 4, the lane's columns c0..
// at word c0 * sizeof(XT) / 4 of it -- half the bytes for 2-byte x, and a read
// gives the same segment a gather would
template <int K, typename XT, int CPL, bool MASK>
__device__ __forceinline__ void stage_seg_rows(const XT *__restrict__ X, int ldx, int N, int w0, int nrows, int kv,
                                               uint32_t *smw)
{
    constexpr int LPR = K / CPL;
    constexpr int RW = K * (int)sizeof(XT) / 4, SW = CPL * (int)sizeof(XT) / 4;
    const int nw = nrows * LPR;
    const int tid = threadIdx.x;
    for (int b = 0; b < nw; b += kWG * 4) {
        XSeg<XT, CPL> tv[4];
        bool tl[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int v = min(b + u * kWG + tid, nw - 1);
            const XT *xp = X + (size_t)(w0 + v / LPR) * ldx + (v % LPR) * CPL;
            tl[u] = false;
            if constexpr (MASK)
                tl[u] = load_seg_m<XT, CPL>(X, w0 + v / LPR, ldx, (v % LPR) * CPL, min(max(kv - (v % LPR) * CPL, 0), CPL), N, tv[u]);
            else load_seg<XT, CPL>(xp, tv[u]);
        }
        if constexpr (MASK) {
            // one wave-uniform test skips the fix-ups when no lane has a tail
            if (__builtin_amdgcn_ballot_w64(tl[0] || tl[1] || tl[2] || tl[3]) != 0) {
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (tl[u]) {
                        const int v = min(b + u * kWG + tid, nw - 1);
                        seg_m_tail<XT, CPL>(X + (size_t)(w0 + v / LPR) * ldx + (v % LPR) * CPL,
                                            min(max(kv - (v % LPR) * CPL, 0), CPL), tv[u]);
                    }
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int v = min(b + u * kWG + tid, nw - 1);
            st_words(smw + (v / LPR) * RW + (v % LPR) * SW, tv[u].w);
        }
    }
}

template <int K, typename XT, int RI, bool LONG, bool MASK = false, bool NEAR = false>
__device__ __forceinline__ void wide_tile(const MatView &A, const XT *__restrict__ X, int ldx, float *__restrict__ Y,
                                          int ldy, int t, float *smem, int kv = K)
{
    static_assert(!NEAR || K <= 8, "the near window is for K <= 8");
    constexpr int CPL = wide_cpl(K, (int)sizeof(XT));
    constexpr int LPR = K / CPL;
    constexpr int GPW = kWG / LPR;  // row groups per workgroup
    constexpr int TRW = GPW * RI;   // rows per tile
    constexpr int P = CPL * RI;     // (row, column) pairs per lane
    constexpr int U = 4;            // remainder entries in flight per lane
    // a shape's LDS window: two planes of 4 columns when one lane owns a
    // row's 8 columns, else rows of K columns; RS floats a window row
    constexpr bool PLANAR = CPL == 8 && LPR == 1;
    constexpr int RS = PLANAR ? 4 : K;
    const int tid = threadIdx.x;
    const int g = tid / LPR, c0 = (tid % LPR) * CPL;
    const int nv = MASK ? min(max(kv - c0, 0), CPL) : CPL;  // this lane's columns that exist
    const uint8_t *ro8 = (const uint8_t *)A.roff;
    const uint16_t *ro16 = (const uint16_t *)A.roff;
    {
        const int u0 = t * TRW;
        acc_t acc[P];
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
                acc[i * CPL + m] = 0.0;
                lq[i * CPL + m] = (GPW * i + g) * K + c0 + m;
                pv[i * CPL + m] = valid;
            }
        }

        // diagonal terms: shapes in order, offsets in order; the rows' loads
        // of one offset are issued together. A shape with two or more
        // offsets first stages the tile's x window (rows [w0, w1), K columns)
        // in LDS, so each x row segment is loaded once per tile, not once
        // per offset. Shapes with a presence bitmap (PM) skip absent terms.
        for (int s = 0; s < A.S; s++) {
            const ShapeDev sh = s == 0 ? A.sh0 : load_shape(A, s);
            if (u0 + TRW <= sh.xs || u0 >= sh.xe) continue;
            auto shape_terms = [&](auto pmc) {
                constexpr bool PM = decltype(pmc)::value;
                const int side = sh.xe - sh.xs;
                const int shift = sh.ys - sh.xs;
                const int ra = max(u0, sh.xs), rb = min(min(u0 + TRW, sh.xe), A.M);
                const int w0 = max(sh.ys, ra + shift + sh.omin);
                const int w1 = min(sh.ys + side, rb - 1 + shift + sh.omax + 1);
                uint32_t plo[RI], phi[RI];
#pragma unroll
                for (int i = 0; i < RI; i++) {
                    plo[i] = phi[i] = 0;
                    if (PM) pres_words(A, sh, u0 + GPW * i + g, plo[i], phi[i]);
                }
                if (sh.oe - sh.ob >= 2 && w1 > w0 && w1 - w0 <= TRW + kWideHalo) {
                    // the window starts past kWideZero floats of +0.0: a term
                    // a bitmap marks absent reads them (keep_idx), so one
                    // mask per row and offset replaces one per column. With 8
                    // columns a lane (one lane a row) the window is two
                    // planes of 4 columns, ps floats apart, each with its
                    // zero run: every read of a wavefront is then 16
                    // contiguous bytes a lane (32-byte rows measured heavy
                    // bank conflicts)
                    constexpr int ps = PLANAR ? wide_plane(TRW) : 0;
                    stage_x_rows<K, XT, CPL, MASK, PLANAR>(X, ldx, A.N, w0, w1 - w0, kv, smem + kWideZero, ps);
                    for (int q = tid; q < kWideZero; q += kWG) {
                        smem[q] = 0.0f;
                        if (PLANAR) smem[ps + q] = 0.0f;
                    }
                    __syncthreads();
                    // the same at an LDS byte address
                    auto win_at = [&](uint32_t a, float (&q)[CPL]) {
                        if constexpr (PLANAR) {
                            float lo[4], hi[4];
                            lds_get_at<4>(a, lo);
                            lds_get_at<4>(a + ps * 4, hi);
#pragma unroll
                            for (int m = 0; m < 4; m++) q[m] = lo[m], q[4 + m] = hi[m];
                        } else {
                            lds_get_at<CPL>(a, q);
                        }
                    };
                    auto win_get = [&](int idx, float (&q)[CPL]) {
                        if constexpr (PLANAR) {
                            float lo[4], hi[4];
                            lds_get<4>(smem + idx, lo);
                            lds_get<4>(smem + ps + idx, hi);
#pragma unroll
                            for (int m = 0; m < 4; m++) q[m] = lo[m], q[4 + m] = hi[m];
                        } else {
                            lds_get<CPL>(smem + idx, q);
                        }
                    };
                    // every row of the wavefront inside the shape with all its
                    // columns: DU offsets per step, their scalar loads, then
                    // the DU * RI LDS vectors, then the adds in offset order
                    bool inner = true;
#pragma unroll
                    for (int i = 0; i < RI; i++) {
                        const int r = u0 + GPW * i + g;
                        inner = inner && r < A.M && r >= sh.xs && r < sh.xe && r + shift + sh.omin >= sh.ys &&
                                r + shift + sh.omax < sh.ys + side;
                    }
                    const int nd = sh.oe - sh.ob;
                    const bool contig = RS <= 4 && nd <= 64 && sh.seq;
                    if (__builtin_amdgcn_ballot_w64(!inner) == 0 && contig) {
                        // contiguous offsets omin + j (a band): offset j of a
                        // row reads its x at b0 + j * RS, or, absent, at
                        // (b0 mod 64) + j * RS inside the zero run (same
                        // banks). One bit extract and one multiply-add pick
                        // the base; j * RS is the LDS instruction's immediate
                        // offset. Offsets in order, DU reads in flight. The
                        // pick is one bit-select (v_bitop3) on the
                        // sign-extended presence bit.
                        uint32_t mlo[RI], mhi[RI];
                        uint32_t ap[RI], az[RI];  // LDS byte addresses: present base, zero-run base
                        const uint32_t klo = nd >= 32 ? ~0u : (1u << nd) - 1u;
                        const uint32_t khi = nd >= 64 ? ~0u : nd > 32 ? (1u << (nd - 32)) - 1u : 0u;
#pragma unroll
                        for (int i = 0; i < RI; i++) {
                            mlo[i] = (PM ? plo[i] : ~0u) & klo;
                            mhi[i] = (PM ? phi[i] : ~0u) & khi;
                            // in bytes: the bank period is 256 bytes
                            const int b0 = 4 * (kWideZero + (u0 + GPW * i + g + shift + sh.omin - w0) * RS +
                                                (PLANAR ? 0 : c0));
                            ap[i] = lds_addr(smem) + (uint32_t)b0;
                            az[i] = lds_addr(smem) + (uint32_t)(b0 & 255);
                        }
                        constexpr int DU = RI * CPL <= 4 ? 8 : 4;
                        // groups of DU offsets, unrolled at compile time (j
                        // a constant: immediate bit positions and offsets)
                        offset_groups<DU>(nd, [&](auto jbc) {
                            constexpr int jb = decltype(jbc)::value;
                            float q[DU][RI][CPL];
#pragma unroll
                            for (int u = 0; u < DU; u++)
#pragma unroll
                                for (int i = 0; i < RI; i++) {
                                    const int j = jb + u;
                                    const uint32_t m =
                                        (uint32_t)__builtin_amdgcn_sbfe((int)(j < 32 ? mlo[i] : mhi[i]), j & 31, 1);
                                    // m ? ap : az (truth table of s1 ? s0 : s2)
                                    const uint32_t a = __builtin_amdgcn_bitop3_b32(ap[i], m, az[i], 0xE2);
                                    win_at(a + j * RS * 4, q[u][i]);
                                }
#pragma unroll
                            for (int u = 0; u < DU; u++)
#pragma unroll
                                for (int i = 0; i < RI; i++)
#pragma unroll
                                    for (int m = 0; m < CPL; m++) accum(acc[i * CPL + m], q[u][i][m]);
                        });
                    } else if (__builtin_amdgcn_ballot_w64(!inner) == 0) {
                        constexpr int DU = 4;
                        const int nseg = PM ? (sh.oe - sh.ob + 31) >> 5 : 1;
                        int base[RI];
#pragma unroll
                        for (int i = 0; i < RI; i++)
                            base[i] = kWideZero + (u0 + GPW * i + g + shift - w0) * RS + (PLANAR ? 0 : c0);
                        for (int seg = 0; seg < nseg; seg++) {
                            const int nb = sh.ob + 32 * seg, ne = PM ? min(sh.oe, nb + 32) : sh.oe;
                            uint32_t w[RI];
#pragma unroll
                            for (int i = 0; i < RI; i++) w[i] = seg ? phi[i] : plo[i];
                            // whole groups of DU offsets with no exits: one
                            // scalar load brings the group's offsets (an exit
                            // per offset makes the compiler wait on a scalar
                            // load for each); each offset's LDS reads and adds
                            // then stay together (sched_barrier), so only
                            // RI * CPL x values are live and occupancy holds.
                            // The rest run one offset at a time.
                            int n = nb;
                            // up to 8 columns with at most 4 x values a row
                            // per lane: the group's DU * RI LDS vectors are
                            // issued together, then the adds in offset order
                            // (16 registers in flight, occupancy holds)
                            constexpr bool kAll = K <= 8 && RI * CPL <= 4;
                            if constexpr (kAll) {
                                for (; n + DU <= ne; n += DU) {
                                    int o[DU];
#pragma unroll
                                    for (int u = 0; u < DU; u++) o[u] = cload(A.offs, n + u) * RS;
                                    float q[DU][RI][CPL];
#pragma unroll
                                    for (int u = 0; u < DU; u++)
#pragma unroll
                                        for (int i = 0; i < RI; i++)
                                            win_get(PM ? keep_idx(base[i] + o[u], w[i], n - nb + u) : base[i] + o[u],
                                                    q[u][i]);
#pragma unroll
                                    for (int u = 0; u < DU; u++)
#pragma unroll
                                        for (int i = 0; i < RI; i++)
#pragma unroll
                                            for (int m = 0; m < CPL; m++) accum(acc[i * CPL + m], q[u][i][m]);
                                }
                            }
                            for (; n + DU <= ne; n += DU) {
                                int o[DU];
#pragma unroll
                                for (int u = 0; u < DU; u++) o[u] = cload(A.offs, n + u) * RS;
#pragma unroll
                                for (int u = 0; u < DU; u++) {
                                    float q[RI][CPL];
#pragma unroll
                                    for (int i = 0; i < RI; i++)
                                        win_get(PM ? keep_idx(base[i] + o[u], w[i], n - nb + u) : base[i] + o[u], q[i]);
#pragma unroll
                                    for (int i = 0; i < RI; i++)
#pragma unroll
                                        for (int m = 0; m < CPL; m++) accum(acc[i * CPL + m], q[i][m]);
                                    __builtin_amdgcn_sched_barrier(0);
                                }
                            }
                            for (; n < ne; n++) {
                                const int o = cload(A.offs, n) * RS;
                                float q[RI][CPL];
#pragma unroll
                                for (int i = 0; i < RI; i++)
                                    win_get(PM ? keep_idx(base[i] + o, w[i], n - nb) : base[i] + o, q[i]);
#pragma unroll
                                for (int i = 0; i < RI; i++)
#pragma unroll
                                    for (int m = 0; m < CPL; m++) accum(acc[i * CPL + m], q[i][m]);
                            }
                        }
                    } else {
                        // some row of this wavefront is off the square or
                        // its columns: every term is tested
                        for (int n = sh.ob; n < sh.oe; n++) {
                            const int o = cload(A.offs, n);
#pragma unroll
                            for (int i = 0; i < RI; i++) {
                                const int r = u0 + GPW * i + g;
                                const int c = r + shift + o;
                                if (r < A.M && r >= sh.xs && r < sh.xe && c >= sh.ys && c < sh.ys + side) {
                                    const int j = n - sh.ob;
                                    const uint32_t w = j < 32 ? plo[i] : phi[i];
                                    float q[CPL];
                                    win_get(kWideZero + (c - w0) * RS + (PLANAR ? 0 : c0), q);
#pragma unroll
                                    for (int m = 0; m < CPL; m++)
                                        accum(acc[i * CPL + m], PM ? keep_if(q[m], w, j & 31) : q[m]);
                                }
                            }
                        }
                    }
                    // one barrier for both branches (the window is reused)
                    __syncthreads();
                    return;
                }
                for (int n = sh.ob; n < sh.oe; n++) {
                    const int o = cload(A.offs, n);
                    XSeg<XT, CPL> xs[RI];
                    bool in[RI], tl[RI];
#pragma unroll
                    for (int i = 0; i < RI; i++) {
                        const int r = u0 + GPW * i + g;
                        const int c = r - sh.xs + sh.ys + o;
                        in[i] = r < A.M && r >= sh.xs && r < sh.xe && c >= sh.ys && c < sh.ys + side;
                        tl[i] = false;
                        if (in[i]) {
                            if constexpr (MASK)
                                tl[i] = load_seg_m<XT, CPL>(X, c, ldx, c0, nv, A.N, xs[i]);
                            else load_seg<XT, CPL>(X + (size_t)c * ldx + c0, xs[i]);
                        }
                    }
                    if constexpr (MASK) {
#pragma unroll
                        for (int i = 0; i < RI; i++)
                            if (tl[i]) seg_m_tail<XT, CPL>(X + (size_t)(A.N - 1) * ldx + c0, nv, xs[i]);
                    }
#pragma unroll
                    for (int i = 0; i < RI; i++)
                        if (in[i]) {
                            const int j = n - sh.ob;
                            const uint32_t w = j < 32 ? plo[i] : phi[i];
#pragma unroll
                            for (int m = 0; m < CPL; m++) {
                                const float xv = seg_at(xs[i], m);
                                accum(acc[i * CPL + m], PM ? keep_if(xv, w, j & 31) : xv);
                            }
                        }
                }
            };
            if (sh.pm >= 0) shape_terms(std::true_type{});
            else shape_terms(std::false_type{});
        }

        // NEAR (the matrix's A.near, K <= 8): its remainder entries mostly
        // lie near their rows, so x rows [u0 - kWideHalo, u0 + TRW +
        // kWideHalo) are staged in LDS once and an entry inside them reads x
        // there, not through the texture path
        int wl = 0, wn = 0;
        constexpr int RW = K * (int)sizeof(XT) / 4;  // words of a near-window row
        const int c0w = c0 * (int)sizeof(XT) / 4;    // the lane's first word in it
        uint32_t *smw = (uint32_t *)smem;
        if constexpr (NEAR) {
            wl = max(0, u0 - kWideHalo);
            wn = max(0, min(A.N, u0 + TRW + kWideHalo) - wl);
            stage_seg_rows<K, XT, CPL, MASK>(X, ldx, A.N, wl, wn, kv, smw);
            __syncthreads();
        }

        // remainder entries, U in flight, added in stored order; the next U
        // column indices load before this step's gathers, so a long row pays
        // one memory trip per step, not two
#pragma unroll
        for (int i = 0; i < RI; i++) {
            int nx[U];
#pragma unroll
            for (int u = 0; u < U; u++) nx[u] = rs[i] + u < re[i] ? A.cols[rs[i] + u] : 1;
            for (int j0 = rs[i]; j0 < re[i]; j0 += U) {
                int v[U];
                XSeg<XT, CPL> xs[U];
#pragma unroll
                for (int u = 0; u < U; u++) v[u] = nx[u];
                if (j0 + U < re[i])
#pragma unroll
                    for (int u = 0; u < U; u++) nx[u] = j0 + U + u < re[i] ? A.cols[j0 + U + u] : 1;
                bool tl[U], nr[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    tl[u] = false;
                    nr[u] = NEAR && (unsigned)((v[u] >> 1) - wl) < (unsigned)wn;
                    if (j0 + u < re[i] && !nr[u]) {
                        const int c = v[u] >> 1;
                        if constexpr (MASK)
                            tl[u] = load_seg_m<XT, CPL>(X, c, ldx, c0, nv, A.N, xs[u]);
                        else load_seg<XT, CPL>(X + (size_t)c * ldx + c0, xs[u]);
                    }
                }
                // near entries read the staged rows after every gather is
                // issued, into registers of their own (a read into the
                // gather's register would wait for that gather)
                if constexpr (NEAR) {
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        XSeg<XT, CPL> xn;
                        ld_words(smw + (nr[u] ? (v[u] >> 1) - wl : 0) * RW + c0w, xn.w);
#pragma unroll
                        for (int w = 0; w < XSeg<XT, CPL>::NW; w++) xs[u].w[w] = nr[u] ? xn.w[w] : xs[u].w[w];
                    }
                }
                if constexpr (MASK) {
                    bool any = false;
#pragma unroll
                    for (int u = 0; u < U; u++) any = any || tl[u];
                    if (__builtin_amdgcn_ballot_w64(any) != 0) {
#pragma unroll
                        for (int u = 0; u < U; u++)
                            if (tl[u]) seg_m_tail<XT, CPL>(X + (size_t)(A.N - 1) * ldx + c0, nv, xs[u]);
                    }
                }
                // the entry's sign as +1.0 or -1.0 (sign bit from bit 0 of the
                // index), then one fused multiply-add a column: (+-1) x is
                // exact, so fma(s, x, a) rounds once exactly as a + (+-x)
                // does, at one instruction a column instead of a select and
                // an add
#pragma unroll
                for (int u = 0; u < U; u++)
                    if (j0 + u < re[i]) {
                        const float sg = __uint_as_float(0xBF800000u ^ ((uint32_t)(v[u] & 1) << 31));
#pragma unroll
                        for (int m = 0; m < CPL; m++)
                            acc[i * CPL + m] = __builtin_fmaf(sg, seg_at(xs[u], m), acc[i * CPL + m]);
                    }
            }
        }

        if constexpr (NEAR) __syncthreads();  // the window is no longer read

        if constexpr (LONG)
            long_phase<K, XT, P>(A, X, ldx, u0, min(TRW, A.M - u0), lq, pv, acc, (acc_t *)smem, TRW * K,
                                 MASK ? kv : K);

#pragma unroll
        for (int i = 0; i < RI; i++) {
            const int r = u0 + GPW * i + g;
            if (r >= A.M) continue;
            float *yp = Y + (size_t)r * ldy + c0;
            // a masked 8-wide row (5..7 of its 8 columns stored) leaves
            // through plain stores, every lane of it: its bytes then meet in
            // L2 and leave as one line, where non-temporal partial stores
            // measured up to 1.8x slower; everything else streams past L2
            constexpr bool NT = !(MASK && K == 8);
            if (MASK && nv < CPL) {
                float f[CPL];
#pragma unroll
                for (int m = 0; m < CPL; m++) f[m] = (float)acc[i * CPL + m];
                store_first<CPL, NT>(yp, f, nv);
                continue;
            }
            if constexpr (CPL == 2) {
                st2<NT>(yp, (float)acc[i * 2], (float)acc[i * 2 + 1]);
            } else {
#pragma unroll
                for (int h = 0; h < CPL; h += 4)
                    st4<NT>(yp + h, (float)acc[i * CPL + h], (float)acc[i * CPL + h + 1], (float)acc[i * CPL + h + 2],
                            (float)acc[i * CPL + h + 3]);
            }
        }
    }
}

// LDS: s_val[capv] | s_lsum[256*P] | s_start[TRW+1]
__host__ __device__ constexpr size_t lds_bytes(int capv, int P, int trw)
{
    return (size_t)lsum_at(capv) * 4 + (size_t)kWG * P * sizeof(acc_t) + ((size_t)trw + 1) * 4;
}

}  // namespace tcd
