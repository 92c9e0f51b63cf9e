// SpMV kernels for one right-hand side (k = 1): one row per lane
// (lane-strided) and quad rows (4 consecutive rows per lane), with their
// launchers.
#include "spmv_host.h"

namespace tcd {

// STR: x and y are one column of row-major X and Y (leading dimensions ldx,
// ldy), as a served request's single column is; otherwise plain vectors.
template <typename XT, int RPT, int OW, bool FAST, bool LONG, bool STR>
__global__ __launch_bounds__(kWG) void spmv_rows_kernel(MatView A, const XT *__restrict__ X, int ldx,
                                                        float *__restrict__ Y, int ldy, int ntiles, int capv)
{
    extern __shared__ float smem[];
    int tb, te;
    block_units(ntiles, &tb, &te);
    for (int t = tb; t < te; t++)
        rows_tile<XT, RPT, OW, FAST, LONG>(A, X, STR ? ldx : 1, Y, STR ? ldy : 1, t, smem, capv, StoreY{}, t + 1 < te);
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
// LDS: s_val[capv] (also the long-row sums afterwards: TRW floats; capv
// >= 2*TRW when the matrix has long rows) | s_win[TRW] (DIAG).
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

    acc_t acc[Q];
#pragma unroll
    for (int i = 0; i < Q; i++) acc[i] = 0.0;
    if constexpr (DIAG) {
        const ShapeDev &sh = A.sh0;
#pragma unroll
        for (int i = 0; i < Q; i++) {
            const int r = r0 + i, c = w0 + Q * tid + i;
            if (r < A.M && r >= sh.xs && r < sh.xe && c >= sh.ys && c < sh.ys + (sh.xe - sh.xs))
                accum(acc[i], xw[i]);
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
            if constexpr (sizeof(XT) == 4 && !DIAG) {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    xv[k] = 0.0f;
                    if (q + k < ne) xv[k] = load_x<XT>(X, (size_t)(vv[k] >> 1));
                }
            } else {
                // bits behind the branch, the window value and the conversion
                // after all four gathers are issued (a conversion, or a window
                // read into the gather's register, inside it serialises them)
                uint32_t xb[4];
                bool inw[4];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int col = vv[k] >> 1;
                    inw[k] = DIAG && (unsigned)(col - w0) < (unsigned)TRW;
                    xb[k] = 0u;
                    if (!inw[k] && q + k < ne) xb[k] = load_xbits<XT>(X, (size_t)col);
                }
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    xv[k] = xbits_f<XT>(xb[k]);
                    if constexpr (DIAG) xv[k] = inw[k] ? s_win[inw[k] ? (vv[k] >> 1) - w0 : 0] : xv[k];
                }
            }
            if (q < ne)
                *(float4 *)(s_val + q) = make_float4((vv[0] & 1) ? xv[0] : -xv[0], (vv[1] & 1) ? xv[1] : -xv[1],
                                                     (vv[2] & 1) ? xv[2] : -xv[2], (vv[3] & 1) ? xv[3] : -xv[3]);
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < Q; i++) {
            const int lo = max(rs[i], e0), hi = min(re[i], e0 + ne);
            acc_t a = acc[i];
            sum_run<1>(a, s_val, lo - e0, hi - e0);
            acc[i] = a;
        }
        __syncthreads();
    }

    if constexpr (LONG) {
        int prow[Q], lq[Q];
        bool pv[Q];
#pragma unroll
        for (int i = 0; i < Q; i++) prow[i] = Q * tid + i, lq[i] = prow[i], pv[i] = r0 + i < A.M;
        long_phase<1, XT, Q>(A, X, 1, u0, min(TRW, A.M - u0), lq, pv, acc, (acc_t *)s_val, TRW);
    }

    if (full) {
        if constexpr (Q == 4)
            st_nt4(Y + r0, (float)acc[0], (float)acc[1], (float)acc[2], (float)acc[3]);
        else st_nt2(Y + r0, (float)acc[0], (float)acc[1]);
    } else {
#pragma unroll
        for (int i = 0; i < Q; i++)
            if (r0 + i < A.M) st_nt(&Y[r0 + i], (float)acc[i]);
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

// ---- launchers ----

template <typename XT, int RPT, int OW, bool FAST, bool LONG>
void launch_rows_k(const MatView &A, const void *X, int ldx, float *Y, int ldy, int ntiles, int capv, int G,
                   size_t lds, hipStream_t s)
{
    if (ldx == 1 && ldy == 1)
        spmv_rows_kernel<XT, RPT, OW, FAST, LONG, false><<<G, kWG, lds, s>>>(A, (const XT *)X, 1, Y, 1, ntiles, capv);
    else if constexpr (RPT <= 2)
        spmv_rows_kernel<XT, RPT, OW, FAST, LONG, true><<<G, kWG, lds, s>>>(A, (const XT *)X, ldx, Y, ldy, ntiles,
                                                                            capv);
}

template <typename XT, int RPT>
int launch_rows_t(const MatView &A, const void *X, int ldx, float *Y, int ldy, int capv, int grid, hipStream_t s)
{
    constexpr int TRW = kWG * RPT;
    const int ntiles = (A.M + TRW - 1) / TRW;
    if (ntiles == 0) return TC_OK;
    const int G = grid > 0 ? std::min(grid, ntiles) : ntiles;
    const bool fast = A.S == 1 && A.sh0.oe - A.sh0.ob == 1 && A.sh0.pm < 0;
    const size_t lds = ((size_t)lsum_at(capv) + (A.L > 0 ? 2 * TRW : 0)) * 4;
    if (lds > max_lds_bytes())
        return tc::fail(TC_ERR_UNSUPPORTED, "rows tiles of %d with %d LDS values need %zu B of LDS", TRW, capv, lds);
    const bool lng = A.L > 0;
    // specialised on offset width, the one-shape-one-offset case and long
    // rows, so the common case carries no registers for the general paths
    if (A.ow == 1) {
        if (fast) lng ? launch_rows_k<XT, RPT, 1, true, true>(A, X, ldx, Y, ldy, ntiles, capv, G, lds, s)
                      : launch_rows_k<XT, RPT, 1, true, false>(A, X, ldx, Y, ldy, ntiles, capv, G, lds, s);
        else lng ? launch_rows_k<XT, RPT, 1, false, true>(A, X, ldx, Y, ldy, ntiles, capv, G, lds, s)
                 : launch_rows_k<XT, RPT, 1, false, false>(A, X, ldx, Y, ldy, ntiles, capv, G, lds, s);
    } else {
        if (fast) lng ? launch_rows_k<XT, RPT, 2, true, true>(A, X, ldx, Y, ldy, ntiles, capv, G, lds, s)
                      : launch_rows_k<XT, RPT, 2, true, false>(A, X, ldx, Y, ldy, ntiles, capv, G, lds, s);
        else lng ? launch_rows_k<XT, RPT, 2, false, true>(A, X, ldx, Y, ldy, ntiles, capv, G, lds, s)
                 : launch_rows_k<XT, RPT, 2, false, false>(A, X, ldx, Y, ldy, ntiles, capv, G, lds, s);
    }
    HIPCHECK(hipGetLastError());
    return TC_OK;
}

// strided columns are compiled for tiles of 256 and 512 rows only; other
// tile sizes run them as 512
template <typename XT>
int launch_rows_x(const MatView &A, const void *X, int ldx, float *Y, int ldy, const tc_config &c, int capv,
                  hipStream_t s)
{
    const bool str = ldx != 1 || ldy != 1;
    switch (c.rows_per_tile) {
    case 256: return launch_rows_t<XT, 1>(A, X, ldx, Y, ldy, capv, c.grid, s);
    case 512: return launch_rows_t<XT, 2>(A, X, ldx, Y, ldy, capv, c.grid, s);
    }
    if (str) return launch_rows_t<XT, 2>(A, X, ldx, Y, ldy, capv, c.grid, s);
    switch (c.rows_per_tile) {
    case 1024: return launch_rows_t<XT, 4>(A, X, 1, Y, 1, capv, c.grid, s);
    case 4096: return launch_rows_t<XT, 16>(A, X, 1, Y, 1, capv, c.grid, s);
    default: return launch_rows_t<XT, 8>(A, X, 1, Y, 1, capv, c.grid, s);
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
    if (!(A.S == 0 || (A.S == 1 && A.sh0.oe - A.sh0.ob == 1 && A.sh0.pm < 0))) return false;
    const int Q = c.rows_per_tile == 512 ? 2 : c.rows_per_tile == 1024 ? 4 : 0;
    if (Q == 0 || capv % 4 != 0 || capv < kWG * Q) return false;
    if (A.L > 0 && capv < 2 * kWG * Q) return false;  // the long-row sums reuse s_val
    return ((uintptr_t)X % ((uintptr_t)Q * esize)) == 0 && ((uintptr_t)Y % ((uintptr_t)Q * 4)) == 0;
}

template <typename XT>
int launch_quad_x(const MatView &A, const void *X, float *Y, const tc_config &c, int capv, hipStream_t s)
{
    if (c.rows_per_tile == 1024) return launch_quad_t<XT, 4>(A, X, Y, capv, c.grid, s);
    return launch_quad_t<XT, 2>(A, X, Y, capv, c.grid, s);
}

int launch_rows(const MatView &A, const void *X, int ldx, float *Y, int ldy, int dtype, const tc_config &c, int capv,
                hipStream_t s)
{
    if (dtype == TC_BF16) return launch_rows_x<bf16_t>(A, X, ldx, Y, ldy, c, capv, s);
    if (dtype == TC_F16) return launch_rows_x<f16_t>(A, X, ldx, Y, ldy, c, capv, s);
    return launch_rows_x<float>(A, X, ldx, Y, ldy, c, capv, s);
}

int launch_quad(const MatView &A, const void *X, float *Y, int dtype, const tc_config &c, int capv, hipStream_t s)
{
    if (dtype == TC_BF16) return launch_quad_x<bf16_t>(A, X, Y, c, capv, s);
    if (dtype == TC_F16) return launch_quad_x<f16_t>(A, X, Y, c, capv, s);
    return launch_quad_x<float>(A, X, Y, c, capv, s);
}

}  // namespace tcd
