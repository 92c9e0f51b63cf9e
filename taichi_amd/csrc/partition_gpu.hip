// Partition and format build on the device (thumbnail and its morphology
// included: thumb.hip, morph_gpu.hip). The O(nnz) passes of
// partition_dia_shape run as kernels: the per-diagonal histogram
// (TaiChi.cu:723-734), the claim of dense-diagonal entries (:850-871), and
// the zero rows (:809-827). So does the merge into the encoded CSR
// (:480-612). The decisions (contours of the morphed thumbnail, cal_shape_size, the
// FP32 density test, the shape arrays) run on the host through the same
// functions as the host build (partition.cpp), so the TaiChi_new is the same.
//
// Claims are one byte per entry. The reference's zero-row walk consumes its
// ascending row list one entry per row. Without duplicate entries that is
// "rows of [st, en) holding no entry of the diagonal", which one bitmap per
// dense diagonal gives. A duplicate on a dense diagonal (an atomicOr finding
// its bit already set) makes the walk differ. The build then returns
// TC_ERR_UNSUPPORTED, and tc_partition_shapes runs the host build instead.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "partition.h"
#include "spmv_host.h"

namespace {

constexpr int kWG = 256;
constexpr int kHistBins = 2048;  // LDS bins; a wider band counts in global memory

int grid_rows(long long rows) { return (int)std::max(1LL, std::min(8192LL, (rows + kWG - 1) / kWG)); }

// device allocation freed on scope exit
struct DevMem {
    void *p = nullptr;
    DevMem() = default;
    DevMem(const DevMem &) = delete;
    DevMem &operator=(const DevMem &) = delete;
    DevMem(DevMem &&o) noexcept : p(o.p) { o.p = nullptr; }
    ~DevMem()
    {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 4); }
    template <class T>
    T *as() const
    {
        return (T *)p;
    }
};

// one dense diagonal of a kept shape: offset, local row range, zero bitmap
struct DiagDev {
    int off, st, en, pad;
    uint32_t *bits;  // bit (r - st) set: local row r holds an entry
};

// rows [xs, xe) of a kept shape and its diagonals [d0, d1) in the table
struct ShapeRows {
    int xs, xe, ys, d0, d1;
};

// diagonal index n = o + w - 1 of entry (x + rr, c), or -1 outside the band
__device__ inline int band_diag(int rr, int c, int y, int side, int w)
{
    if (c < y || c >= y + side) return -1;
    const int o = (c - y) - rr;
    if (o < -(w - 1) || o > w - 1) return -1;
    return o + w - 1;
}

// per-diagonal counts of the unclaimed entries in the square's band
__global__ __launch_bounds__(kWG) void part_hist(const int *__restrict__ rowptr, const int *__restrict__ col,
                                                 const uint8_t *__restrict__ claimed, int x, int y, int side, int w,
                                                 unsigned long long *__restrict__ cnt)
{
    __shared__ unsigned int h[kHistBins];
    const int nd = 2 * w - 1;
    const bool lds = nd <= kHistBins;
    if (lds)
        for (int n = threadIdx.x; n < nd; n += kWG) h[n] = 0;
    __syncthreads();
    for (int rr = blockIdx.x * kWG + threadIdx.x; rr < side; rr += gridDim.x * kWG) {
        const int r = x + rr;
        const int e = rowptr[r + 1];
        for (int j = rowptr[r]; j < e; j++) {
            if (claimed[j]) continue;
            const int n = band_diag(rr, col[j], y, side, w);
            if (n < 0) continue;
            if (lds) atomicAdd(&h[n], 1u);
            else atomicAdd(&cnt[n], 1ull);
        }
    }
    __syncthreads();
    if (lds)
        for (int n = threadIdx.x; n < nd; n += kWG)
            if (h[n]) atomicAdd(&cnt[n], (unsigned long long)h[n]);
}

// claim the unclaimed entries on dense diagonals and mark their rows
__global__ __launch_bounds__(kWG) void part_claim(const int *__restrict__ rowptr, const int *__restrict__ col,
                                                  uint8_t *__restrict__ claimed, int x, int y, int side, int w,
                                                  const int *__restrict__ dense_of, const DiagDev *__restrict__ dg,
                                                  int *__restrict__ dup)
{
    for (int rr = blockIdx.x * kWG + threadIdx.x; rr < side; rr += gridDim.x * kWG) {
        const int r = x + rr;
        const int e = rowptr[r + 1];
        for (int j = rowptr[r]; j < e; j++) {
            if (claimed[j]) continue;
            const int n = band_diag(rr, col[j], y, side, w);
            if (n < 0) continue;
            const int d = dense_of[n];
            if (d < 0) continue;
            claimed[j] = 1;
            const int b = rr - dg[d].st;  // in [0, en - st): the diagonal lies in the square there
            const uint32_t bit = 1u << (b & 31);
            if (atomicOr(&dg[d].bits[b >> 5], bit) & bit) *dup = 1;
        }
    }
}

// entries of row r in the format: zero rows of the shapes covering it (shape
// order, then diagonal order) followed by its unclaimed entries
template <bool FILL>
__global__ __launch_bounds__(kWG) void part_rows(int M, const int *__restrict__ rowptr, const int *__restrict__ col,
                                                 const uint8_t *__restrict__ claimed,
                                                 const ShapeRows *__restrict__ sh, int S,
                                                 const DiagDev *__restrict__ dg, long long *__restrict__ count,
                                                 const int *__restrict__ crow, int *__restrict__ ccol)
{
    for (int r = blockIdx.x * kWG + threadIdx.x; r < M; r += gridDim.x * kWG) {
        int pos = FILL ? crow[r] : 0;
        int k = 0;
        for (int s = 0; s < S; s++) {
            const ShapeRows q = sh[s];
            if (r < q.xs || r >= q.xe) continue;
            const int rr = r - q.xs;
            for (int g = q.d0; g < q.d1; g++) {
                const DiagDev d = dg[g];
                if (rr < d.st || rr >= d.en) continue;
                const int b = rr - d.st;
                if ((d.bits[b >> 5] >> (b & 31)) & 1u) continue;
                if (FILL) ccol[pos++] = (q.ys + rr + d.off) << 1;
                else k++;
            }
        }
        const int e = rowptr[r + 1];
        for (int j = rowptr[r]; j < e; j++) {
            if (claimed[j]) continue;
            if (FILL) ccol[pos++] = (col[j] << 1) + 1;
            else k++;
        }
        if (!FILL) count[r] = k;
    }
}

__global__ void narrow_rows(int n, const long long *__restrict__ a, int *__restrict__ b)
{
    for (int i = blockIdx.x * kWG + threadIdx.x; i < n; i += gridDim.x * kWG) b[i] = (int)a[i];
}

#define DCHECK(expr)                                                                                        \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess) return tc::fail(TC_ERR_HIP, "partition (device): %s: %s", #expr,              \
                                              hipGetErrorString(e_));                                       \
    } while (0)

// The passes above for the squares of one thumbnail; fills out like
// tc_partition_from_image with a NULL val, and claimed_out (nnz bytes, host;
// may be null) with each entry's claim byte.
int partition_on_device(int M, int N, int nnz, int *rowptr, int *col, const int *d_rowptr, const int *d_col,
                        const std::vector<int> &pts, TaiChi_new *out, hipStream_t s, uint8_t *claimed_out)
{
    std::vector<tc::Square> cand;
    tc::candidate_squares_from_pts(M, N, pts, cand);
    int max_nd = 1;
    for (const tc::Square &q : cand) max_nd = std::max(max_nd, 2 * tc::width_tier(q.side) - 1);

    DevMem claimed, cnt, dense_of, dup, shape_dg;
    DCHECK(claimed.alloc((size_t)nnz));
    DCHECK(cnt.alloc((size_t)max_nd * 8));
    DCHECK(dense_of.alloc((size_t)max_nd * 4));
    DCHECK(shape_dg.alloc((size_t)max_nd * sizeof(DiagDev)));
    DCHECK(dup.alloc(4));
    DCHECK(hipMemsetAsync(claimed.p, 0, (size_t)nnz, s));
    DCHECK(hipMemsetAsync(dup.p, 0, 4, s));

    std::vector<tc::DiaShape> kept;
    std::vector<DevMem> bitmaps;
    std::vector<DiagDev> diags;
    std::vector<unsigned long long> hcnt;
    std::vector<long long> lcnt;
    std::vector<int> dof;
    for (const tc::Square &q : cand) {
        const int w = tc::width_tier(q.side);
        const int nd = 2 * w - 1;
        if (nd <= 0) continue;
        DCHECK(hipMemsetAsync(cnt.p, 0, (size_t)nd * 8, s));
        part_hist<<<grid_rows(q.side), kWG, 0, s>>>(d_rowptr, d_col, claimed.as<uint8_t>(), q.x, q.y, q.side, w,
                                                    cnt.as<unsigned long long>());
        hcnt.resize((size_t)nd);
        DCHECK(hipMemcpyAsync(hcnt.data(), cnt.p, (size_t)nd * 8, hipMemcpyDeviceToHost, s));
        DCHECK(hipStreamSynchronize(s));
        lcnt.assign(hcnt.begin(), hcnt.end());
        tc::DiaShape sh;
        sh.xs = q.x;
        sh.ys = q.y;
        sh.side = q.side;
        sh.width = w;
        tc::decide_dense(q.side, w, lcnt.data(), sh, dof);
        const int neg = (int)sh.offs.size();
        if (neg == 0) continue;
        size_t words = 0;
        std::vector<size_t> woff((size_t)neg);
        for (int d = 0; d < neg; d++) {
            woff[d] = words;
            words += ((size_t)(sh.en[d] - sh.st[d]) + 31) / 32;
        }
        DevMem bm;
        DCHECK(bm.alloc(words * 4));
        DCHECK(hipMemsetAsync(bm.p, 0, words * 4, s));
        const size_t first = diags.size();
        for (int d = 0; d < neg; d++)
            diags.push_back({sh.offs[d], sh.st[d], sh.en[d], 0, bm.as<uint32_t>() + woff[d]});
        DCHECK(hipMemcpyAsync(dense_of.p, dof.data(), (size_t)nd * 4, hipMemcpyHostToDevice, s));
        DCHECK(hipMemcpyAsync(shape_dg.p, diags.data() + first, (size_t)neg * sizeof(DiagDev),
                              hipMemcpyHostToDevice, s));
        part_claim<<<grid_rows(q.side), kWG, 0, s>>>(d_rowptr, d_col, claimed.as<uint8_t>(), q.x, q.y, q.side, w,
                                                     dense_of.as<int>(), shape_dg.as<DiagDev>(), dup.as<int>());
        DCHECK(hipGetLastError());
        // the host copies above must finish before dof/diags change
        DCHECK(hipStreamSynchronize(s));
        bitmaps.push_back(std::move(bm));
        kept.push_back(std::move(sh));
    }
    int hdup = 0;
    DCHECK(hipMemcpyAsync(&hdup, dup.p, 4, hipMemcpyDeviceToHost, s));
    DCHECK(hipStreamSynchronize(s));
    if (hdup) return tc::fail(TC_ERR_UNSUPPORTED, "partition (device): duplicate entries on a dense diagonal");

    if (claimed_out && nnz > 0) {
        DCHECK(hipMemcpyAsync(claimed_out, claimed.p, (size_t)nnz, hipMemcpyDeviceToHost, s));
        DCHECK(hipStreamSynchronize(s));
    }
    memset(out, 0, sizeof *out);
    tc::shape_arrays(kept, nnz, out);
    if (out->dense_nnz == 0) {  // no dense diagonal: the input CSR is the format (TaiChi.cu:575-578)
        out->csrRow = rowptr;
        out->csrCol = col;
        return TC_OK;
    }
    auto drop = [&](int status) {
        free(out->xStart);
        free(out->yStart);
        free(out->xStop);
        free(out->nDiasPtr);
        free(out->neg_offsets);
        memset(out, 0, sizeof *out);
        return status;
    };
    const int S = (int)kept.size();
    std::vector<ShapeRows> rows_tab;
    int g = 0;
    for (const tc::DiaShape &sh : kept) {
        rows_tab.push_back({sh.xs, sh.xs + sh.side, sh.ys, g, g + (int)sh.offs.size()});
        g += (int)sh.offs.size();
    }
    DevMem d_rows, d_dg, count, crow64, crow, scan_tmp;
    if (d_rows.alloc(rows_tab.size() * sizeof(ShapeRows)) != hipSuccess ||
        d_dg.alloc(diags.size() * sizeof(DiagDev)) != hipSuccess || count.alloc((size_t)M * 8) != hipSuccess ||
        crow64.alloc(((size_t)M + 1) * 8) != hipSuccess || crow.alloc(((size_t)M + 1) * 4) != hipSuccess)
        return drop(tc::fail(TC_ERR_NOMEM, "partition (device): out of device memory"));
    size_t tmp_bytes = 0;
    if (hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, count.as<long long>(), crow64.as<long long>() + 1, M,
                                         s) != hipSuccess ||
        scan_tmp.alloc(tmp_bytes) != hipSuccess)
        return drop(tc::fail(TC_ERR_HIP, "partition (device): scan set-up failed"));
    hipError_t e = hipMemcpyAsync(d_rows.p, rows_tab.data(), rows_tab.size() * sizeof(ShapeRows),
                                  hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_dg.p, diags.data(), diags.size() * sizeof(DiagDev), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemsetAsync(crow64.p, 0, 8, s);
    if (e != hipSuccess) return drop(tc::fail(TC_ERR_HIP, "partition (device): %s", hipGetErrorString(e)));
    part_rows<false><<<grid_rows(M), kWG, 0, s>>>(M, d_rowptr, d_col, claimed.as<uint8_t>(), d_rows.as<ShapeRows>(),
                                                  S, d_dg.as<DiagDev>(), count.as<long long>(), nullptr, nullptr);
    e = hipcub::DeviceScan::InclusiveSum(scan_tmp.p, tmp_bytes, count.as<long long>(), crow64.as<long long>() + 1, M,
                                         s);
    long long R = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&R, crow64.as<long long>() + M, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return drop(tc::fail(TC_ERR_HIP, "partition (device): %s", hipGetErrorString(e)));
    const long long sparse = (long long)nnz - out->dense_nnz;
    if (R > INT_MAX || R < sparse)
        return drop(tc::fail(TC_ERR_UNSUPPORTED, "partition (device): %lld format entries exceed int32", R));
    DevMem ccol;
    if (ccol.alloc((size_t)R * 4) != hipSuccess)
        return drop(tc::fail(TC_ERR_NOMEM, "partition (device): out of device memory"));
    narrow_rows<<<grid_rows(M + 1LL), kWG, 0, s>>>(M + 1, crow64.as<long long>(), crow.as<int>());
    part_rows<true><<<grid_rows(M), kWG, 0, s>>>(M, d_rowptr, d_col, claimed.as<uint8_t>(), d_rows.as<ShapeRows>(),
                                                 S, d_dg.as<DiagDev>(), nullptr, crow.as<int>(), ccol.as<int>());
    int *hrow = (int *)malloc(((size_t)M + 1) * 4);
    int *hcol = (int *)malloc((size_t)(R ? R : 1) * 4);
    if (!hrow || !hcol) {
        free(hrow);
        free(hcol);
        return drop(tc::fail(TC_ERR_NOMEM, "partition (device): out of host memory"));
    }
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(hrow, crow.p, ((size_t)M + 1) * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && R) e = hipMemcpyAsync(hcol, ccol.p, (size_t)R * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        free(hrow);
        free(hcol);
        return drop(tc::fail(TC_ERR_HIP, "partition (device): %s", hipGetErrorString(e)));
    }
    out->csrRow = hrow;
    out->csrCol = hcol;
    out->zero_elements = (int)(R - sparse);
    return TC_OK;
}

}  // namespace

int tc::partition_device_claims(int M, int N, int nnz, int *rowptr, int *col, const int *d_rowptr, const int *d_col,
                                TaiChi_new *out, void *stream, uint8_t *claimed)
{
    if (M < 0 || N < 0 || nnz < 0 || !rowptr || (nnz > 0 && !col) || !out || (M > 0 && !d_rowptr) ||
        (nnz > 0 && !d_col))
        return tc::fail(TC_ERR_INVALID, "tc_partition_device: bad argument");
    if (rowptr[0] != 0 || rowptr[M] != nnz)
        return tc::fail(TC_ERR_INVALID, "tc_partition_device: rowptr[0]=%d rowptr[M]=%d nnz=%d", rowptr[0],
                        rowptr[M], nnz);
    hipStream_t s = (hipStream_t)stream;
    int H, W, pM, pN;
    std::vector<int> pts;  // contour points; none when M <= 512 (no thumbnail)
    if (tc_thumb_dims(M, N, &H, &W, &pM, &pN) == 0) {
        DevMem dimg;
        DCHECK(dimg.alloc((size_t)H * W));
        int rc = tc_thumbnail(M, N, d_rowptr, d_col, dimg.as<unsigned char>(), s);
        if (rc == TC_OK) rc = tc::sdf_on_device(dimg.as<unsigned char>(), H, W, pts, s);
        if (rc != TC_OK) return rc;
    }
    return partition_on_device(M, N, nnz, rowptr, col, d_rowptr, d_col, pts, out, s, claimed);
}

extern "C" int tc_partition_device(int M, int N, int nnz, int *rowptr, int *col, const int *d_rowptr,
                                   const int *d_col, TaiChi_new *out, void *stream)
try {
    return tc::partition_device_claims(M, N, nnz, rowptr, col, d_rowptr, d_col, out, stream, nullptr);
} catch (...) {
    return tc::exception_status(__func__);
}
