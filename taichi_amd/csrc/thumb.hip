// Kernel (a): density thumbnail of getSdf (morohutils.cu:153-229) on gfx950.
//
// The reference starts a 512 x W byte image at 255 and lets every nonzero
// (i, c) store 0 at [i/pM][c/pN] from a thread-per-row-chunk kernel. Here
// each workgroup owns a contiguous row range, which lands on a few pixel rows
// (pM = M/512 rows per pixel row); for each such pixel row the workgroup
// streams that row band's column indices with coalesced loads and sets bits
// of an LDS bitset with integer atomicOr, then merges its non-empty words into
// a global bitset with atomicOr. A second pass expands bits to the 0/255 image.
// OR is order independent, so the image is bit-exact against the closed form.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <vector>

#include "common.h"
#include "partition.h"

namespace {

constexpr int kWG = 256;
constexpr int kThumbH = 512;             // morohutils.cu:191
constexpr int kMaxLdsWords = 16384;      // 64 KiB of LDS bitset per workgroup

__global__ __launch_bounds__(kWG) void thumb_clear(unsigned *__restrict__ bits, int nwords)
{
    for (int i = blockIdx.x * kWG + threadIdx.x; i < nwords; i += gridDim.x * kWG) bits[i] = 0u;
}

// rows [blockIdx*rpb, +rpb) -> pixel rows [pr0, pr1]; use_lds selects the
// LDS-privatised path (needs (pr1-pr0+1)*WW <= kMaxLdsWords).
__global__ __launch_bounds__(kWG) void thumb_bin(int M, int pM, int pN, int W, int WW, int rpb, int use_lds,
                                                 const int *__restrict__ rowptr, const int *__restrict__ col,
                                                 unsigned *__restrict__ gbits)
{
    extern __shared__ unsigned sbits[];
    const int r0 = blockIdx.x * rpb;
    if (r0 >= M) return;
    const int r1 = min(M, r0 + rpb);
    const int pr0 = r0 / pM;
    if (pr0 >= kThumbH) return;  // rows past 512*pM are dropped (morohutils.cu:174)
    const int pr1 = min(kThumbH - 1, (r1 - 1) / pM);
    const int nwords = (pr1 - pr0 + 1) * WW;
    const int tid = threadIdx.x;
    if (use_lds) {
        for (int i = tid; i < nwords; i += kWG) sbits[i] = 0u;
        __syncthreads();
    }
    for (int pr = pr0; pr <= pr1; pr++) {
        const int ra = max(r0, pr * pM), rb = min(r1, (pr + 1) * pM);
        const int e0 = rowptr[ra], e1 = rowptr[rb];
        unsigned *dst = use_lds ? sbits + (pr - pr0) * WW : gbits + (size_t)pr * WW;
        for (int e = e0 + tid; e < e1; e += kWG) {
            const int y = col[e] / pN;
            if (y < W) atomicOr(&dst[y >> 5], 1u << (y & 31));
        }
    }
    if (use_lds) {
        __syncthreads();
        for (int i = tid; i < nwords; i += kWG) {
            const unsigned v = sbits[i];
            if (v) atomicOr(&gbits[(size_t)pr0 * WW + i], v);
        }
    }
}

__global__ __launch_bounds__(kWG) void thumb_expand(const unsigned *__restrict__ bits, int H, int W, int WW,
                                                    unsigned char *__restrict__ img)
{
    const long long n = (long long)H * W;
    for (long long p = (long long)blockIdx.x * kWG + threadIdx.x; p < n; p += (long long)gridDim.x * kWG) {
        const int x = (int)(p / W), y = (int)(p % W);
        img[p] = ((bits[(size_t)x * WW + (y >> 5)] >> (y & 31)) & 1u) ? 0 : 255;
    }
}

}  // namespace

#define HIPCHECK(expr)                                                                                      \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess) return tc::fail(TC_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));          \
    } while (0)

extern "C" int tc_thumbnail(int M, int N, const int *d_rowptr, const int *d_col, unsigned char *d_img, void *stream)
try {
    int H, W, pM, pN;
    int rc = tc_thumb_dims(M, N, &H, &W, &pM, &pN);
    if (rc != 0) return tc::fail(TC_ERR_INVALID, "tc_thumbnail: no thumbnail for M=%d N=%d (code %d)", M, N, rc);
    if (!d_rowptr || !d_col || !d_img) return tc::fail(TC_ERR_INVALID, "tc_thumbnail: null pointer");
    if ((long long)H * W > (1ll << 31)) return tc::fail(TC_ERR_UNSUPPORTED, "tc_thumbnail: image too wide");
    hipStream_t s = (hipStream_t)stream;
    const int WW = (W + 31) / 32;
    unsigned *bits = nullptr;
    HIPCHECK(hipMallocAsync((void **)&bits, (size_t)H * WW * sizeof(unsigned), s));
    const int nwords = H * WW;
    thumb_clear<<<std::min(1024, (nwords + kWG - 1) / kWG), kWG, 0, s>>>(bits, nwords);
    // ~2048 workgroups, at least one pixel row of work each when pM is large
    int rpb = std::max(256, (M + 2047) / 2048);
    int nblk = (M + rpb - 1) / rpb;
    int max_prows = std::min(kThumbH, rpb / pM + 2);
    int use_lds = (long long)max_prows * WW <= kMaxLdsWords;
    size_t lds = use_lds ? (size_t)max_prows * WW * sizeof(unsigned) : 0;
    thumb_bin<<<nblk, kWG, lds, s>>>(M, pM, pN, W, WW, rpb, use_lds, d_rowptr, d_col, bits);
    long long npix = (long long)H * W;
    int eblk = (int)std::min<long long>(4096, (npix + kWG - 1) / kWG);
    thumb_expand<<<eblk, kWG, 0, s>>>(bits, H, W, WW, d_img);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipFreeAsync(bits, s));
    return TC_OK;
} catch (...) {
    return tc::exception_status(__func__);
}

extern "C" int tc_partition_shapes(const char *path, int M, int N, int nnz, int *rowptr, int *col, float *val,
                                   TaiChi_new *out)
try {
    (void)path;
    if (M < 0 || N < 0 || nnz < 0 || !rowptr || (nnz && !col) || !out)
        return tc::fail(TC_ERR_INVALID, "tc_partition_shapes: bad argument");
    int H, W, pM, pN;
    std::vector<unsigned char> img;
    if (tc_thumb_dims(M, N, &H, &W, &pM, &pN) == 0) {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
            return tc::fail(TC_ERR_NODEVICE, "tc_partition_shapes: no HIP device for the thumbnail kernel");
        struct Buffers {  // freed on every return path
            int *drow = nullptr, *dcol = nullptr;
            unsigned char *dimg = nullptr;
            ~Buffers()
            {
                (void)hipFree(drow);
                (void)hipFree(dcol);
                (void)hipFree(dimg);
            }
        } b;
        img.resize((size_t)H * W);
        HIPCHECK(hipMalloc((void **)&b.drow, ((size_t)M + 1) * sizeof(int)));
        HIPCHECK(hipMalloc((void **)&b.dcol, (size_t)(nnz ? nnz : 1) * sizeof(int)));
        HIPCHECK(hipMalloc((void **)&b.dimg, img.size()));
        HIPCHECK(hipMemcpy(b.drow, rowptr, ((size_t)M + 1) * sizeof(int), hipMemcpyHostToDevice));
        if (nnz) HIPCHECK(hipMemcpy(b.dcol, col, (size_t)nnz * sizeof(int), hipMemcpyHostToDevice));
        // the device build; with val, its claim bytes become the FLT_MAX
        // markers the reference writes (TaiChi.cu:866). An input whose val
        // already holds markers (entries the reference counts as claimed
        // from the start, :730) takes the host build, as do duplicates on a
        // dense diagonal (the device build reports them as unsupported).
        bool premarked = false;
        if (val)
            for (int j = 0; j < nnz && !premarked; j++) premarked = !(std::fabs(val[j] - FLT_MAX) > 1e-8);
        int rc = TC_ERR_UNSUPPORTED;
        if (!premarked) {
            std::vector<uint8_t> cl(val ? (size_t)nnz : 0);
            rc = tc::partition_device_claims(M, N, nnz, rowptr, col, b.drow, b.dcol, out, nullptr,
                                             val ? cl.data() : nullptr);
            if (rc == TC_OK && val)
                for (int j = 0; j < nnz; j++)
                    if (cl[j]) val[j] = FLT_MAX;
        }
        const bool host_build = rc == TC_ERR_UNSUPPORTED;
        if (host_build) {
            rc = tc_thumbnail(M, N, b.drow, b.dcol, b.dimg, nullptr);
            if (rc == TC_OK) HIPCHECK(hipMemcpy(img.data(), b.dimg, img.size(), hipMemcpyDeviceToHost));
        }
        if (!host_build || rc != TC_OK) return rc;
    }
    return tc_partition_from_image(M, N, nnz, rowptr, col, val, img.empty() ? nullptr : img.data(), out);
} catch (...) {
    return tc::exception_status(__func__);
}
