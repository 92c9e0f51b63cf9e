// Thumbnail morphology of getSdf (morohutils.cu:231-251) on the device: the
// same OpenCV semantics as the host code (morph.cpp), pixel for pixel.
//   rev  = 255 - mat                                  (:232)
//   fill = imfill(rev)                                (:235, imfill :73-83)
//   o1   = open(fill, 30x1 line), o2 = open(o1, 1x30) (:239-243)
//   del  = rev - o2, saturated                        (:245)
//   thin = open(del, eye(5))                          (:250-251)
// Contour following (:257-258) stays on the host: it is a sequential walk
// over the few border pixels of a 512-row image.
//
// imfill floods, 4-connected, the pixels equal to 0 that connect to the
// one-pixel zero frame around the image. Here that region is the connected
// component of padded pixel 0 among the zero pixels, found by union-find.
// Parent labels only decrease (atomicMin), and a root is only linked to a
// smaller root, so the component holding pixel 0 has root 0. A find chain
// strictly decreases (L[x] <= x, even when a read is stale). Each failed
// union retry strictly lowers one of its two labels, so every loop ends.
// The union loop also carries a hard cap that raises an error flag instead
// of looping on.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "partition.h"
#include "spmv_host.h"

namespace {

constexpr int kWG = 256;

int grid_for(long long n) { return (int)std::max(1LL, std::min(4096LL, (n + kWG - 1) / kWG)); }

// padded image (zero frame), union-find labels over its zero pixels, rev
__global__ __launch_bounds__(kWG) void mo_init(const uint8_t *__restrict__ mat, int H, int W,
                                               uint8_t *__restrict__ pad, int *__restrict__ lab,
                                               uint8_t *__restrict__ rev)
{
    const int Wp = W + 2, n = (H + 2) * Wp;
    for (int p = blockIdx.x * kWG + threadIdx.x; p < n; p += gridDim.x * kWG) {
        const int py = p / Wp, px = p - py * Wp;
        const bool inside = py >= 1 && py <= H && px >= 1 && px <= W;
        const uint8_t v = inside ? (uint8_t)(255 - mat[(py - 1) * W + (px - 1)]) : (uint8_t)0;
        pad[p] = v;
        lab[p] = v == 0 ? p : -1;
        if (inside) rev[(py - 1) * W + (px - 1)] = v;
    }
}

__device__ inline int uf_find(const int *lab, int x)
{
    int p;
    while ((p = lab[x]) != x) x = p;  // p < x: strictly decreasing
    return x;
}

// link the components of zero pixels a and b; false if the cap was reached
__device__ inline bool uf_union(int *lab, int a, int b, int cap)
{
    for (int it = 0; it < cap; it++) {
        a = uf_find(lab, a);
        b = uf_find(lab, b);
        if (a == b) return true;
        if (a < b) {
            const int old = atomicMin(&lab[b], a);
            if (old == b) return true;
            b = old;  // b was no longer a root: old < b
        } else {
            const int old = atomicMin(&lab[a], b);
            if (old == a) return true;
            a = old;
        }
    }
    return false;
}

// 4-connected unions of each zero pixel with its east and south neighbours
__global__ __launch_bounds__(kWG) void mo_union(int H, int W, const uint8_t *__restrict__ pad, int *lab,
                                                int *__restrict__ err)
{
    const int Wp = W + 2, Hp = H + 2, n = Hp * Wp;
    const int cap = 2 * n + 2;
    for (int p = blockIdx.x * kWG + threadIdx.x; p < n; p += gridDim.x * kWG) {
        if (pad[p] != 0) continue;
        const int py = p / Wp, px = p - py * Wp;
        bool ok = true;
        if (px + 1 < Wp && pad[p + 1] == 0) ok = uf_union(lab, p, p + 1, cap) && ok;
        if (py + 1 < Hp && pad[p + Wp] == 0) ok = uf_union(lab, p, p + Wp, cap) && ok;
        if (!ok) atomicOr(err, 1);
    }
}

// fill = rev | ~cut, cut = 255 where the flood from the frame reached, else pad
__global__ __launch_bounds__(kWG) void mo_fill(int H, int W, const uint8_t *__restrict__ pad,
                                               const int *__restrict__ lab, const uint8_t *__restrict__ rev,
                                               uint8_t *__restrict__ fill)
{
    const int n = H * W, Wp = W + 2;
    for (int i = blockIdx.x * kWG + threadIdx.x; i < n; i += gridDim.x * kWG) {
        const int y = i / W, x = i - y * W;
        const int p = (y + 1) * Wp + x + 1;
        const uint8_t pv = pad[p];
        const uint8_t cut = (pv == 0 && uf_find(lab, p) == 0) ? (uint8_t)255 : pv;
        fill[i] = (uint8_t)(rev[i] | (uint8_t)~cut);
    }
}

// erode (min, outside 255) or dilate (max, outside 0) over the 30-long line
// element with anchor 15: the window [i - 15, i + 14] along rows or columns
template <bool ROWS, bool MIN>
__global__ __launch_bounds__(kWG) void mo_line(int H, int W, const uint8_t *__restrict__ src,
                                               uint8_t *__restrict__ dst)
{
    const int n = H * W;
    for (int i = blockIdx.x * kWG + threadIdx.x; i < n; i += gridDim.x * kWG) {
        const int y = i / W, x = i - y * W;
        int m = MIN ? 255 : 0;
        for (int t = -15; t <= 14; t++) {
            const int xx = ROWS ? x + t : x, yy = ROWS ? y : y + t;
            const bool out = ROWS ? (xx < 0 || xx >= W) : (yy < 0 || yy >= H);
            const int v = out ? (MIN ? 255 : 0) : src[yy * W + xx];
            m = MIN ? min(m, v) : max(m, v);
        }
        dst[i] = (uint8_t)m;
    }
}

// erode of del = sat(rev - o2) with eye(5), anchor (2, 2): taps (d, d)
__global__ __launch_bounds__(kWG) void mo_eye_erode(int H, int W, const uint8_t *__restrict__ rev,
                                                    const uint8_t *__restrict__ o2, uint8_t *__restrict__ dst)
{
    const int n = H * W;
    for (int i = blockIdx.x * kWG + threadIdx.x; i < n; i += gridDim.x * kWG) {
        const int y = i / W, x = i - y * W;
        int m = 255;
        for (int d = -2; d <= 2; d++) {
            const int xx = x + d, yy = y + d;
            int v = 255;
            if (xx >= 0 && xx < W && yy >= 0 && yy < H) {
                const int a = rev[yy * W + xx], b = o2[yy * W + xx];
                v = a > b ? a - b : 0;
            }
            m = min(m, v);
        }
        dst[i] = (uint8_t)m;
    }
}

__global__ __launch_bounds__(kWG) void mo_eye_dilate(int H, int W, const uint8_t *__restrict__ src,
                                                     uint8_t *__restrict__ dst)
{
    const int n = H * W;
    for (int i = blockIdx.x * kWG + threadIdx.x; i < n; i += gridDim.x * kWG) {
        const int y = i / W, x = i - y * W;
        int m = 0;
        for (int d = -2; d <= 2; d++) {
            const int xx = x + d, yy = y + d;
            const int v = (xx >= 0 && xx < W && yy >= 0 && yy < H) ? src[yy * W + xx] : 0;
            m = max(m, v);
        }
        dst[i] = (uint8_t)m;
    }
}

size_t a256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

// scratch: pad[(H+2)(W+2)] | lab[(H+2)(W+2)] | rev | fill | t1 | t2 | err
extern "C" size_t tc_morph_work_bytes(int H, int W)
{
    if (H <= 0 || W <= 0) return 0;
    const size_t np = (size_t)(H + 2) * (W + 2), n = (size_t)H * W;
    return a256(np) + a256(np * 4) + 4 * a256(n) + 256;
}

extern "C" int tc_morph_device(const unsigned char *d_img, int H, int W, unsigned char *d_thin, void *work,
                               void *stream)
try {
    if (!d_img || !d_thin || !work || H <= 0 || W <= 0 || (long long)(H + 2) * (W + 2) > (1LL << 30))
        return tc::fail(TC_ERR_INVALID, "tc_morph_device: bad argument");
    hipStream_t s = (hipStream_t)stream;
    const size_t np = (size_t)(H + 2) * (W + 2), n = (size_t)H * W;
    char *p = (char *)work;
    uint8_t *pad = (uint8_t *)p;
    p += a256(np);
    int *lab = (int *)p;
    p += a256(np * 4);
    uint8_t *rev = (uint8_t *)p, *fill = (uint8_t *)p + a256(n), *t1 = (uint8_t *)p + 2 * a256(n),
            *t2 = (uint8_t *)p + 3 * a256(n);
    int *err = (int *)(p + 4 * a256(n));
    const int gp = grid_for((long long)np), gn = grid_for((long long)n);
    HIPCHECK(hipMemsetAsync(err, 0, 4, s));
    mo_init<<<gp, kWG, 0, s>>>(d_img, H, W, pad, lab, rev);
    mo_union<<<gp, kWG, 0, s>>>(H, W, pad, lab, err);
    mo_fill<<<gn, kWG, 0, s>>>(H, W, pad, lab, rev, fill);
    mo_line<true, true><<<gn, kWG, 0, s>>>(H, W, fill, t1);   // o1 = open(fill, 30x1)
    mo_line<true, false><<<gn, kWG, 0, s>>>(H, W, t1, t2);
    mo_line<false, true><<<gn, kWG, 0, s>>>(H, W, t2, t1);   // o2 = open(o1, 1x30)
    mo_line<false, false><<<gn, kWG, 0, s>>>(H, W, t1, fill);
    mo_eye_erode<<<gn, kWG, 0, s>>>(H, W, rev, fill, t1);     // thin = open(rev - o2, eye(5))
    mo_eye_dilate<<<gn, kWG, 0, s>>>(H, W, t1, d_thin);
    HIPCHECK(hipGetLastError());
    int herr = 0;
    HIPCHECK(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    if (herr) return tc::fail(TC_ERR_HIP, "tc_morph_device: union-find did not settle");
    return TC_OK;
} catch (...) {
    return tc::exception_status(__func__);
}

// getSdf with the morphology on the device: same points as tc_sdf_from_image
int tc::sdf_on_device(const unsigned char *d_img, int H, int W, std::vector<int> &pts, hipStream_t s)
{
    void *work = nullptr, *thin = nullptr;
    HIPCHECK(hipMalloc(&work, tc_morph_work_bytes(H, W)));
    hipError_t e = hipMalloc(&thin, (size_t)H * W);
    if (e != hipSuccess) {
        (void)hipFree(work);
        return tc::fail(TC_ERR_NOMEM, "sdf (device): %s", hipGetErrorString(e));
    }
    std::vector<uint8_t> h((size_t)H * W);
    int rc = tc_morph_device(d_img, H, W, (unsigned char *)thin, work, s);
    if (rc == TC_OK) {
        e = hipMemcpy(h.data(), thin, h.size(), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = tc::fail(TC_ERR_HIP, "sdf (device): %s", hipGetErrorString(e));
    }
    (void)hipFree(work);
    (void)hipFree(thin);
    if (rc != TC_OK) return rc;
    pts = tc::contour_points(h.data(), H, W);
    return TC_OK;
}
