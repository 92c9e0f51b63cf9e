// Host side shared by the SpMV translation units: the resident-matrix
// handle, the launch-configuration rules, and the launcher each kernel file
// provides (spmv_k1.hip, spmv_staged.hip, spmv_wide.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "common.h"
#include "spmv_tiles.h"
#include "taichi_amd.h"

#define HIPCHECK(expr)                                                                                      \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess) return tc::fail(TC_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));          \
    } while (0)

struct tc_matrix {
    int device = 0;
    tcd::MatView v{};
    void *mem = nullptr;
    size_t mem_bytes = 0;
    tc_matrix_stats st{};
    long long diag_terms = 0;  // sum over shapes of offsets x rows
    std::mutex mu;
    std::map<std::pair<int, int>, tc_config> cfg;
};

namespace tc {
// getSdf with the morphology on the device (morph_gpu.hip): the points
// tc_sdf_from_image gives for the same thumbnail, d_img in device memory
int sdf_on_device(const unsigned char *d_img, int H, int W, std::vector<int> &pts, hipStream_t s);
}  // namespace tc

// a captured launch sequence (tc_graph_capture, tc_pagerank_capture)
struct tc_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace tcd {

inline int elem_size(int dtype) { return dtype == TC_F32 ? 4 : 2; }

// supported compile-time column counts, largest first
constexpr int kKs[] = {64, 32, 16, 8, 7, 6, 5, 4, 3, 2, 1};

inline int pick_k(int k)
{
    for (int K : kKs)
        if (K <= k) return K;
    return 1;
}

// k > 1: the wide-rows kernel serves K in {2, 4, 8, 16, 32, 64} unless
// config flag bit 2 (kStagedK) asks for the staged tile kernel, which also
// serves every other K.
constexpr int kStagedK = 4;

inline bool wide_k(int K) { return K == 2 || K == 4 || K == 8 || K == 16 || K == 32 || K == 64; }

// rows of a wide-rows tile with RI rows per lane group (x elements of esize bytes)
inline int wide_rows(int K, int RI, int esize) { return (kWG / (K / wide_cpl(K, esize))) * RI; }

// A column count with no wide kernel of its own (3, 5, 6, 7) runs as one
// masked pass of this wide width (0: none) when X and Y rows split into
// aligned vectors of it: one pass over the indices instead of the staged
// kernel or two or three narrower chunks. (A single strided column measured
// slower masked than on the rows or staged kernels, so it is not masked.)
inline int masked_width(int K)
{
    if (K == 3) return 4;
    if (K >= 5 && K <= 7) return 8;
    return 0;
}

constexpr int kNoMask = 8;   // config flag (of the wide width's config): no masked chunks
constexpr int kQuadOff = 2;  // config flag: use the one-row-per-lane kernel at k = 1

// P from a rows_per_tile request for column count K (1, 2 or 4)
inline int p_for(int K, int rows_per_tile)
{
    int best = 2;
    for (int P : {1, 2, 4})
        if ((kWG * P) / K == rows_per_tile) best = P;
    return best;
}

int cu_count();
// the stored or default configuration of (K, dtype)
tc_config get_cfg(tc_matrix *A, int K, int dtype);
int check_spmv_args(const tc_matrix *A, const void *X, int ldx, int dtype, const float *Y, int ldy, int k);

// k = 1, one row per lane (spmv_k1.hip)
// per-workgroup LDS limit of the current device (cached per device)
size_t max_lds_bytes();

// X column: X[r * ldx], Y column: Y[r * ldy]
int launch_rows(const MatView &A, const void *X, int ldx, float *Y, int ldy, int dtype, const tc_config &c, int capv,
                hipStream_t s);
// k = 1, quad rows: whether they apply, and the launch (spmv_k1.hip)
bool quad_applies(const MatView &A, const void *X, const float *Y, int esize, const tc_config &c, int capv);
int launch_quad(const MatView &A, const void *X, float *Y, int dtype, const tc_config &c, int capv, hipStream_t s);
// staged tiles of P*256/K rows, any compiled K (spmv_staged.hip)
int launch_staged(const MatView &A, const void *X, int ldx, float *Y, int ldy, int dtype, int K, int P, int capv,
                  int grid, hipStream_t s);
// wide rows, K in {2, 4, 8, 16, 32, 64}: whether they apply, and the launch (spmv_wide.hip)
bool wide_applies(int K, const void *X, int ldx, const float *Y, int ldy, int esize, const tc_config &c);
// kv < K: only the first kv columns exist (a masked chunk, K <= 8)
int launch_wide(const MatView &A, const void *X, int ldx, float *Y, int ldy, int dtype, int K, const tc_config &c,
                hipStream_t s, int kv = 0);

// Host copy of a matrix's device layout (layout.hip)
struct Layout {
    int M = 0, N = 0, S = 0, D = 0, G = 0, L = 0, ow = 1;
    long long R = 0;      // entries kept in cols
    long long holes = 0;  // cancelled zero cells absorbed by presence bitmaps
    long long near = 0;   // kept entries whose column is within kNearRows of their row
    std::vector<ShapeDev> shapes;
    std::vector<uint32_t> pres;
    std::vector<int> gbase, glong, lrow, lptr, cols;
    std::vector<uint16_t> roff16;  // local offsets when ow == 2
    std::vector<uint8_t> roff8;    // local offsets when ow == 1
};
struct LayoutPart {
    const void *src;
    size_t bytes;
};
int plan_layout(int M, int N, const TaiChi_new *T, int long_threshold, Layout &P);
// device arrays in upload order: shapes, offs, gbase, roff, cols, glong,
// lrow, lptr, pres (offs: the format's neg_offsets)
std::vector<LayoutPart> layout_parts(const Layout &P, const int *offs);

}  // namespace tcd
