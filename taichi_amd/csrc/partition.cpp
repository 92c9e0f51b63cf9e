// Morphology-based partition and hybrid-format build (host, g++):
// partition_shapes (TaiChi.cu:313-638) with cal_shape_size (:280-311),
// partition_dia_shape (:677-879) and cal_start_end (:881-916).
//
// Semantics are the reference's with its shapes processed serially in
// contour order (the OpenMP loop at TaiChi.cu:360 races on shared counters
// and on the claim markers). Differences in method, not in result:
//   * each shape scans only its own rows [x, x+side) (the reference scans
//     all M rows and filters on the same condition);
//   * claims live in a byte map; `val`, when given, is still updated with
//     FLT_MAX on claimed entries like the reference (:866);
//   * the merged CSR is written directly (zeros of a row first, ordered by
//     shape then diagonal, then the unclaimed entries in CSR order), which is
//     the order the reference's bucket + merge produces (:508-612).
// Arithmetic that decides shapes is kept in the reference's types: float
// gain and products (:283-287), int*double width tiers (:689-699), float
// density against double 0.6 (:756-757). Built with -ffp-contract=off.
#include "common.h"

#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "partition.h"

namespace {

constexpr int kThumbH = 512;     // morohutils.cu:191
constexpr int kMinShape = 1000;  // TaiChi.cu:379

template <class T>
T *dup_array(const std::vector<T> &v)
{
    T *p = (T *)malloc((v.empty() ? 1 : v.size()) * sizeof(T));
    if (p && !v.empty()) memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

}  // namespace

// diagonal band half-width of a square of this side (TaiChi.cu:689-699)
int tc::width_tier(int sd)
{
    if (sd >= 10000000) return (int)(sd * 0.000001);
    if (sd >= 1000000) return (int)(sd * 0.00001);
    if (sd >= 100000) return (int)(sd * 0.001);
    if (sd >= 10000) return (int)(sd * 0.01);
    return (int)(sd * 0.1);
}

// Candidate squares in contour order (TaiChi.cu:360-381): each contour's
// corners scaled by cal_shape_size, too-small ones dropped.
int tc::candidate_squares(int M, int N, const unsigned char *img, std::vector<Square> &out)
{
    int H, W, pM, pN;
    std::vector<int> pts;
    if (tc_thumb_dims(M, N, &H, &W, &pM, &pN) == 0) {
        if (!img) return tc::fail(TC_ERR_INVALID, "partition: thumbnail required for M > 512");
        pts = tc::sdf_from_image(img, H, W);
    }
    candidate_squares_from_pts(M, N, pts, out);
    return TC_OK;
}

void tc::candidate_squares_from_pts(int M, int N, const std::vector<int> &pts, std::vector<Square> &out)
{
    out.clear();
    const int num_shapes = (int)pts.size() / 4;
    for (int i = 0; i < num_shapes; i++) {
        int x1 = pts[4 * i], y1 = pts[4 * i + 1];
        int a1 = pts[4 * i + 2] - x1, b1 = pts[4 * i + 3] - y1;  // TaiChi.cu:368-371
        int x, y, a, b;
        float gain;
        tc_cal_shape_size(M, N, kThumbH, x1, y1, a1, b1, &x, &y, &a, &b, &gain);
        if (a < kMinShape || b < kMinShape) continue;
        out.push_back({x, y, b});
    }
}

// The density test of partition_dia_shape (TaiChi.cu:736-760) from the
// per-diagonal counts of unclaimed entries in the square's band: fills offs,
// st, en and nnz; dense_of[n] is the dense index of diagonal n, or -1.
void tc::decide_dense(int side, int w, const long long *cnt, DiaShape &sh, std::vector<int> &dense_of)
{
    const int nd = 2 * w - 1;
    dense_of.assign((size_t)(nd > 0 ? nd : 0), -1);
    for (int n = 0; n < nd; n++) {
        int off = n - w + 1, s, e;
        tc_cal_start_end(side, side, w, off, n, &s, &e);
        float density = (float)cnt[n] / (float)(e - s);
        if ((double)density >= 0.6) {
            dense_of[n] = (int)sh.offs.size();
            sh.offs.push_back(off);
            sh.st.push_back(s);
            sh.en.push_back(e);
            sh.nnz += cnt[n];
        }
    }
}

// Shape arrays and counters of the format (TaiChi.cu:423-481) from the kept
// shapes; csrRow, csrCol and zero_elements are left to the caller.
void tc::shape_arrays(const std::vector<DiaShape> &kept, int nnz, TaiChi_new *out)
{
    const int S = (int)kept.size();
    long long D = 0, dense = 0;
    for (auto &s : kept) {
        D += (long long)s.offs.size();
        dense += s.nnz;
    }
    out->dia_shapes = S;
    out->num_negs = (int)D;
    out->dense_nnz = (int)dense;
    out->sparse_nnz = (int)(nnz - dense);
    if (S == 0) return;
    std::vector<int> xs, ys, xe, ptr{0}, offs;
    for (auto &s : kept) {
        xs.push_back(s.xs);
        ys.push_back(s.ys);
        xe.push_back(s.xs + s.side);
        ptr.push_back(ptr.back() + (int)s.offs.size());
        offs.insert(offs.end(), s.offs.begin(), s.offs.end());
    }
    out->xStart = dup_array(xs);
    out->yStart = dup_array(ys);
    out->xStop = dup_array(xe);
    out->nDiasPtr = dup_array(ptr);
    out->neg_offsets = dup_array(offs);
    int lo = xs[0], hi = xe[0];
    for (int i = 1; i < S; i++) {
        lo = xs[i] < lo ? xs[i] : lo;
        hi = xe[i] > hi ? xe[i] : hi;
    }
    out->row_start = lo;
    out->row_stop = hi;
}

// One square shape [x, x+side) x [y, y+side) (partition_dia_shape,
// TaiChi.cu:677-879).
void tc::partition_square(const int *rowptr, const int *col, std::vector<uint8_t> &claimed, int x, int y,
                      int side, DiaShape &sh)
{
    sh.xs = x;
    sh.ys = y;
    sh.side = side;
    const int w = width_tier(side);
    sh.width = w;
    const int nd = 2 * w - 1;
    if (nd <= 0) return;
    std::vector<long long> cnt((size_t)nd, 0);
    // diagonal index of entry (r, c), or -1 when outside the square's band
    auto diag = [&](int r, int c) -> int {
        if (c < y || c >= y + side) return -1;
        int o = (c - y) - (r - x);
        if (o < -(w - 1) || o > w - 1) return -1;
        return o + w - 1;
    };
    for (int r = x; r < x + side; r++)
        for (int j = rowptr[r]; j < rowptr[r + 1]; j++) {
            if (claimed[j]) continue;
            int n = diag(r, col[j]);
            if (n >= 0) cnt[n]++;
        }
    std::vector<int> dense_of;
    decide_dense(side, w, cnt.data(), sh, dense_of);
    const std::vector<int> &st = sh.st, &en = sh.en;
    const int neg = (int)sh.offs.size();
    if (neg == 0) return;
    // rows holding a nonzero on each dense diagonal (ascending), claiming them
    std::vector<std::vector<int>> rows((size_t)neg);
    for (int d = 0; d < neg; d++) rows[d].reserve((size_t)cnt[sh.offs[d] + w - 1]);
    for (int r = x; r < x + side; r++)
        for (int j = rowptr[r]; j < rowptr[r + 1]; j++) {
            if (claimed[j]) continue;
            int n = diag(r, col[j]);
            if (n < 0 || dense_of[n] < 0) continue;
            rows[dense_of[n]].push_back(r - x);
            claimed[j] = 1;
        }
    // zero rows: the walk of TaiChi.cu:809-827 (a list entry is consumed by
    // the first row at or after it)
    sh.zeros.resize((size_t)neg);
    for (int d = 0; d < neg; d++) {
        const std::vector<int> &lst = rows[d];
        std::vector<int> &z = sh.zeros[d];
        size_t ri = 0;
        for (int r = st[d]; r < en[d]; r++) {
            if (ri < lst.size() && r >= lst[ri]) ri++;
            else z.push_back(r);
        }
    }
}

extern "C" {

// thumbnail sizing (morohutils.cu:191-200); the thumbnail kernel is in thumb.hip
int tc_thumb_dims(int M, int N, int *H, int *W, int *pM, int *pN)
{
    if (!H || !W || !pM || !pN) return tc::fail(TC_ERR_INVALID, "tc_thumb_dims: null argument");
    *H = kThumbH;
    *W = *pM = *pN = 0;
    if (M <= kThumbH) return 1;  // morohutils.cu:192-196
    *pM = M / kThumbH;
    const double scale = (double)kThumbH / (double)M;
    *W = (int)(scale * N);       // double product, truncated (:198-199)
    if (*W <= 0) return 2;       // the reference would divide by zero at :200
    *pN = N / *W;
    return 0;
}

void tc_cal_shape_size(int M, int N, int mF, int x1, int y1, int a1, int b1, int *x, int *y, int *a,
                       int *b, float *gain)
{
    float g;
    int X, Y, A, B;
    if (M > mF) {
        g = (float)M / (float)mF;
        X = (int)((float)x1 * g);
        Y = (int)((float)y1 * g);
        A = (int)((float)a1 * g);
        B = (int)((float)b1 * g);
    } else if (M < mF) {
        g = (float)mF / (float)M;
        X = (int)((float)x1 / g);
        Y = (int)((float)y1 / g);
        A = (int)((float)a1 / g);
        B = (int)((float)b1 / g);
    } else {
        g = 1.0f;
        X = x1;
        Y = y1;
        A = a1;
        B = b1;
    }
    if (X < 0) X = 0;
    if (X > M - 1) X = M - 1;
    if (Y < 0) Y = 0;
    if (Y > N - 1) Y = N - 1;
    if (Y + A > N) A = N - Y;
    if (X + B > M) B = M - X;
    if (A > B) A = B;
    else B = A;
    *x = X;
    *y = Y;
    *a = A;
    *b = B;
    *gain = g;
}

void tc_cal_start_end(int M_sub, int N_sub, int width, int off, int n, int *start, int *end)
{
    const int ao = off < 0 ? -off : off;
    if (M_sub >= N_sub) {
        if (n >= width - 1) {
            *start = 0;
            *end = N_sub - off;
        } else {
            *start = ao;
            *end = ao > M_sub - N_sub ? M_sub : N_sub + ao;
        }
    } else {
        if (n <= width - 1) {
            *start = ao;
            *end = M_sub;
        } else {
            *start = 0;
            *end = ao > N_sub - M_sub ? N_sub - ao : M_sub;
        }
    }
}

int tc_partition_from_image(int M, int N, int nnz, int *rowptr, int *col, float *val,
                            const unsigned char *img, TaiChi_new *out)
try {
    if (M < 0 || N < 0 || nnz < 0 || !rowptr || (nnz > 0 && !col) || !out)
        return tc::fail(TC_ERR_INVALID, "tc_partition_from_image: bad argument");
    if (rowptr[0] != 0 || rowptr[M] != nnz)
        return tc::fail(TC_ERR_INVALID, "tc_partition_from_image: rowptr[0]=%d rowptr[M]=%d nnz=%d", rowptr[0],
                        rowptr[M], nnz);
    memset(out, 0, sizeof *out);
    std::vector<uint8_t> claimed((size_t)nnz, 0);
    if (val)
        for (int j = 0; j < nnz; j++) {
            float d = val[j] - FLT_MAX;
            claimed[j] = !(std::fabs((double)d) > 1e-8);  // TaiChi.cu:730
        }
    std::vector<uint8_t> pre = claimed;

    std::vector<tc::Square> cand;
    int rc = tc::candidate_squares(M, N, img, cand);
    if (rc != TC_OK) return rc;
    std::vector<tc::DiaShape> kept;
    for (const tc::Square &q : cand) {
        tc::DiaShape sh;
        tc::partition_square(rowptr, col, claimed, q.x, q.y, q.side, sh);
        if (sh.nnz > 0) kept.push_back(std::move(sh));
    }

    // ---- format (TaiChi.cu:423-612) ----
    tc::shape_arrays(kept, nnz, out);
    const long long dense = out->dense_nnz;
    long long zeros = 0;
    for (auto &s : kept)
        for (auto &z : s.zeros) zeros += (long long)z.size();
    out->zero_elements = (int)zeros;
    if (dense == 0) {  // no dense diagonal: the input CSR is the format (:575-578)
        out->csrRow = rowptr;
        out->csrCol = col;
    } else {
        std::vector<int> cntr((size_t)M + 1, 0);
        for (auto &s : kept)
            for (auto &z : s.zeros)
                for (int r : z) cntr[s.xs + r]++;
        for (int r = 0; r < M; r++)
            for (int j = rowptr[r]; j < rowptr[r + 1]; j++)
                if (!claimed[j]) cntr[r]++;
        int *crow = (int *)malloc(((size_t)M + 1) * sizeof(int));
        if (!crow) return tc::fail(TC_ERR_NOMEM, "tc_partition_from_image: out of host memory");
        crow[0] = 0;
        for (int r = 0; r < M; r++) crow[r + 1] = crow[r] + cntr[r];
        const long long R = crow[M];
        int *ccol = R ? (int *)malloc((size_t)R * sizeof(int)) : nullptr;
        if (R && !ccol) {
            free(crow);
            return tc::fail(TC_ERR_NOMEM, "tc_partition_from_image: out of host memory");
        }
        std::vector<int> cur(crow, crow + M);
        for (auto &s : kept)
            for (size_t d = 0; d < s.offs.size(); d++)
                for (int r : s.zeros[d]) ccol[cur[s.xs + r]++] = (s.ys + r + s.offs[d]) << 1;
        for (int r = 0; r < M; r++)
            for (int j = rowptr[r]; j < rowptr[r + 1]; j++)
                if (!claimed[j]) ccol[cur[r]++] = (col[j] << 1) + 1;
        out->csrRow = crow;
        out->csrCol = ccol;
    }
    if (val)
        for (int j = 0; j < nnz; j++)
            if (claimed[j] && !pre[j]) val[j] = FLT_MAX;
    return TC_OK;
} catch (...) {
    return tc::exception_status(__func__);
}

TaiChi_new *tc_taichi_new_alloc(void) { return (TaiChi_new *)calloc(1, sizeof(TaiChi_new)); }

int tc_free_taichi_new(TaiChi_new *t)
{
    if (!t) return TC_OK;
    free(t->xStart);
    free(t->yStart);
    free(t->xStop);
    free(t->nDiasPtr);
    free(t->neg_offsets);
    if (t->dense_nnz > 0) {  // otherwise they alias the caller's CSR (TaiChi.cu:100-113)
        free(t->csrRow);
        free(t->csrCol);
    }
    free(t);
    return TC_OK;
}

}  // extern "C"
