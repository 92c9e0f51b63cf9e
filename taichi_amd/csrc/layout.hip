// Device layout of a resident matrix (host side): from a host TaiChi_new to
// the arrays tc_matrix_create uploads (spmv_tiles.h describes them), and
// their sizes for tc_matrix_footprint.
//
// Replaces the upload half of taichi_SpMV (TaiChi.cu:1360-1383), which copies
// the TaiChi_new arrays as they are. Here the int rowptr becomes 64-row
// group bases plus 1- or 2-byte local offsets, rows longer than the long-row
// threshold move to a table of their own, and the sign-0 entries that cancel
// zero cells of a shape's dense diagonals (TaiChi.cu:500-520 writes them)
// become cleared bits of a per-row presence bitmap. A shape gets a bitmap
// (ceil(D_s/32) words a row) when more of its zero cells are cancelled than
// it has bitmap words, and also whenever one of its cancellations would be
// inexact in FP32 accumulation: that is every zero cell except one that is
// its row's only diagonal term on a row summed by one lane (there the term
// comes first and its sign-0 entry second, and x - x is exactly +0.0). The
// values y = A x stay the same; only the terms a kernel reads change.
#include <algorithm>
#include <mutex>
#include <vector>

#include "spmv_host.h"

namespace tcd {

namespace {

constexpr int kMaxPresShapes = 64;   // shapes considered for a bitmap
constexpr int kMaxPresOffsets = 64;  // offsets a shape with a bitmap may have

// offsets of one shape, sorted with their index in the shape: (offset, j)
struct ShapeOffs {
    int s;
    std::vector<std::pair<int, int>> offs;
};

// index j of the term of shape c at (row r, column col), or -1; for an offset
// stored twice, the first j whose bit is still set (bits == nullptr: the
// first j)
int match_term(const ShapeDev &h, const ShapeOffs &c, int r, int col, const uint32_t *pres)
{
    if (r < h.xs || r >= h.xe) return -1;
    const int shift = h.ys - h.xs;
    if (col < h.ys || col >= h.ys + (h.xe - h.xs)) return -1;
    const long long o = (long long)col - r - shift;
    if (o < h.omin || o > h.omax) return -1;
    auto it = std::lower_bound(c.offs.begin(), c.offs.end(), std::make_pair((int)o, -1));
    const int rows = h.xe - h.xs, i = r - h.xs;
    for (; it != c.offs.end() && it->first == (int)o; ++it) {
        const int j = it->second;
        if (!pres) return j;
        if (pres[h.pm + (j >> 5) * rows + i] >> (j & 31) & 1u) return j;
    }
    return -1;
}

}  // namespace

int plan_layout(int M, int N, const TaiChi_new *T, int long_threshold, Layout &P)
{
    if (!T || M < 0 || N < 0) return tc::fail(TC_ERR_INVALID, "matrix layout: bad argument");
    if (M > 0 && !T->csrRow) return tc::fail(TC_ERR_INVALID, "matrix layout: null csrRow");
    const int LT = long_threshold > 0 ? std::min(long_threshold, 1023) : tc::kDefaultLongRow;
    const int S = T->dia_shapes;
    const int R = M > 0 ? T->csrRow[M] : 0;
    if (S < 0 || R < 0 || (R > 0 && !T->csrCol)) return tc::fail(TC_ERR_INVALID, "matrix layout: bad format");
    if (S > 0 && (!T->xStart || !T->xStop || !T->yStart || !T->nDiasPtr || (T->nDiasPtr[S] > 0 && !T->neg_offsets)))
        return tc::fail(TC_ERR_INVALID, "matrix layout: null shape arrays");
    const bool encoded = T->dense_nnz > 0;
    if (!encoded && N >= (1 << 30)) return tc::fail(TC_ERR_UNSUPPORTED, "matrix layout: N >= 2^30");
    const int D = S > 0 ? T->nDiasPtr[S] : 0;
    for (int r = 0; r < M; r++)
        if (T->csrRow[r + 1] < T->csrRow[r])
            return tc::fail(TC_ERR_INVALID, "matrix layout: csrRow not monotone at %d", r);
    P = Layout();
    P.M = M;
    P.N = N;
    P.S = S;
    P.D = D;

    // shapes: validated so that every column the kernel forms lies in [0, N)
    P.shapes.resize((size_t)S);
    for (int s = 0; s < S; s++) {
        ShapeDev &h = P.shapes[s];
        h.xs = T->xStart[s];
        h.xe = T->xStop[s];
        h.ys = T->yStart[s];
        h.ob = T->nDiasPtr[s];
        h.oe = T->nDiasPtr[s + 1];
        if (h.xs < 0 || h.xe > M || h.xs > h.xe || h.ys < 0 || (long long)h.ys + (h.xe - h.xs) > N || h.ob < 0 ||
            h.oe < h.ob || h.oe > D)
            return tc::fail(TC_ERR_INVALID, "matrix layout: shape %d out of range", s);
        h.omin = 0;
        h.omax = 0;
        h.seq = 0;
        h.pad0 = h.pad1 = h.pad2 = 0;
        if (h.oe > h.ob) {
            h.omin = h.omax = T->neg_offsets[h.ob];
            h.seq = 1;
            for (int n = h.ob; n < h.oe; n++) {
                h.omin = std::min(h.omin, T->neg_offsets[n]);
                h.omax = std::max(h.omax, T->neg_offsets[n]);
                if ((long long)T->neg_offsets[n] != (long long)T->neg_offsets[h.ob] + (n - h.ob)) h.seq = 0;
            }
        }
        h.pm = -1;
    }

    // presence bitmaps: candidates, then the cancelled cells each would absorb
    std::vector<ShapeOffs> cand;
    if (encoded && S <= kMaxPresShapes)
        for (int s = 0; s < S; s++) {
            const ShapeDev &h = P.shapes[s];
            const int ds = h.oe - h.ob;
            if (ds < 1 || ds > kMaxPresOffsets || h.xe == h.xs) continue;
            ShapeOffs c;
            c.s = s;
            for (int j = 0; j < ds; j++) c.offs.push_back({T->neg_offsets[h.ob + j], j});
            std::sort(c.offs.begin(), c.offs.end());
            cand.push_back(std::move(c));
        }
    // diagonal terms a kernel adds for row r (the columns of every shape's
    // offsets that fall inside its square), counted up to 2
    auto diag_terms_of = [&](int r) {
        int n = 0;
        for (int s = 0; s < S && n < 2; s++) {
            const ShapeDev &h = P.shapes[s];
            if (r < h.xs || r >= h.xe) continue;
            const int side = h.xe - h.xs;
            for (int k = h.ob; k < h.oe && n < 2; k++) {
                const long long c = (long long)r - h.xs + h.ys + T->neg_offsets[k];
                n += c >= h.ys && c < (long long)h.ys + side;
            }
        }
        return n;
    };
    std::vector<ShapeOffs> pm_shapes;  // candidates that get a bitmap, in shape order
    if (!cand.empty()) {
        std::vector<long long> cnt(cand.size(), 0);
        std::vector<char> need(cand.size(), 0);  // a cancellation of the shape is inexact in FP32
        std::mutex mu;
        tc::parallel_for(M, [&](long long b, long long e) {
            std::vector<long long> lc(cand.size(), 0);
            std::vector<char> ln(cand.size(), 0);
            for (long long r = b; r < e; r++) {
                int inexact = -1;  // computed at the row's first sign-0 entry
                for (int j = T->csrRow[r]; j < T->csrRow[r + 1]; j++) {
                    const int raw = T->csrCol[j];
                    if (raw < 0 || (raw & 1)) continue;
                    for (size_t ci = 0; ci < cand.size(); ci++)
                        if (match_term(P.shapes[cand[ci].s], cand[ci], (int)r, raw >> 1, nullptr) >= 0) {
                            lc[ci]++;
                            if (inexact < 0)
                                inexact = T->csrRow[r + 1] - T->csrRow[r] > LT || diag_terms_of((int)r) >= 2;
                            ln[ci] |= (char)inexact;
                            break;
                        }
                }
            }
            std::lock_guard<std::mutex> g(mu);
            for (size_t ci = 0; ci < cand.size(); ci++) {
                cnt[ci] += lc[ci];
                need[ci] |= ln[ci];
            }
        });
        long long words = 0;
        for (size_t ci = 0; ci < cand.size(); ci++) {
            ShapeDev &h = P.shapes[cand[ci].s];
            const long long w = (long long)(h.xe - h.xs) * (((h.oe - h.ob) + 31) / 32);
            if ((cnt[ci] <= w && !need[ci]) || words + w > (1ll << 30)) continue;
            h.pm = (int)words;
            words += w;
            pm_shapes.push_back(cand[ci]);
        }
        P.pres.assign((size_t)words, 0u);
        for (const ShapeOffs &c : pm_shapes) {
            const ShapeDev &h = P.shapes[c.s];
            const int rows = h.xe - h.xs, ds = h.oe - h.ob;
            for (int w = 0; w * 32 < ds; w++) {
                const int nb = std::min(32, ds - 32 * w);
                const uint32_t full = nb == 32 ? 0xFFFFFFFFu : ((1u << nb) - 1u);
                std::fill(P.pres.begin() + h.pm + (size_t)w * rows, P.pres.begin() + h.pm + (size_t)(w + 1) * rows,
                          full);
            }
        }
    }

    // kept entries per row (a cancelled cell absorbed by a bitmap is dropped;
    // each row clears only its own bitmap words)
    std::vector<uint8_t> keep((size_t)R, 1);
    std::vector<int> len((size_t)M);
    tc::parallel_for(M, [&](long long b, long long e) {
        for (long long r = b; r < e; r++) {
            int n = 0;
            for (int j = T->csrRow[r]; j < T->csrRow[r + 1]; j++) {
                const int raw = T->csrCol[j];
                if (!pm_shapes.empty() && raw >= 0 && !(raw & 1)) {
                    for (const ShapeOffs &c : pm_shapes) {
                        const ShapeDev &h = P.shapes[c.s];
                        const int t = match_term(h, c, (int)r, raw >> 1, P.pres.data());
                        if (t < 0) continue;
                        P.pres[h.pm + (size_t)(t >> 5) * (h.xe - h.xs) + (r - h.xs)] &= ~(1u << (t & 31));
                        keep[j] = 0;
                        break;
                    }
                }
                n += keep[j];
            }
            len[r] = n;
        }
    });

    // rows: group offsets, long rows
    const int G = (M + 63) / 64;
    P.G = G;
    P.gbase.assign((size_t)G + 1, 0);
    P.roff16.resize((size_t)M);
    long long short_total = 0;
    int maxoff = 0;
    for (int r = 0; r < M; r++) {
        if ((r & 63) == 0) P.gbase[r >> 6] = (int)short_total;
        if (len[r] > LT) P.lrow.push_back(r);
        else short_total += len[r];
        const int o = (int)(short_total - P.gbase[r >> 6]);
        P.roff16[r] = (uint16_t)o;
        maxoff = std::max(maxoff, o);
    }
    P.gbase[G] = (int)short_total;
    P.L = (int)P.lrow.size();
    P.ow = maxoff <= 255 ? 1 : 2;
    if (P.L > 0) {
        P.glong.assign((size_t)G + 1, 0);
        for (int r : P.lrow) P.glong[(r >> 6) + 1]++;
        for (int g = 0; g < G; g++) P.glong[g + 1] += P.glong[g];
        P.lptr.resize((size_t)P.L + 1);
        P.lptr[0] = (int)short_total;
        for (int l = 0; l < P.L; l++) P.lptr[l + 1] = P.lptr[l] + len[P.lrow[l]];
    }
    if (P.ow == 1) {
        P.roff8.resize((size_t)M);
        for (int r = 0; r < M; r++) P.roff8[r] = (uint8_t)P.roff16[r];
        P.roff16.clear();
    }

    // encoded cols: short rows in row order, then long rows; checked against N
    long long kept = 0;
    for (int r = 0; r < M; r++) kept += len[r];
    P.R = kept;
    P.holes = (long long)R - kept;
    P.cols.resize((size_t)kept);
    std::vector<long long> dst((size_t)M);
    {
        long long sp = 0, lp = short_total;
        size_t li = 0;
        for (int r = 0; r < M; r++) {
            if (li < P.lrow.size() && P.lrow[li] == r) {
                dst[r] = lp;
                lp += len[r];
                li++;
            } else {
                dst[r] = sp;
                sp += len[r];
            }
        }
    }
    int bad = 0;
    std::mutex bm;
    tc::parallel_for(M, [&](long long b, long long e) {
        int lbad = 0;
        long long lnear = 0;
        for (long long r = b; r < e; r++) {
            long long o = dst[r];
            for (int j = T->csrRow[r]; j < T->csrRow[r + 1]; j++) {
                if (!keep[j]) continue;
                const int raw = T->csrCol[j];
                const int c = encoded ? raw >> 1 : raw;
                if (raw < 0 || c >= N) {
                    lbad = 1;
                    P.cols[o++] = 1;
                    continue;
                }
                lnear += c >= r - tcd::kNearRows && c <= r + tcd::kNearRows;
                P.cols[o++] = encoded ? raw : ((raw << 1) | 1);
            }
        }
        std::lock_guard<std::mutex> g(bm);
        P.near += lnear;
        bad |= lbad;
    });
    if (bad) return tc::fail(TC_ERR_INVALID, "matrix layout: column index out of range");
    return TC_OK;
}

// The device arrays of a layout in upload order (each 256-byte aligned on
// the device): shapes, offsets, group bases, local offsets, cols, long-row
// tables, presence bitmaps.
std::vector<LayoutPart> layout_parts(const Layout &P, const int *offs)
{
    return {
        {P.shapes.data(), P.shapes.size() * sizeof(ShapeDev)},
        {offs, (size_t)P.D * sizeof(int)},
        {P.gbase.data(), P.gbase.size() * sizeof(int)},
        {P.ow == 1 ? (const void *)P.roff8.data() : (const void *)P.roff16.data(), (size_t)P.M * P.ow},
        {P.cols.data(), P.cols.size() * sizeof(int)},
        {P.glong.data(), P.glong.size() * sizeof(int)},
        {P.lrow.data(), P.lrow.size() * sizeof(int)},
        {P.lptr.data(), P.lptr.size() * sizeof(int)},
        {P.pres.data(), P.pres.size() * sizeof(uint32_t)},
    };
}

}  // namespace tcd
