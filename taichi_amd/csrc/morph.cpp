// Thumbnail morphology and contour extraction of getSdf
// (morohutils.cu:231-282), host code. The reference calls OpenCV; this file
// implements the documented OpenCV semantics it relies on:
//   * 255 - mat and a - b saturate for 8-bit pixels;
//   * floodFill(seed (0,0), newVal 255) on the image padded by one 0 pixel,
//     4-connected, filling pixels equal to the seed value (imfill, :73-83);
//   * morphologyEx(MORPH_OPEN) = erode then dilate with the same element and
//     anchor (ksize/2), both reading src(x + x' - ax, y + y' - ay) (no
//     reflection for dilate); outside pixels read as 255 for erode and 0 for
//     dilate (the default morphology border);
//   * findContours(RETR_CCOMP, CHAIN_APPROX_NONE): Suzuki-Abe border
//     following on the zero-padded image, 8-connected foreground, every
//     border pixel listed, outer borders returned in reverse discovery order
//     each followed by its holes (reverse discovery order).
// OpenCV's version is unpinned by the reference (Makefile:11-12).
#include "common.h"
#include "partition.h"

#include <algorithm>
#include <cstring>
#include <deque>
#include <vector>

namespace {

using Img = std::vector<uint8_t>;

// Sliding-window min (erode) or max (dilate) along rows or columns over the
// window [i + lo, i + hi], with out-of-image pixels reading as `border`.
// Monotonic deque: O(1) amortized per pixel.
void window_1d(const Img &src, Img &dst, int H, int W, bool along_rows, int lo, int hi, bool take_min,
               uint8_t border)
{
    const int len = along_rows ? W : H, lines = along_rows ? H : W;
    std::vector<int> line((size_t)len), out((size_t)len);
    std::deque<int> dq;  // indices into the padded line
    for (int l = 0; l < lines; l++) {
        for (int i = 0; i < len; i++) line[i] = along_rows ? src[(size_t)l * W + i] : src[(size_t)i * W + l];
        auto at = [&](int j) -> int { return (j < 0 || j >= len) ? border : line[j]; };
        auto better = [&](int a, int b) { return take_min ? a <= b : a >= b; };
        dq.clear();
        int next = lo;  // next padded index to push
        for (int i = 0; i < len; i++) {
            while (next <= i + hi) {
                int v = at(next);
                while (!dq.empty() && better(v, at(dq.back()))) dq.pop_back();
                dq.push_back(next);
                next++;
            }
            while (dq.front() < i + lo) dq.pop_front();
            out[i] = at(dq.front());
        }
        for (int i = 0; i < len; i++) {
            if (along_rows) dst[(size_t)l * W + i] = (uint8_t)out[i];
            else dst[(size_t)i * W + l] = (uint8_t)out[i];
        }
    }
}

// Open with a 30-long line element: anchor 15 puts the window at [x-15, x+14].
void open_line30(const Img &src, Img &dst, int H, int W, bool horizontal)
{
    Img tmp(src.size());
    window_1d(src, tmp, H, W, horizontal, -15, 14, true, 255);
    window_1d(tmp, dst, H, W, horizontal, -15, 14, false, 0);
}

// Open with eye(5), anchor (2,2): the element is (d,d) for d in [-2, 2].
void open_eye5(const Img &src, Img &dst, int H, int W)
{
    Img tmp(src.size());
    for (int pass = 0; pass < 2; pass++) {
        const Img &in = pass == 0 ? src : tmp;
        Img &out = pass == 0 ? tmp : dst;
        const bool erode = pass == 0;
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                int m = erode ? 255 : 0;
                for (int d = -2; d <= 2; d++) {
                    int xx = x + d, yy = y + d;
                    int v = (xx < 0 || xx >= W || yy < 0 || yy >= H) ? (erode ? 255 : 0) : in[(size_t)yy * W + xx];
                    m = erode ? std::min(m, v) : std::max(m, v);
                }
                out[(size_t)y * W + x] = (uint8_t)m;
            }
    }
}

// imfill: holes of `src` (background not 4-connected to the padded border)
// become 255.
void fill_holes(const Img &src, Img &dst, int H, int W)
{
    const int Wp = W + 2, Hp = H + 2;
    std::vector<uint8_t> pad((size_t)Hp * Wp, 0), reach((size_t)Hp * Wp, 0);
    for (int y = 0; y < H; y++) memcpy(&pad[(size_t)(y + 1) * Wp + 1], &src[(size_t)y * W], W);
    const uint8_t seedv = pad[0];
    std::vector<int> st;
    st.push_back(0);
    reach[0] = 1;
    while (!st.empty()) {
        int p = st.back();
        st.pop_back();
        int y = p / Wp, x = p % Wp;
        int nb[4][2] = {{x - 1, y}, {x + 1, y}, {x, y - 1}, {x, y + 1}};
        for (auto &q : nb) {
            if (q[0] < 0 || q[0] >= Wp || q[1] < 0 || q[1] >= Hp) continue;
            int i = q[1] * Wp + q[0];
            if (!reach[i] && pad[i] == seedv) {
                reach[i] = 1;
                st.push_back(i);
            }
        }
    }
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            size_t i = (size_t)(y + 1) * Wp + x + 1;
            uint8_t cut = reach[i] ? 255 : pad[i];
            dst[(size_t)y * W + x] = (uint8_t)(src[(size_t)y * W + x] | (uint8_t)~cut);
        }
}

struct Contour {
    std::vector<int> xy;  // x,y pairs in image coordinates
    bool hole;
    int parent;           // outer contour id for holes under RETR_CCOMP
};

// Border following (Suzuki & Abe 1985, steps 3.1-3.5). Direction codes
// 0=E 1=NE 2=N 3=NW 4=W 5=SW 6=S 7=SE with y growing downwards. Pixel values:
// 0 background, 1 unvisited foreground, >=2 visited, <0 visited with a 0-pixel
// to its east examined by this border. owner[] records which contour marked
// a pixel, for the LNBD parent rule.
class Tracer {
  public:
    Tracer(int H, int W) : Wp(W + 2), Hp(H + 2), img((size_t)(H + 2) * (W + 2), 0), owner(img.size(), -1) {}
    int Wp, Hp;
    std::vector<int> img, owner;

    void follow(int x0, int y0, bool hole, int id, std::vector<int> &xy)
    {
        static const int dx[8] = {1, 1, 0, -1, -1, -1, 0, 1};
        static const int dy[8] = {0, -1, -1, -1, 0, 1, 1, 1};
        auto idx = [&](int x, int y) { return y * Wp + x; };
        // 3.1: clockwise from the 0-pixel that started the border
        int s0 = hole ? 0 : 4, s = s0;
        int x1 = x0, y1 = y0;
        bool found = false;
        for (int t = 1; t <= 8; t++) {
            s = (s0 - t) & 7;
            x1 = x0 + dx[s];
            y1 = y0 + dy[s];
            if (img[idx(x1, y1)] != 0) {
                found = true;
                break;
            }
            if (s == s0) break;
        }
        if (!found) {
            img[idx(x0, y0)] = -2;
            owner[idx(x0, y0)] = id;
            xy.push_back(x0 - 1);
            xy.push_back(y0 - 1);
            return;
        }
        int x3 = x0, y3 = y0;
        int back = s;  // direction from (x3,y3) to the previous border pixel
        for (;;) {
            int x4 = x3, y4 = y3, sf = back;
            for (int t = 1; t <= 8; t++) {  // 3.3: counter-clockwise
                sf = (back + t) & 7;
                x4 = x3 + dx[sf];
                y4 = y3 + dy[sf];
                if (img[idx(x4, y4)] != 0) break;
            }
            // 3.4: the east neighbour was examined iff the search wrapped
            // past direction 0, i.e. 1 <= sf <= back
            int &p3 = img[idx(x3, y3)];
            if (sf >= 1 && sf <= back) {
                p3 = -2;
                owner[idx(x3, y3)] = id;
            } else if (p3 == 1) {
                p3 = 2;
                owner[idx(x3, y3)] = id;
            }
            xy.push_back(x3 - 1);
            xy.push_back(y3 - 1);
            // 3.5
            if (x4 == x0 && y4 == y0 && x3 == x1 && y3 == y1) break;
            x3 = x4;
            y3 = y4;
            back = (sf + 4) & 7;
        }
    }
};

std::vector<Contour> find_contours_ccomp(const Img &bin, int H, int W)
{
    Tracer T(H, W);
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) T.img[(size_t)(y + 1) * T.Wp + x + 1] = bin[(size_t)y * W + x] ? 1 : 0;
    std::vector<Contour> cs;
    for (int y = 1; y < T.Hp - 1; y++) {
        int prev = 0, lnbd = 0;  // last border pixel met on this row (0 = frame)
        for (int x = 1; x < T.Wp - 1; x++) {
            int p = T.img[(size_t)y * T.Wp + x];
            if (p == prev) continue;
            bool outer = prev == 0 && p == 1;
            bool hole = !outer && p == 0 && prev >= 1;
            if (outer || hole) {
                int xs = hole ? x - 1 : x;
                if (hole && prev != 1) lnbd = x - 1;
                Contour c;
                c.hole = hole;
                c.parent = -1;
                if (hole && lnbd > 0) {
                    // Suzuki-Abe table 1 reduced to two levels: a hole's
                    // parent is the LNBD border if that is outer, else its parent
                    int b = T.owner[(size_t)y * T.Wp + lnbd];
                    c.parent = (b >= 0 && cs[b].hole) ? cs[b].parent : b;
                }
                T.follow(xs, y, hole, (int)cs.size(), c.xy);
                cs.push_back(std::move(c));
                lnbd = xs;
                prev = T.img[(size_t)y * T.Wp + x];
            } else {
                prev = p;
                if (prev != 0 && prev != 1) lnbd = x;
            }
        }
    }
    // output order of OpenCV's contour tree (head insertion, depth first)
    std::vector<std::vector<int>> holes_of(cs.size());
    for (int i = (int)cs.size() - 1; i >= 0; i--)
        if (cs[i].hole && cs[i].parent >= 0) holes_of[cs[i].parent].push_back(i);
    std::vector<Contour> out;
    out.reserve(cs.size());
    for (int i = (int)cs.size() - 1; i >= 0; i--) {
        if (cs[i].hole && cs[i].parent >= 0) continue;  // listed under its parent
        out.push_back(cs[i]);
        for (int h : holes_of[i]) out.push_back(cs[h]);
    }
    return out;
}

}  // namespace

namespace tc {

// getSdf's image pipeline (morohutils.cu:231-282): returns x1,y1,x2,y2 per
// kept contour (contour[0] and contour[size/2], size >= 80).
// getSdf's morphology (morohutils.cu:231-251): the binary image whose
// contours give the shapes.
void morph_thin(const uint8_t *mat, int H, int W, uint8_t *thin_out)
{
    const size_t n = (size_t)H * W;
    Img rev(n), fill(n), o1(n), o2(n), del(n), thin(n);
    for (size_t i = 0; i < n; i++) rev[i] = (uint8_t)(255 - mat[i]);
    fill_holes(rev, fill, H, W);
    open_line30(fill, o1, H, W, true);
    open_line30(o1, o2, H, W, false);
    for (size_t i = 0; i < n; i++) del[i] = rev[i] > o2[i] ? (uint8_t)(rev[i] - o2[i]) : 0;
    open_eye5(del, thin, H, W);
    memcpy(thin_out, thin.data(), n);
}

// x1,y1,x2,y2 of each contour of at least 80 points (morohutils.cu:138-151,
// 257-270): its first point and the one halfway along.
std::vector<int> contour_points(const uint8_t *thin, int H, int W)
{
    Img bin(thin, thin + (size_t)H * W);
    std::vector<Contour> cs = find_contours_ccomp(bin, H, W);
    std::vector<int> pts;
    for (const Contour &c : cs) {
        int len = (int)c.xy.size() / 2;
        if (len < 80) continue;
        pts.push_back(c.xy[0]);
        pts.push_back(c.xy[1]);
        pts.push_back(c.xy[2 * (len / 2)]);
        pts.push_back(c.xy[2 * (len / 2) + 1]);
    }
    return pts;
}

std::vector<int> sdf_from_image(const uint8_t *mat, int H, int W)
{
    Img thin((size_t)H * W);
    morph_thin(mat, H, W, thin.data());
    return contour_points(thin.data(), H, W);
}

// Contours alone, for tests: flattened as [n, (len, hole, x0,y0,...)...].
std::vector<int> contours_flat(const uint8_t *img, int H, int W)
{
    Img bin(img, img + (size_t)H * W);
    std::vector<Contour> cs = find_contours_ccomp(bin, H, W);
    std::vector<int> out{(int)cs.size()};
    for (const Contour &c : cs) {
        out.push_back((int)c.xy.size() / 2);
        out.push_back(c.hole ? 1 : 0);
        out.insert(out.end(), c.xy.begin(), c.xy.end());
    }
    return out;
}

}  // namespace tc

extern "C" int tc_sdf_from_image(const unsigned char *img, int H, int W, int *pts, int cap, int *num_shapes)
try {
    if (!img || H <= 0 || W <= 0 || !num_shapes || cap < 0 || (cap > 0 && !pts))
        return tc::fail(TC_ERR_INVALID, "tc_sdf_from_image: bad argument");
    std::vector<int> v = tc::sdf_from_image(img, H, W);
    int k = (int)v.size() / 4;
    *num_shapes = k;
    if (k > cap) return tc::fail(TC_ERR_CAPACITY, "tc_sdf_from_image: %d shapes > cap %d", k, cap);
    if (k) memcpy(pts, v.data(), v.size() * sizeof(int));
    return TC_OK;
} catch (...) {
    return tc::exception_status(__func__);
}

extern "C" int tc_morph_from_image(const unsigned char *img, int H, int W, unsigned char *thin)
try {
    if (!img || !thin || H <= 0 || W <= 0) return tc::fail(TC_ERR_INVALID, "tc_morph_from_image: bad argument");
    tc::morph_thin(img, H, W, thin);
    return TC_OK;
} catch (...) {
    return tc::exception_status(__func__);
}

extern "C" int tc_contours(const unsigned char *img, int H, int W, int *out, int cap, int *needed)
try {
    if (!img || H <= 0 || W <= 0 || !needed || cap < 0 || (cap > 0 && !out))
        return tc::fail(TC_ERR_INVALID, "tc_contours: bad argument");
    std::vector<int> v = tc::contours_flat(img, H, W);
    *needed = (int)v.size();
    if ((int)v.size() > cap) return tc::fail(TC_ERR_CAPACITY, "tc_contours: need %zu ints", v.size());
    memcpy(out, v.data(), v.size() * sizeof(int));
    return TC_OK;
} catch (...) {
    return tc::exception_status(__func__);
}
