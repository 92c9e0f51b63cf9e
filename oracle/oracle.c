/*
 * oracle.c -- CPU restatement of TaiChi's binary-SpMV path (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this file's library. It is the checker, never the product: the HIP library
 * in taichi_amd/ neither links nor calls it.
 *
 * The restatement is deliberately literal and slow (two full CSR passes per
 * shape, brute-force morphology) so that it reads line-by-line against the
 * reference. Build with -ffp-contract=off: the partition's float arithmetic
 * (gain, density) must round exactly like the reference's.
 *
 * Parity status per function (how each is pinned is in tests/):
 *   orc_read_mtx ............ fixtures + reference mmio.cpp (oracle/_ref)
 *   orc_thumb_dims/thumbnail  closed form + brute force
 *   orc_imfill/erode/dilate . hand-derived images (OpenCV semantics)
 *   orc_find_contours ....... hand-derived images; parity unpinned vs OpenCV
 *   orc_sdf_from_image ...... hand-derived images; parity unpinned vs OpenCV
 *   orc_cal_* / width ....... table tests at every tier edge
 *   orc_partition_shapes .... format invariants + interpreter == CSR y
 *   orc_csr_spmv_* .......... dense brute force / numpy
 *   orc_taichi_spmv_* ....... equality with CSR y (bit-exact, integer x)
 */
#include "oracle.h"

#include <ctype.h>
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_THUMB_H 512            /* morohutils.cu:191 */
#define ORC_ZERO 1e-8              /* TaiChi.h:76 */
#define ORC_MIN_CONTOUR 80         /* morohutils.cu:140 */
#define ORC_MIN_SHAPE 1000         /* TaiChi.cu:379 */

void orc_free(void *p) { free(p); }

/* ===================================================================== */
/* a1: Matrix Market load                                                 */
/* ===================================================================== */

/* mm_read_banner (mmio.cpp:96-179). typecode: [0]'M' [1]'C'/'A'
   [2]'R'/'C'/'P'/'I' [3]'G'/'S'/'H'/'K'. Returns 0 or an MM_* code. */
static int orc_banner(FILE *f, char tc[4])
{
    char line[1025];
    char banner[64], mtx[64], crd[64], dtype[64], storage[64];
    tc[0] = tc[1] = tc[2] = ' ';
    tc[3] = 'G';
    if (fgets(line, sizeof line, f) == NULL)
        return 12; /* MM_PREMATURE_EOF */
    if (sscanf(line, "%63s %63s %63s %63s %63s", banner, mtx, crd, dtype, storage) != 5)
        return 12;
    for (char *p = mtx; *p; ++p) *p = (char)tolower((unsigned char)*p);
    for (char *p = crd; *p; ++p) *p = (char)tolower((unsigned char)*p);
    for (char *p = dtype; *p; ++p) *p = (char)tolower((unsigned char)*p);
    for (char *p = storage; *p; ++p) *p = (char)tolower((unsigned char)*p);
    if (strncmp(banner, "%%MatrixMarket", 14) != 0)
        return 14; /* MM_NO_HEADER */
    if (strcmp(mtx, "matrix") != 0)
        return 15; /* MM_UNSUPPORTED_TYPE */
    tc[0] = 'M';
    if (strcmp(crd, "coordinate") == 0) tc[1] = 'C';
    else if (strcmp(crd, "array") == 0) tc[1] = 'A';
    else return 15;
    if (strcmp(dtype, "real") == 0) tc[2] = 'R';
    else if (strcmp(dtype, "complex") == 0) tc[2] = 'C';
    else if (strcmp(dtype, "pattern") == 0) tc[2] = 'P';
    else if (strcmp(dtype, "integer") == 0) tc[2] = 'I';
    else return 15;
    if (strcmp(storage, "general") == 0) tc[3] = 'G';
    else if (strcmp(storage, "symmetric") == 0) tc[3] = 'S';
    else if (strcmp(storage, "hermitian") == 0) tc[3] = 'H';
    else if (strcmp(storage, "skew-symmetric") == 0) tc[3] = 'K';
    else return 15;
    return 0;
}

int orc_mm_read_banner_file(const char *path, char *typecode4)
{
    FILE *f = fopen(path, "r");
    if (!f) return 11; /* MM_COULD_NOT_READ_FILE */
    int rc = orc_banner(f, typecode4);
    fclose(f);
    return rc;
}

/* mm_read_mtx_crd_size (mmio.cpp:189-217): skip '%' lines, then "M N nz". */
static int orc_crd_size(FILE *f, int *M, int *N, unsigned long *nz)
{
    char line[1025];
    long long nzl = 0;
    *M = *N = 0;
    *nz = 0;
    do {
        if (fgets(line, sizeof line, f) == NULL)
            return 12;
    } while (line[0] == '%');
    if (sscanf(line, "%d %d %lld", M, N, &nzl) == 3) {
        *nz = (unsigned long)nzl;
        return 0;
    }
    /* The reference loops until three numbers are read (mmio.cpp:210-214);
       a matching failure leaves the stream where it is, so that loop never
       ends on malformed input. Stop there instead. */
    int got;
    do {
        got = fscanf(f, "%d %d %lld", M, N, &nzl);
        if (got == EOF || got == 0) return 12;
    } while (got != 3);
    *nz = (unsigned long)nzl;
    return 0;
}

/* readMtx (TaiChi.cu:120-278). Deviations, both undefined behaviour in the
   reference and rejected here with -4: an 'array' (dense) file, whose size
   line the reference misreads, and a data section that ends early or holds an
   index outside [1,m] x [1,n]. */
int orc_read_mtx(const char *path, int *m, int *n, unsigned long *nnzA,
                 int **rowptr, int **colidx, float **val)
{
    char tc[4];
    FILE *f = fopen(path, "r");
    if (f == NULL) return -1;                       /* :130-131 */
    if (orc_banner(f, tc) != 0) { fclose(f); return -2; }   /* :133-136 */
    if (tc[2] == 'C') { fclose(f); return -3; }             /* :138-141 */
    int isPattern = tc[2] == 'P', isReal = tc[2] == 'R', isInteger = tc[2] == 'I';
    unsigned long nrep = 0;
    if (tc[1] != 'C') { fclose(f); return -4; }  /* 'array': see the note above */
    if (orc_crd_size(f, m, n, &nrep) != 0) { fclose(f); return -4; }  /* :158-160 */
    if (*m < 0 || *n < 0) { fclose(f); return -4; }
    int isSym = (tc[3] == 'S' || tc[3] == 'H');     /* :162-165 */

    int M = *m;
    int *cnt = (int *)calloc((size_t)M + 1, sizeof(int));
    int *ri = (int *)malloc((nrep ? nrep : 1) * sizeof(int));
    int *ci = (int *)malloc((nrep ? nrep : 1) * sizeof(int));
    float *vi = (float *)malloc((nrep ? nrep : 1) * sizeof(float));
    for (unsigned long i = 0; i < nrep; i++) {      /* :180-205 */
        int a = 0, b = 0, iv = 0, got = 0;
        double fv = 0.0;
        if (isReal) got = fscanf(f, "%d %d %lg\n", &a, &b, &fv) == 3;
        else if (isInteger) { got = fscanf(f, "%d %d %d\n", &a, &b, &iv) == 3; fv = iv; }
        else if (isPattern) { got = fscanf(f, "%d %d\n", &a, &b) == 2; fv = 1.0; }
        a--; b--;
        if (!got || a < 0 || a >= M || b < 0 || b >= *n) {
            free(cnt); free(ri); free(ci); free(vi); fclose(f);
            return -4;
        }
        cnt[a]++;
        ri[i] = a; ci[i] = b; vi[i] = (float)fv;
    }
    fclose(f);
    if (isSym)                                       /* :210-216 */
        for (unsigned long i = 0; i < nrep; i++)
            if (ri[i] != ci[i]) cnt[ci[i]]++;
    /* exclusive scan :218-228 */
    int old = cnt[0];
    cnt[0] = 0;
    for (int i = 1; i <= M; i++) {
        int nw = cnt[i];
        cnt[i] = old + cnt[i - 1];
        old = nw;
    }
    int total = cnt[M];
    int *rp = (int *)malloc(((size_t)M + 1) * sizeof(int));
    memcpy(rp, cnt, ((size_t)M + 1) * sizeof(int));
    memset(cnt, 0, ((size_t)M + 1) * sizeof(int));
    int *cc = (int *)calloc(total ? (size_t)total : 1, sizeof(int));
    float *vv = (float *)calloc(total ? (size_t)total : 1, sizeof(float));
    for (unsigned long i = 0; i < nrep; i++) {      /* :240-270 */
        int r = ri[i], c = ci[i];
        int o = rp[r] + cnt[r]++;
        cc[o] = c; vv[o] = vi[i];
        if (isSym && r != c) {
            o = rp[c] + cnt[c]++;
            cc[o] = r; vv[o] = vi[i];
        }
    }
    free(cnt); free(ri); free(ci); free(vi);
    *nnzA = (unsigned long)total;
    *rowptr = rp; *colidx = cc; *val = vv;
    return 0;
}

/* ===================================================================== */
/* a2: thumbnail                                                          */
/* ===================================================================== */

/* morohutils.cu:191-200 */
int orc_thumb_dims(int M, int N, int *H, int *W, int *pM, int *pN)
{
    *H = ORC_THUMB_H; *W = 0; *pM = 0; *pN = 0;
    if (M <= ORC_THUMB_H) return 1;
    *pM = M / ORC_THUMB_H;
    double scale = (double)ORC_THUMB_H / (double)M;
    *W = (int)(scale * N);
    if (*W <= 0) return 2;
    *pN = N / *W;
    return 0;
}

/* getSdf_GPU1 (morohutils.cu:153-180) as a closed form: a pixel is 0 iff
   some nonzero (i,c) has i/pM < 512 and c/pN < W landing on it. */
void orc_thumbnail(int M, int N, const int *rowptr, const int *col, uint8_t *img)
{
    int H, W, pM, pN;
    if (orc_thumb_dims(M, N, &H, &W, &pM, &pN) != 0) return;
    memset(img, 255, (size_t)H * W);
    for (int i = 0; i < M; i++)
        for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
            int x = i / pM, y = col[k] / pN;
            if (x < H && y < W) img[(size_t)x * W + y] = 0;
        }
}

/* ===================================================================== */
/* a3: morphology with OpenCV semantics                                   */
/* ===================================================================== */

/* imfill (morohutils.cu:73-83): pad by one zero pixel, floodFill from the
   corner with 4-connectivity over pixels equal to the seed value (0), then
   dst = src | ~interior. */
void orc_imfill(const uint8_t *src, uint8_t *dst, int H, int W)
{
    int Hp = H + 2, Wp = W + 2;
    uint8_t *t = (uint8_t *)calloc((size_t)Hp * Wp, 1);
    for (int y = 0; y < H; y++) memcpy(t + (size_t)(y + 1) * Wp + 1, src + (size_t)y * W, W);
    int *stack = (int *)malloc((size_t)Hp * Wp * sizeof(int));
    int sp = 0;
    uint8_t seed = t[0];
    if (seed != 255) {
        t[0] = 255;
        stack[sp++] = 0;
        while (sp) {
            int p = stack[--sp];
            int y = p / Wp, x = p % Wp;
            const int nx[4] = {x - 1, x + 1, x, x};
            const int ny[4] = {y, y, y - 1, y + 1};
            for (int k = 0; k < 4; k++) {
                if (nx[k] < 0 || nx[k] >= Wp || ny[k] < 0 || ny[k] >= Hp) continue;
                int q = ny[k] * Wp + nx[k];
                if (t[q] == seed) { t[q] = 255; stack[sp++] = q; }
            }
        }
    }
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            uint8_t cut = t[(size_t)(y + 1) * Wp + x + 1];
            dst[(size_t)y * W + x] = (uint8_t)(src[(size_t)y * W + x] | (uint8_t)~cut);
        }
    free(stack);
    free(t);
}

/* erode: dst(x,y) = min over kernel elements of src(x+dx, y+dy), pixels
   outside the image read as 255 (OpenCV's default morphology border). */
void orc_erode(const uint8_t *src, uint8_t *dst, int H, int W,
               const int *dx, const int *dy, int nk)
{
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            int m = 255;
            for (int k = 0; k < nk; k++) {
                int xx = x + dx[k], yy = y + dy[k];
                int v = (xx < 0 || xx >= W || yy < 0 || yy >= H) ? 255 : src[(size_t)yy * W + xx];
                if (v < m) m = v;
            }
            dst[(size_t)y * W + x] = (uint8_t)m;
        }
}

/* dilate: max over the SAME (unreflected) offsets; outside reads as 0. */
void orc_dilate(const uint8_t *src, uint8_t *dst, int H, int W,
                const int *dx, const int *dy, int nk)
{
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            int m = 0;
            for (int k = 0; k < nk; k++) {
                int xx = x + dx[k], yy = y + dy[k];
                int v = (xx < 0 || xx >= W || yy < 0 || yy >= H) ? 0 : src[(size_t)yy * W + xx];
                if (v > m) m = v;
            }
            dst[(size_t)y * W + x] = (uint8_t)m;
        }
}

/* structuring elements of morohutils.cu:239, :242, :250 with their default
   anchors (ksize/2): Size(30,1) -> (15,0), Size(1,30) -> (0,15), eye(5) -> (2,2). */
static int orc_kernel(int which, int *dx, int *dy)
{
    int n = 0;
    if (which == 0) for (int i = 0; i < 30; i++) { dx[n] = i - 15; dy[n] = 0; n++; }
    else if (which == 1) for (int i = 0; i < 30; i++) { dx[n] = 0; dy[n] = i - 15; n++; }
    else for (int i = 0; i < 5; i++) { dx[n] = i - 2; dy[n] = i - 2; n++; }
    return n;
}

/* morphologyEx(MORPH_OPEN) = erode then dilate with the same element. */
void orc_open(const uint8_t *src, uint8_t *dst, int H, int W, int which)
{
    int dx[30], dy[30];
    int nk = orc_kernel(which, dx, dy);
    uint8_t *tmp = (uint8_t *)malloc((size_t)H * W);
    orc_erode(src, tmp, H, W, dx, dy, nk);
    orc_dilate(tmp, dst, H, W, dx, dy, nk);
    free(tmp);
}

void orc_sub_sat(const uint8_t *a, const uint8_t *b, uint8_t *dst, int n)
{
    for (int i = 0; i < n; i++) dst[i] = a[i] > b[i] ? (uint8_t)(a[i] - b[i]) : 0;
}

/* ===================================================================== */
/* a3: findContours(RETR_CCOMP, CHAIN_APPROX_NONE)                        */
/* Suzuki & Abe (1985) border following as OpenCV applies it: the image  */
/* is zero-padded by one pixel, nonzero = 1, 8-connected foreground.      */
/* Direction codes: 0=E 1=NE 2=N 3=NW 4=W 5=SW 6=S 7=SE (y grows down).  */
/* ===================================================================== */

typedef struct { int *v; int n, cap; } ivec;
static void iv_push(ivec *a, int x)
{
    if (a->n == a->cap) { a->cap = a->cap ? 2 * a->cap : 64; a->v = (int *)realloc(a->v, (size_t)a->cap * sizeof(int)); }
    a->v[a->n++] = x;
}

static const int DX8[8] = {1, 1, 0, -1, -1, -1, 0, 1};
static const int DY8[8] = {0, -1, -1, -1, 0, 1, 1, 1};

/* Follow one border starting at pixel (x0,y) of the padded image P.
   Marks: 2 = visited, -2 = visited and its east neighbour is a 0-pixel
   examined by this border (Suzuki-Abe step 3.4). Appends the points in
   original coordinates. */
static void orc_follow(int *P, int Wp, int x0, int y0, int hole, ivec *pts)
{
    int i0 = y0 * Wp + x0;
    int d[8];
    for (int k = 0; k < 8; k++) d[k] = DY8[k] * Wp + DX8[k];
    /* clockwise search for the first nonzero neighbour, starting just after
       the 0-pixel that triggered the border: W for outer, E for holes. */
    int s = hole ? 0 : 4, s_end = s, i1;
    do {
        s = (s - 1) & 7;
        i1 = i0 + d[s];
    } while (P[i1] == 0 && s != s_end);
    if (s == s_end) {               /* isolated pixel */
        P[i0] = -2;
        iv_push(pts, x0 - 1); iv_push(pts, y0 - 1);
        return;
    }
    int i3 = i0, x3 = x0, y3 = y0;
    for (;;) {
        s_end = s;                  /* direction from i3 back to the previous pixel */
        int i4 = i3, t;
        for (t = 1; t <= 8; t++) {  /* counter-clockwise from the next direction */
            s = (s_end + t) & 7;
            i4 = i3 + d[s];
            if (P[i4] != 0) break;
        }
        if ((unsigned)(s - 1) < (unsigned)s_end) P[i3] = -2;   /* east examined, 0 */
        else if (P[i3] == 1) P[i3] = 2;
        iv_push(pts, x3 - 1); iv_push(pts, y3 - 1);
        if (i4 == i0 && i3 == i1) break;
        i3 = i4; x3 += DX8[s]; y3 += DY8[s];
        s = (s + 4) & 7;
    }
}

int orc_find_contours(const uint8_t *img, int H, int W, int **out_pts, int **out_ptr, int **out_hole)
{
    int Hp = H + 2, Wp = W + 2;
    int *P = (int *)calloc((size_t)Hp * Wp, sizeof(int));
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) P[(y + 1) * Wp + x + 1] = img[(size_t)y * W + x] != 0;
    /* 8-connected component labels: a hole's parent under RETR_CCOMP is the
       outer border of the component holding the pixel left of the hole. */
    int *lab = (int *)malloc((size_t)Hp * Wp * sizeof(int));
    for (int i = 0; i < Hp * Wp; i++) lab[i] = -1;
    int *stack = (int *)malloc((size_t)Hp * Wp * sizeof(int));
    int ncomp = 0;
    for (int i = 0; i < Hp * Wp; i++) {
        if (!P[i] || lab[i] >= 0) continue;
        int sp = 0;
        stack[sp++] = i; lab[i] = ncomp;
        while (sp) {
            int p = stack[--sp], y = p / Wp, x = p % Wp;
            for (int k = 0; k < 8; k++) {
                int xx = x + DX8[k], yy = y + DY8[k];
                if (xx < 0 || xx >= Wp || yy < 0 || yy >= Hp) continue;
                int q = yy * Wp + xx;
                if (P[q] && lab[q] < 0) { lab[q] = ncomp; stack[sp++] = q; }
            }
        }
        ncomp++;
    }
    free(stack);
    int *comp_outer = (int *)malloc((size_t)(ncomp ? ncomp : 1) * sizeof(int));
    for (int c = 0; c < ncomp; c++) comp_outer[c] = -1;

    ivec pts = {0}, beg = {0}, holes = {0}, parent = {0};
    for (int y = 1; y < Hp - 1; y++) {
        int prev = 0;
        for (int x = 1; x < Wp - 1; x++) {
            int p = P[y * Wp + x];
            if (p == prev) continue;
            int outer = (prev == 0 && p == 1);
            int hole = !outer && p == 0 && prev >= 1;
            if (outer || hole) {
                int xs = hole ? x - 1 : x;
                int c = lab[y * Wp + xs];
                iv_push(&beg, pts.n / 2);
                iv_push(&holes, hole);
                iv_push(&parent, hole ? comp_outer[c] : -1);
                if (!hole) comp_outer[c] = beg.n - 1;
                orc_follow(P, Wp, xs, y, hole, &pts);
                prev = P[y * Wp + x];
            } else {
                prev = p;
            }
        }
    }
    int nc = beg.n;
    iv_push(&beg, pts.n / 2);
    /* output order: tree built by head insertion, read depth-first */
    int *order = (int *)malloc((size_t)(nc ? nc : 1) * sizeof(int));
    int no = 0;
    for (int c = nc - 1; c >= 0; c--) {
        if (holes.v[c]) continue;
        order[no++] = c;
        for (int h = nc - 1; h > c; h--)
            if (holes.v[h] && parent.v[h] == c) order[no++] = h;
    }
    int *optr = (int *)malloc(((size_t)nc + 1) * sizeof(int));
    int *ohole = (int *)malloc((size_t)(nc ? nc : 1) * sizeof(int));
    int *opts = (int *)malloc((size_t)(pts.n ? pts.n : 2) * sizeof(int));
    int w = 0;
    optr[0] = 0;
    for (int i = 0; i < no; i++) {
        int c = order[i];
        int len = beg.v[c + 1] - beg.v[c];
        memcpy(opts + 2 * (size_t)w, pts.v + 2 * (size_t)beg.v[c], 2 * (size_t)len * sizeof(int));
        w += len;
        optr[i + 1] = w;
        ohole[i] = holes.v[c];
    }
    free(order); free(pts.v); free(beg.v); free(holes.v); free(parent.v);
    free(comp_outer); free(lab); free(P);
    *out_pts = opts; *out_ptr = optr; *out_hole = ohole;
    return no;
}

/* getSdf image part (morohutils.cu:231-282). */
int orc_sdf_from_image(const uint8_t *mat, int H, int W, int **out)
{
    size_t n = (size_t)H * W;
    uint8_t *rev = (uint8_t *)malloc(n), *fill = (uint8_t *)malloc(n);
    uint8_t *o1 = (uint8_t *)malloc(n), *o2 = (uint8_t *)malloc(n);
    uint8_t *del = (uint8_t *)malloc(n), *thin = (uint8_t *)malloc(n);
    for (size_t i = 0; i < n; i++) rev[i] = (uint8_t)(255 - mat[i]);  /* :232 */
    orc_imfill(rev, fill, H, W);                                      /* :235 */
    orc_open(fill, o1, H, W, 0);                                      /* :239-240 */
    orc_open(o1, o2, H, W, 1);                                        /* :242-243 */
    orc_sub_sat(rev, o2, del, (int)n);                                /* :245 */
    orc_open(del, thin, H, W, 2);                                     /* :250-251 */
    int *pts, *ptr, *hole;
    int nc = orc_find_contours(thin, H, W, &pts, &ptr, &hole);        /* :257-258 */
    int *res = (int *)malloc(((size_t)nc * 4 + 4) * sizeof(int));
    int k = 0;
    for (int c = 0; c < nc; c++) {
        int len = ptr[c + 1] - ptr[c];
        if (len < ORC_MIN_CONTOUR) continue;                          /* :138-151 */
        const int *cp = pts + 2 * (size_t)ptr[c];
        res[4 * k + 0] = cp[0];                                       /* :269 */
        res[4 * k + 1] = cp[1];
        res[4 * k + 2] = cp[2 * (len / 2)];                           /* :270 */
        res[4 * k + 3] = cp[2 * (len / 2) + 1];
        k++;
    }
    free(pts); free(ptr); free(hole);
    free(rev); free(fill); free(o1); free(o2); free(del); free(thin);
    *out = res;
    return k;
}

/* ===================================================================== */
/* a4/a5: shape sizing and diagonal partition                             */
/* ===================================================================== */

/* cal_shape_size (TaiChi.cu:280-311), float arithmetic as written. */
void orc_cal_shape_size(int M, int N, int mF, int x1, int y1, int a1, int b1,
                        int *x, int *y, int *a, int *b, float *gain)
{
    if (M > mF) {
        *gain = (float)M / (float)mF;
        *x = (int)((float)x1 * *gain); *y = (int)((float)y1 * *gain);
        *a = (int)((float)a1 * *gain); *b = (int)((float)b1 * *gain);
    } else if (M < mF) {
        *gain = (float)mF / (float)M;
        *x = (int)((float)x1 / *gain); *y = (int)((float)y1 / *gain);
        *a = (int)((float)a1 / *gain); *b = (int)((float)b1 / *gain);
    } else {
        *gain = 1.0f;
        *x = x1; *y = y1; *a = a1; *b = b1;
    }
    if (*x < 0) *x = 0;
    if (*x > M - 1) *x = M - 1;
    if (*y < 0) *y = 0;
    if (*y > N - 1) *y = N - 1;
    if (*y + *a > N) *a = N - *y;
    if (*x + *b > M) *b = M - *x;
    if (*a > *b) *a = *b; else *b = *a;
}

/* width tiers (TaiChi.cu:689-699): int * double, truncated. */
int orc_width_for(int sd)
{
    if (sd >= 10000000) return (int)(sd * 0.000001);
    if (sd >= 1000000) return (int)(sd * 0.00001);
    if (sd >= 100000) return (int)(sd * 0.001);
    if (sd >= 10000) return (int)(sd * 0.01);
    return (int)(sd * 0.1);
}

/* cal_start_end (TaiChi.cu:881-916) */
void orc_cal_start_end(int M_sub, int N_sub, int width, int off, int n, int *s, int *e)
{
    int ao = off < 0 ? -off : off;
    if (M_sub >= N_sub) {
        if (n >= width - 1) { *s = 0; *e = N_sub - off; }
        else if (ao > M_sub - N_sub) { *s = ao; *e = M_sub; }
        else { *s = ao; *e = N_sub + ao; }
    } else {
        if (n <= width - 1) { *s = ao; *e = M_sub; }
        else if (ao > N_sub - M_sub) { *s = 0; *e = N_sub - ao; }
        else { *s = 0; *e = M_sub; }
    }
}

static int orc_unclaimed(const float *val, int j)
{
    float dlt = val[j] - FLT_MAX;
    return fabs((double)dlt) > ORC_ZERO;             /* TaiChi.cu:730, 859, 560 */
}

typedef struct {
    int M_sub, N_sub, xStart, yStart, width, neg, nnz;
    int *neg_offsets, *numZeroNeg, **rowZeroNeg;
} orc_shape;

/* partition_dia_shape (TaiChi.cu:677-879) for the square [x,x+b) x [y,y+a). */
static void orc_partition_dia(int M, const int *rowptr, const int *col, float *val,
                              int x, int y, int a, int b, orc_shape *t)
{
    t->M_sub = b; t->N_sub = a; t->xStart = x; t->yStart = y;
    int M_sub = b, N_sub = a;
    int sd = M_sub > N_sub ? N_sub : M_sub;
    int width = orc_width_for(sd);
    t->width = width;
    int nd = 2 * width - 1;
    int dofs = y - x;
    /* pass 1 (:723-734): per-diagonal row lists, in row order */
    int *cnt = (int *)calloc((size_t)nd, sizeof(int));
    for (int i = 0; i < M; i++)
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) {
            int r = i, c = col[j], new_r = r + dofs;
            if (r >= x && r < x + b && c >= y && c < y + a &&
                c >= new_r - width + 1 && c <= new_r + width - 1 && orc_unclaimed(val, j))
                cnt[c - r - (y - x) + width - 1]++;
        }
    int *lptr = (int *)malloc(((size_t)nd + 1) * sizeof(int));
    lptr[0] = 0;
    for (int k = 0; k < nd; k++) lptr[k + 1] = lptr[k] + cnt[k];
    int *rows = (int *)malloc((size_t)(lptr[nd] ? lptr[nd] : 1) * sizeof(int));
    memset(cnt, 0, (size_t)nd * sizeof(int));
    for (int i = 0; i < M; i++)
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) {
            int r = i, c = col[j], new_r = r + dofs;
            if (r >= x && r < x + b && c >= y && c < y + a &&
                c >= new_r - width + 1 && c <= new_r + width - 1 && orc_unclaimed(val, j)) {
                int nn = c - r - (y - x) + width - 1;
                rows[lptr[nn] + cnt[nn]++] = r - x;
            }
        }
    /* dense diagonals (:752-766) */
    int *offs = (int *)malloc((size_t)nd * sizeof(int));
    int *st = (int *)malloc((size_t)nd * sizeof(int)), *en = (int *)malloc((size_t)nd * sizeof(int));
    int neg = 0;
    for (int n = 0; n < nd; n++) {
        int off = n - width + 1, s0, e0;
        orc_cal_start_end(M_sub, N_sub, width, off, n, &s0, &e0);
        float density = (float)cnt[n] / (float)(e0 - s0);
        if ((double)density >= 0.6) { st[neg] = s0; en[neg] = e0; offs[neg] = off; neg++; }
    }
    t->neg = neg;
    t->nnz = 0;
    if (neg > 0) {
        t->neg_offsets = (int *)malloc((size_t)neg * sizeof(int));
        t->numZeroNeg = (int *)malloc((size_t)neg * sizeof(int));
        t->rowZeroNeg = (int **)calloc((size_t)neg, sizeof(int *));
        for (int d = 0; d < neg; d++) {
            int nn = offs[d] + width - 1;
            t->neg_offsets[d] = offs[d];
            t->nnz += cnt[nn];
            /* zero rows (:809-827); reading past the list end is undefined in
               the reference but both of its branches then record a zero. */
            int *zr = (int *)malloc((size_t)(en[d] - st[d] + 1) * sizeof(int));
            int zi = 0, ri = 0;
            const int *lst = rows + lptr[nn];
            for (int r = st[d]; r < en[d]; r++) {
                if (ri < cnt[nn] && r >= lst[ri]) ri++;
                else zr[zi++] = r;
            }
            t->rowZeroNeg[d] = zr;
            /* :787 stores end-start-cnt; that equals zi unless the CSR holds
               duplicate entries on this diagonal, where the reference then
               reads or writes past rowZeroNeg. zi keeps it defined. */
            t->numZeroNeg[d] = zi;
        }
        /* pass 2 (:848-874): claim the covered nonzeros */
        for (int i = 0; i < M; i++)
            for (int j = rowptr[i]; j < rowptr[i + 1]; j++) {
                int r = i, c = col[j], new_r = r + dofs;
                if (r >= x && r < x + b && c >= y && c < y + a &&
                    c >= new_r - width + 1 && c <= new_r + width - 1 && orc_unclaimed(val, j)) {
                    int rs = r - x, cs = c - y;
                    for (int d = 0; d < neg; d++)
                        if (cs - rs == offs[d]) val[j] = FLT_MAX;
                }
            }
    }
    free(cnt); free(lptr); free(rows); free(offs); free(st); free(en);
}

static int imin_arr(const int *a, int n) { int m = a[0]; for (int i = 1; i < n; i++) if (a[i] < m) m = a[i]; return m; }
static int imax_arr(const int *a, int n) { int m = a[0]; for (int i = 1; i < n; i++) if (a[i] > m) m = a[i]; return m; }

/* partition_shapes (TaiChi.cu:313-638), shapes processed serially in
   contour order (the reference's OpenMP loop at :360 races). */
int orc_partition_shapes(int M, int N, int nnz, int *rowptr, int *col, float *val_in,
                         const uint8_t *img_in, orc_taichi_new *T)
{
    memset(T, 0, sizeof *T);
    float *val = val_in;
    if (val == NULL) val = (float *)calloc(nnz ? (size_t)nnz : 1, sizeof(float));
    int H, W, pM, pN, num_shapes = 0, *sp = NULL;
    int mF = 0;
    if (orc_thumb_dims(M, N, &H, &W, &pM, &pN) == 0) {
        uint8_t *img = (uint8_t *)img_in;
        if (img == NULL) {
            img = (uint8_t *)malloc((size_t)H * W);
            orc_thumbnail(M, N, rowptr, col, img);
        }
        num_shapes = orc_sdf_from_image(img, H, W, &sp);
        mF = H;
        if (img != img_in) free(img);
    }
    orc_shape *sh = (orc_shape *)calloc(num_shapes ? (size_t)num_shapes : 1, sizeof(orc_shape));
    int dia_shapes = 0, num_negs = 0, dense_nnz = 0, zero_elements = 0;
    for (int i = 0; i < num_shapes; i++) {                     /* :361-421 */
        int x1 = sp[4 * i], y1 = sp[4 * i + 1];
        int a1 = sp[4 * i + 2] - x1, b1 = sp[4 * i + 3] - y1;
        int x, y, a, b;
        float gain;
        orc_cal_shape_size(M, N, mF, x1, y1, a1, b1, &x, &y, &a, &b, &gain);
        if (a < ORC_MIN_SHAPE || b < ORC_MIN_SHAPE) continue;
        orc_partition_dia(M, rowptr, col, val, x, y, a, b, &sh[i]);
        if (sh[i].nnz > 0) { dia_shapes++; num_negs += sh[i].neg; dense_nnz += sh[i].nnz; }
    }
    int *keep = (int *)malloc((size_t)(dia_shapes ? dia_shapes : 1) * sizeof(int));
    if (dia_shapes > 0) {                                      /* :429-460 */
        T->xStart = (int *)malloc((size_t)dia_shapes * sizeof(int));
        T->yStart = (int *)malloc((size_t)dia_shapes * sizeof(int));
        T->xStop = (int *)malloc((size_t)dia_shapes * sizeof(int));
        T->nDiasPtr = (int *)malloc(((size_t)dia_shapes + 1) * sizeof(int));
        T->neg_offsets = (int *)malloc((size_t)num_negs * sizeof(int));
        T->nDiasPtr[0] = 0;
        int d = 0, tn = 0;
        for (int i = 0; i < num_shapes; i++) {
            if (sh[i].nnz <= 0) continue;
            keep[d] = i;
            T->xStart[d] = sh[i].xStart;
            T->yStart[d] = sh[i].yStart;
            T->xStop[d] = sh[i].xStart + sh[i].M_sub;
            T->nDiasPtr[d + 1] = T->nDiasPtr[d] + sh[i].neg;
            memcpy(T->neg_offsets + tn, sh[i].neg_offsets, (size_t)sh[i].neg * sizeof(int));
            tn += sh[i].neg;
            for (int n = 0; n < sh[i].neg; n++) zero_elements += sh[i].numZeroNeg[n];
            d++;
        }
        T->row_start = imin_arr(T->xStart, dia_shapes);
        T->row_stop = imax_arr(T->xStop, dia_shapes);
    }
    int sparse_nnz = nnz - dense_nnz;                          /* :465 */
    T->dia_shapes = dia_shapes; T->dense_nnz = dense_nnz; T->sparse_nnz = sparse_nnz;
    T->zero_elements = zero_elements; T->num_negs = num_negs;

    int *rz = NULL, *cz = NULL, *rsp = NULL, *csp = NULL;
    if (zero_elements > 0) {                                   /* :495-547 */
        rz = (int *)calloc((size_t)M + 1, sizeof(int));
        cz = (int *)calloc((size_t)zero_elements, sizeof(int));
        int *ri = (int *)malloc((size_t)zero_elements * sizeof(int));
        int *ci = (int *)malloc((size_t)zero_elements * sizeof(int));
        int k = 0;
        for (int j = 0; j < dia_shapes; j++) {
            orc_shape *s = &sh[keep[j]];
            for (int d = 0; d < s->neg; d++)
                for (int q = 0; q < s->numZeroNeg[d]; q++) {
                    ri[k] = s->xStart + s->rowZeroNeg[d][q];
                    rz[ri[k]]++;
                    ci[k] = (s->yStart + s->rowZeroNeg[d][q] + s->neg_offsets[d]) << 1;
                    k++;
                }
        }
        int old = rz[0];
        rz[0] = 0;
        for (int i = 1; i <= M; i++) { int nw = rz[i]; rz[i] = old + rz[i - 1]; old = nw; }
        int *tmp = (int *)calloc((size_t)M, sizeof(int));
        for (int i = 0; i < zero_elements; i++) cz[rz[ri[i]] + tmp[ri[i]]++] = ci[i];
        free(ri); free(ci); free(tmp);
    }
    if (sparse_nnz > 0 && dense_nnz > 0) {                     /* :548-567 */
        rsp = (int *)calloc((size_t)M + 1, sizeof(int));
        csp = (int *)calloc((size_t)sparse_nnz, sizeof(int));
        int k = 0;
        for (int r = 0; r < M; r++) {
            rsp[r + 1] = rsp[r];
            for (int j = rowptr[r]; j < rowptr[r + 1]; j++)
                if (orc_unclaimed(val, j) && k < sparse_nnz) { rsp[r + 1]++; csp[k++] = (col[j] << 1) + 1; }
        }
    }
    int nshape = zero_elements + sparse_nnz;                   /* :573-612 */
    if (dense_nnz == 0) {
        T->csrRow = rowptr; T->csrCol = col; T->aliased = 1;
    } else if (nshape > 0) {
        T->csrRow = (int *)calloc((size_t)M + 1, sizeof(int));
        T->csrCol = (int *)calloc((size_t)nshape, sizeof(int));
        if (zero_elements > 0 && sparse_nnz > 0) {
            int w = 0, zi = 0, si = 0;
            for (int r = 0; r < M; r++) {
                int nz = rz[r + 1] - rz[r], ns = rsp[r + 1] - rsp[r];
                T->csrRow[r + 1] = rz[r + 1] + rsp[r + 1];
                memcpy(T->csrCol + w, cz + zi, (size_t)nz * sizeof(int)); w += nz; zi += nz;
                memcpy(T->csrCol + w, csp + si, (size_t)ns * sizeof(int)); w += ns; si += ns;
            }
        } else if (zero_elements == 0) {
            memcpy(T->csrRow, rsp, ((size_t)M + 1) * sizeof(int));
            memcpy(T->csrCol, csp, (size_t)sparse_nnz * sizeof(int));
        } else {
            memcpy(T->csrRow, rz, ((size_t)M + 1) * sizeof(int));
            memcpy(T->csrCol, cz, (size_t)zero_elements * sizeof(int));
        }
    } else {
        /* dense_nnz > 0 with an empty remainder: the reference leaves NULL
           pointers (undefined behaviour at :1382-1383); use an empty CSR. */
        T->csrRow = (int *)calloc((size_t)M + 1, sizeof(int));
        T->csrCol = NULL;
    }
    free(rz); free(cz); free(rsp); free(csp);
    for (int i = 0; i < num_shapes; i++) {
        if (sh[i].neg > 0) {
            for (int d = 0; d < sh[i].neg; d++) free(sh[i].rowZeroNeg[d]);
            free(sh[i].rowZeroNeg); free(sh[i].numZeroNeg); free(sh[i].neg_offsets);
        }
    }
    free(sh); free(keep); free(sp);
    if (val != val_in) free(val);
    return 0;
}

void orc_free_taichi_new_arrays(orc_taichi_new *t)
{
    free(t->xStart); free(t->yStart); free(t->xStop);
    free(t->nDiasPtr); free(t->neg_offsets);
    if (!t->aliased) { free(t->csrRow); free(t->csrCol); }
    memset(t, 0, sizeof *t);
}

/* ===================================================================== */
/* a13: SpMV oracles                                                      */
/* ===================================================================== */

/* main.cu:61-69: sequential FP32 sum over the CSR, values ignored. */
void orc_csr_spmv_f32(int M, const int *rowptr, const int *col, const float *x, float *y)
{
    for (int i = 0; i < M; i++) {
        y[i] = 0.0f;
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) y[i] += x[col[j]];
    }
}

/* The same sums with rows split statically over `threads` OpenMP threads:
   each row is still summed sequentially, so y equals orc_csr_spmv_f32's.
   The CPU baseline across host cores. */
void orc_csr_spmv_f32_mt(int M, const int *rowptr, const int *col, const float *x, float *y, int threads)
{
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int i = 0; i < M; i++) {
        float s = 0.0f;
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) s += x[col[j]];
        y[i] = s;
    }
}

void orc_csr_spmv_f64(int M, const int *rowptr, const int *col, const double *x, double *y)
{
    for (int i = 0; i < M; i++) {
        double s = 0.0;
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) s += x[col[j]];
        y[i] = s;
    }
}

void orc_csr_absrow_f64(int M, const int *rowptr, const int *col, const double *x, double *y)
{
    for (int i = 0; i < M; i++) {
        double s = 0.0;
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) s += fabs(x[col[j]]);
        y[i] = s;
    }
}

void orc_csr_spmm_f64(int M, const int *rowptr, const int *col, const double *X, double *Y, int k)
{
    for (int i = 0; i < M; i++) {
        double *yr = Y + (size_t)i * k;
        for (int c = 0; c < k; c++) yr[c] = 0.0;
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) {
            const double *xr = X + (size_t)col[j] * k;
            for (int c = 0; c < k; c++) yr[c] += xr[c];
        }
    }
}

/* Format interpreter: remainder (signed, TaiChi.cu:1007-1013 / FUNC2 :32-36),
   then every shape covering the row, offsets in stored order, with the
   column-window test of :1015-1024. */
#define ORC_INTERP(T_, NAME)                                                         \
    void NAME(int M, const orc_taichi_new *t, const T_ *x, T_ *y)                     \
    {                                                                                \
        for (int r = 0; r < M; r++) {                                                \
            T_ s = 0;                                                                \
            for (int j = t->csrRow[r]; j < t->csrRow[r + 1]; j++) {                  \
                int v = t->csrCol[j];                                                \
                if (t->dense_nnz == 0) s += x[v];                                    \
                else if (v & 1) s += x[v >> 1];                                      \
                else s -= x[v >> 1];                                                 \
            }                                                                        \
            for (int q = 0; q < t->dia_shapes; q++) {                                \
                int xs = t->xStart[q], xe = t->xStop[q], ys = t->yStart[q];          \
                if (r < xs || r >= xe) continue;                                     \
                for (int n = t->nDiasPtr[q]; n < t->nDiasPtr[q + 1]; n++) {          \
                    int c = r - xs + ys + t->neg_offsets[n];                         \
                    if (c >= ys && c < ys + (xe - xs)) s += x[c];                    \
                }                                                                    \
            }                                                                        \
            y[r] = s;                                                                \
        }                                                                            \
    }
ORC_INTERP(float, orc_taichi_spmv_f32)
ORC_INTERP(double, orc_taichi_spmv_f64)
