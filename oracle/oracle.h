/*
 * oracle.h -- CPU restatement of TaiChi's binary-SpMV path (TEST INFRASTRUCTURE).
 *
 * This library is the checker for the HIP implementation in taichi_amd/. It is
 * linked or loaded ONLY by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py. The product library never links it, and no
 * product code path falls back to it.
 *
 * Every function restates one piece of the reference and cites the file:line
 * it follows. The reference is CUDA + OpenCV + CUSP; none of those exist here,
 * so the morphology and contour steps restate OpenCV's published semantics
 * (erode/dilate with anchor and constant border, 4-connected floodFill,
 * Suzuki-Abe border following as used by findContours) -- OpenCV's version is
 * unpinned by the reference Makefile (Makefile:11-12).
 *
 * Pinning: each function is checked in tests/ against hand-derived fixtures,
 * closed forms or brute force (see tests/test_oracle_*.py); the Matrix Market
 * banner/size parsing is also checked against the reference's own mmio.cpp
 * compiled into oracle/_ref/. Functions whose parity against the real OpenCV
 * binary cannot be pinned are marked "parity unpinned" below.
 */
#ifndef TAICHI_ORACLE_H
#define TAICHI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same field order and meaning as TaiChi_new (TaiChi.h:45-60). */
typedef struct orc_taichi_new {
    int dia_shapes;
    int num_negs;
    int dense_nnz;
    int zero_elements;
    int sparse_nnz;
    int row_start;
    int row_stop;
    int *xStart;
    int *yStart;
    int *xStop;
    int *nDiasPtr;
    int *neg_offsets;
    int *csrRow;
    int *csrCol;
    /* oracle-only bookkeeping: 1 if csrRow/csrCol alias the input CSR
       (dense_nnz == 0, TaiChi.cu:575-578) */
    int aliased;
} orc_taichi_new;

/* ---- a1: Matrix Market load (TaiChi.cu:120-278, mmio.cpp:96-217) ---- */
int  orc_read_mtx(const char *path, int *m, int *n, unsigned long *nnz,
                  int **rowptr, int **col, float **val);
/* banner only (mmio.cpp:96-179): returns 0 or an MM_* code; typecode[4] out */
int  orc_mm_read_banner_file(const char *path, char *typecode4);
void orc_free(void *p);

/* ---- a2: thumbnail (morohutils.cu:191-200 sizing, :153-180 binning) ---- */
/* returns 0 on success, 1 if M <= 512 (no thumbnail, morohutils.cu:192-196),
   2 if the width rounds to 0 (the reference would divide by zero at :200). */
int  orc_thumb_dims(int M, int N, int *H, int *W, int *pM, int *pN);
/* img must hold H*W bytes; 255 = empty pixel, 0 = pixel hit by a nonzero. */
void orc_thumbnail(int M, int N, const int *rowptr, const int *col, uint8_t *img);

/* ---- a3: morphology (OpenCV semantics, morohutils.cu:73-83, 231-251) ---- */
void orc_imfill(const uint8_t *src, uint8_t *dst, int H, int W);
/* kernel given as offsets relative to the anchor: element (dx[i],dy[i]) */
void orc_erode(const uint8_t *src, uint8_t *dst, int H, int W,
               const int *dx, const int *dy, int nk);
void orc_dilate(const uint8_t *src, uint8_t *dst, int H, int W,
                const int *dx, const int *dy, int nk);
/* which: 0 = 30x1 rect (anchor 15,0), 1 = 1x30 rect (anchor 0,15),
          2 = eye(5) (anchor 2,2) */
void orc_open(const uint8_t *src, uint8_t *dst, int H, int W, int which);
void orc_sub_sat(const uint8_t *a, const uint8_t *b, uint8_t *dst, int n);

/* findContours(RETR_CCOMP, CHAIN_APPROX_NONE) -- parity unpinned vs OpenCV.
   Output order is OpenCV's: outer borders in reverse discovery order, each
   followed by its holes (reverse discovery order). pts holds x,y pairs
   (x = column, y = row); ptr[c]..ptr[c+1] are contour c's points.
   Returns the number of contours; arrays are malloc'd (orc_free). */
int  orc_find_contours(const uint8_t *img, int H, int W,
                       int **pts, int **ptr, int **is_hole);

/* Whole getSdf image pipeline (morohutils.cu:231-282) on a thumbnail:
   returns shapeNum and writes 2*shapeNum points (x,y) into *pts (malloc'd):
   contour[0] then contour[size/2] per kept contour (size >= 80). */
int  orc_sdf_from_image(const uint8_t *img, int H, int W, int **pts);

/* ---- a4/a5: partition (TaiChi.cu:280-311, 677-916) ---- */
void orc_cal_shape_size(int M, int N, int mF, int x1, int y1, int a1, int b1,
                        int *x, int *y, int *a, int *b, float *gain);
void orc_cal_start_end(int M_sub, int N_sub, int width, int off, int n,
                       int *start, int *end);
int  orc_width_for(int smaller_dim);

/* ---- a4..a6: partition_shapes (TaiChi.cu:313-638), serial contour order.
   img: thumbnail (H*W) or NULL to compute it with orc_thumbnail.
   val is mutated exactly like the reference (FLT_MAX markers, :866); pass
   NULL to treat every entry as unclaimed. Returns 0. */
int  orc_partition_shapes(int M, int N, int nnz, int *rowptr, int *col,
                          float *val, const uint8_t *img, orc_taichi_new *out);
void orc_free_taichi_new_arrays(orc_taichi_new *t);

/* ---- a13: CPU CSR SpMV (main.cu:61-69), FP64 oracle, format interpreter ---- */
void orc_csr_spmv_f32(int M, const int *rowptr, const int *col,
                      const float *x, float *y);
void orc_csr_spmv_f32_mt(int M, const int *rowptr, const int *col, const float *x, float *y, int threads);
void orc_csr_spmv_f64(int M, const int *rowptr, const int *col,
                      const double *x, double *y);
/* per-row sum of |x[col]| over the CSR (componentwise error bound) */
void orc_csr_absrow_f64(int M, const int *rowptr, const int *col,
                        const double *x, double *y);
/* multi-RHS FP64 CSR: X is [N][k], Y is [M][k] */
void orc_csr_spmm_f64(int M, const int *rowptr, const int *col,
                      const double *X, double *Y, int k);
/* TaiChi_new interpreter (Appendix-A semantics of TaiChi.cu:1007-1024,
   1094-1111, 1181-1198): remainder terms in CSR order then shapes/offsets. */
void orc_taichi_spmv_f32(int M, const orc_taichi_new *t, const float *x, float *y);
void orc_taichi_spmv_f64(int M, const orc_taichi_new *t, const double *x, double *y);

#ifdef __cplusplus
}
#endif
#endif
