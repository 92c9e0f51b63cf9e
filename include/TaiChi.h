/*
 * TaiChi.h -- drop-in C++ surface of the reference's TaiChi.h (TaiChi.h:1-107)
 * for the MI355X build. Same type names, field order, macros and function
 * signatures, so a main.cu-style caller compiles unchanged against it; the
 * implementations (taichi_amd/csrc/compat.cpp) call the C ABI in
 * taichi_amd.h. No CUDA or cuSPARSE header is needed: the reference's
 * checkcuda/checkcusparse helpers (TaiChi.h:100-101) take CUDA types and are
 * not provided.
 */
#ifndef TAICHI_COMPAT_H
#define TAICHI_COMPAT_H

#include "taichi_amd.h"

/* record of one shape found in the thumbnail (TaiChi.h:5-18) */
typedef struct shape {
    int id;
    char category[1024];  /* always "diagonal" (morohutils.cu:186) */
    int x1;               /* contour[0].x in the thumbnail (used as a row start) */
    int y1;               /* contour[0].y in the thumbnail */
    int a1;               /* contour[size/2].x - x1 */
    int b1;               /* contour[size/2].y - y1 */
    int area1;
    int x;                /* row start in the matrix */
    int y;                /* column start in the matrix */
    int a;                /* side (columns) */
    int b;                /* side (rows) */
    int width;
} shape;

/* per-shape partition result (TaiChi.h:20-43) */
typedef struct TaiChi {
    char shape;        /* 0 diagonal */
    int M_sub;
    int N_sub;
    int xStart;
    int yStart;
    int width;
    int neg;           /* dense diagonals */
    int maxNumZero;
    int *neg_offsets;
    int *numZeroNeg;
    int **rowZeroNeg;
    int *start;
    int *end;
    int *cStart;
    int nnz;
    int *csrRow;
    int *csrCol;
} TaiChi;

/* TaiChi_new is declared in taichi_amd.h with the reference's layout
   (TaiChi.h:45-60). */

#ifndef VALUE_TYPE
#define VALUE_TYPE float
#endif
#ifndef NUM_RUN
#define NUM_RUN 500
#endif
#ifndef NUM_TRANSFER
#define NUM_TRANSFER 50
#endif
#define ZERO 1e-8

/* readMtx (TaiChi.h:79-80): returns 0 or -1..-4 */
int readMtx(char *filename, int &m, int &n, unsigned long &nnzA, int *&csrRowPtrA, int *&csrColIdxA,
            float *&csrValA);

/* partition_shapes (TaiChi.h:83-84): fills neg_format; marks claimed entries
   of csrVal with FLT_MAX; prints the reference's summary lines */
void partition_shapes(char *sdf_name, int M, int N, int nnz, int *&csRow, int *&csrCol, float *&csrVal,
                      TaiChi_new *neg_format);

/* cal_shape_size (TaiChi.h:87-88) */
void cal_shape_size(int M, int N, int mF, int x1, int y1, int a1, int b1, int &x, int &y, int &a, int &b,
                    float &gain);

/* partition_dia_shape (TaiChi.h:91-92): one square shape hi.x,hi.y,hi.b
   against the unclaimed (val != FLT_MAX) entries */
void partition_dia_shape(int M, int N, int nnz, int *&csRow, int *&csrCol, float *&csrVal, float gain, shape &hi,
                         TaiChi &neg_format_i);

/* taichi_SpMV (TaiChi.h:95): 10 warm-up + NUM_RUN timed launches on the GPU,
   prints "transfer_time:%.4f spmv_time=%.4f" (ms), copies y back */
void taichi_SpMV(int M, int N, TaiChi_new *neg_format, float *xHostPtr, float *yHostPtr);

/* prints HIP device properties (TaiChi.h:97) */
void queryDevice();

int max(int *a, int len);
int min(int *a, int len);
void cal_start_end(int M_sub, int N_sub, int width, int dia_offsets_temp, int n, int &start_temp, int &end_temp);

/* free_taichi_new (TaiChi.h:106): frees the arrays and the struct */
void free_taichi_new(TaiChi_new *taichi_new);
/* free_taichi (TaiChi.h:107): frees per-shape arrays and the array itself */
void free_taichi(int num_shapes, TaiChi *taichi);

#endif
