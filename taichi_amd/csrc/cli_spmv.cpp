// taichi_spmv: the reference's driver (main.cu:19-119) against the
// MI355X build, through the drop-in TaiChi.h surface. Usage:
//   taichi_spmv matrix.mtx
// Loads the matrix, partitions it, sets x = 1, computes the CPU CSR product
// as main.cu:61-69 does, runs taichi_SpMV on the GPU and counts entries with
// |y - y_ref| > 1e-6. Unlike the reference (main.cu:118) the exit status is
// nonzero when the load fails or any entry is wrong.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "TaiChi.h"

int main(int argc, char *argv[])
{
    if (argc != 2) {
        printf("usage: ./exe MatrixName\n");
        return 2;
    }
    int M = 0, N = 0;
    unsigned long nnz = 0;
    int *rowptr = nullptr, *col = nullptr;
    float *val = nullptr;
    printf("reading file %s\n", argv[1]);
    int rc = readMtx(argv[1], M, N, nnz, rowptr, col, val);
    if (rc != 0) {
        printf("readMtx failed with %d\n", rc);
        return 1;
    }
    printf("M=%d N=%d nnz=%lu\n", M, N, nnz);
    TaiChi_new *fmt = (TaiChi_new *)calloc(1, sizeof(TaiChi_new));
    partition_shapes(argv[1], M, N, (int)nnz, rowptr, col, val, fmt);
    free(val);
    float *x = (float *)malloc(sizeof(float) * (N > 0 ? N : 1));
    float *y = (float *)calloc(M > 0 ? M : 1, sizeof(float));
    float *yref = (float *)malloc(sizeof(float) * (M > 0 ? M : 1));
    for (int i = 0; i < N; i++) x[i] = 1.0f;
    for (int i = 0; i < M; i++) {
        yref[i] = 0.0f;
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) yref[i] += x[col[j]];
    }
    taichi_SpMV(M, N, fmt, x, y);
    int wrong = 0;
    for (int i = 0; i < M; i++)
        if (std::fabs(y[i] - yref[i]) > 1e-6) {
            if (wrong == 0) printf("yHostPtr[%d]=%f, y_ref[%d]=%f\n", i, y[i], i, yref[i]);
            wrong++;
        }
    printf("Warning: %d are wrong!\n", wrong);
    free(x);
    free(y);
    free(yref);
    free_taichi_new(fmt);
    free(rowptr);
    free(col);
    return wrong == 0 ? 0 : 1;
}
