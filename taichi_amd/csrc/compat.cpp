// The reference's C++ entry points (TaiChi.h:79-107) over the C ABI, so a
// main.cu-style caller links against libtaichi_amd.so unchanged. Printing
// follows the reference's lines; errors that the reference only prints
// (checkcuda, TaiChi.cu:42-48) are printed the same way here.
#include "TaiChi.h"

#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "common.h"

#include "partition.h"

int readMtx(char *filename, int &m, int &n, unsigned long &nnzA, int *&csrRowPtrA, int *&csrColIdxA,
            float *&csrValA)
{
    return tc_read_mtx(filename, &m, &n, &nnzA, &csrRowPtrA, &csrColIdxA, &csrValA);
}

void partition_shapes(char *sdf_name, int M, int N, int nnz, int *&csRow, int *&csrCol, float *&csrVal,
                      TaiChi_new *neg_format)
{
    auto t0 = std::chrono::steady_clock::now();
    int rc = tc_partition_shapes(sdf_name, M, N, nnz, csRow, csrCol, csrVal, neg_format);
    auto t1 = std::chrono::steady_clock::now();
    if (rc != TC_OK) {
        printf("partition_shapes failed: %s (%s)\n", tc_status_string(rc), tc_last_error());
        return;
    }
    printf("dia_shapes=%d, num_negs=%d, dense_nnz=%d, sparse_nnz=%d\n", neg_format->dia_shapes,
           neg_format->num_negs, neg_format->dense_nnz, neg_format->sparse_nnz);
    printf("partition:%.4f ms\n", std::chrono::duration<double, std::milli>(t1 - t0).count());
}

void cal_shape_size(int M, int N, int mF, int x1, int y1, int a1, int b1, int &x, int &y, int &a, int &b,
                    float &gain)
{
    tc_cal_shape_size(M, N, mF, x1, y1, a1, b1, &x, &y, &a, &b, &gain);
}

void cal_start_end(int M_sub, int N_sub, int width, int dia_offsets_temp, int n, int &start_temp, int &end_temp)
{
    tc_cal_start_end(M_sub, N_sub, width, dia_offsets_temp, n, &start_temp, &end_temp);
}

void partition_dia_shape(int M, int N, int nnz, int *&csRow, int *&csrCol, float *&csrVal, float gain, shape &hi,
                         TaiChi &t)
{
    (void)N;
    (void)gain;
    std::vector<uint8_t> claimed((size_t)nnz);
    for (int j = 0; j < nnz; j++) {
        float d = csrVal[j] - FLT_MAX;
        claimed[j] = !(std::fabs((double)d) > ZERO);
    }
    std::vector<uint8_t> pre = claimed;
    tc::DiaShape sh;
    if (hi.x >= 0 && hi.b > 0 && hi.x + hi.b <= M)
        tc::partition_square(csRow, csrCol, claimed, hi.x, hi.y, hi.b, sh);
    memset(&t, 0, sizeof t);
    t.shape = 0;
    t.M_sub = hi.b;
    t.N_sub = hi.a;
    t.xStart = hi.x;
    t.yStart = hi.y;
    t.width = sh.width;
    t.neg = (int)sh.offs.size();
    t.nnz = (int)sh.nnz;
    if (t.neg > 0) {
        const size_t n = (size_t)t.neg;
        t.neg_offsets = (int *)malloc(n * sizeof(int));
        t.numZeroNeg = (int *)malloc(n * sizeof(int));
        t.rowZeroNeg = (int **)calloc(n, sizeof(int *));
        t.start = (int *)malloc(n * sizeof(int));
        t.end = (int *)malloc(n * sizeof(int));
        t.cStart = (int *)malloc(n * sizeof(int));
        for (int d = 0; d < t.neg; d++) {
            t.neg_offsets[d] = sh.offs[d];
            t.start[d] = sh.st[d];
            t.end[d] = sh.en[d];
            t.cStart[d] = sh.offs[d] > 0 ? sh.offs[d] : 0;  // TaiChi.cu:789-792
            t.numZeroNeg[d] = (int)sh.zeros[d].size();
            if (t.numZeroNeg[d] > t.maxNumZero) t.maxNumZero = t.numZeroNeg[d];
        }
        for (int d = 0; d < t.neg; d++) {
            t.rowZeroNeg[d] = (int *)malloc(((size_t)t.maxNumZero + 1) * sizeof(int));
            if (!sh.zeros[d].empty())
                memcpy(t.rowZeroNeg[d], sh.zeros[d].data(), sh.zeros[d].size() * sizeof(int));
        }
    }
    for (int j = 0; j < nnz; j++)
        if (claimed[j] && !pre[j]) csrVal[j] = FLT_MAX;
}

void taichi_SpMV(int M, int N, TaiChi_new *neg_format, float *xHostPtr, float *yHostPtr)
{
    float transfer_ms = 0, spmv_ms = 0;
    int rc = tc_taichi_spmv(M, N, neg_format, xHostPtr, yHostPtr, 10, NUM_RUN, &transfer_ms, &spmv_ms);
    if (rc != TC_OK) {
        printf("HIP Runtime Error: %s (%s)\n", tc_status_string(rc), tc_last_error());
        return;
    }
    printf("transfer_time:%.4f spmv_time=%.4f\n", transfer_ms, spmv_ms);
}

void queryDevice()
{
    int n = tc_device_count();
    printf("HIP devices: %d\n", n);
}

int max(int *a, int len)
{
    int m = a[0];
    for (int i = 1; i < len; i++)
        if (a[i] > m) m = a[i];
    return m;
}

int min(int *a, int len)
{
    int m = a[0];
    for (int i = 1; i < len; i++)
        if (a[i] < m) m = a[i];
    return m;
}

void free_taichi_new(TaiChi_new *taichi_new) { tc_free_taichi_new(taichi_new); }

void free_taichi(int num_shapes, TaiChi *taichi)
{
    if (!taichi) return;
    for (int i = 0; i < num_shapes; i++) {
        TaiChi &t = taichi[i];
        if (t.rowZeroNeg)
            for (int d = 0; d < t.neg; d++) free(t.rowZeroNeg[d]);
        free(t.rowZeroNeg);
        free(t.neg_offsets);
        free(t.numZeroNeg);
        free(t.start);
        free(t.end);
        free(t.cStart);
    }
    free(taichi);
}
