// The staged tile kernel: (row, column) pairs per lane, tiles of
// 256*P/K rows, every compiled column count K; the fallback for k > 1 when
// the wide-rows kernel does not apply.
#include "spmv_host.h"

namespace tcd {

template <int K, typename XT, int P>
__global__ __launch_bounds__(kWG) void spmv_kernel(MatView A, const XT *__restrict__ X, int ldx, float *__restrict__ Y,
                                                   int ldy, int nunits, int capv)
{
    extern __shared__ float smem[];
    constexpr int TRW = (kWG * P) / K;
    float *s_val = smem;
    acc_t *s_lsum = (acc_t *)(s_val + lsum_at(capv));
    int *s_start = (int *)(s_lsum + kWG * P);
    int ub, ue;
    block_units(nunits, &ub, &ue);
    for (int u = ub; u < ue; u++) {
        run_tile<K, XT, P>(A, X, ldx, Y, ldy, u * TRW, s_val, capv, s_start, s_lsum);
        __syncthreads();
    }
}

template <int K, typename XT, int P>
int launch_t(const MatView &A, const void *X, int ldx, float *Y, int ldy, int capv, int grid, hipStream_t s)
{
    constexpr int TRW = (kWG * P) / K;
    const int nunits = (A.M + TRW - 1) / TRW;
    if (nunits == 0) return TC_OK;
    int G = grid > 0 ? std::min(grid, nunits) : nunits;
    const size_t lds = lds_bytes(capv, P, TRW);
    spmv_kernel<K, XT, P><<<G, kWG, lds, s>>>(A, (const XT *)X, ldx, Y, ldy, nunits, capv);
    HIPCHECK(hipGetLastError());
    return TC_OK;
}

template <int K, typename XT>
int launch_p(const MatView &A, const void *X, int ldx, float *Y, int ldy, int P, int capv, int grid, hipStream_t s)
{
    if (P == 1) return launch_t<K, XT, 1>(A, X, ldx, Y, ldy, capv, grid, s);
    if (P == 4) return launch_t<K, XT, 4>(A, X, ldx, Y, ldy, capv, grid, s);
    return launch_t<K, XT, 2>(A, X, ldx, Y, ldy, capv, grid, s);
}

template <typename XT>
int launch_k_x(const MatView &A, const void *X, int ldx, float *Y, int ldy, int K, int P, int capv, int grid,
             hipStream_t s)
{
    switch (K) {
    case 1: return launch_p<1, XT>(A, X, ldx, Y, ldy, P, capv, grid, s);
    case 2: return launch_p<2, XT>(A, X, ldx, Y, ldy, P, capv, grid, s);
    case 3: return launch_p<3, XT>(A, X, ldx, Y, ldy, P, capv, grid, s);
    case 4: return launch_p<4, XT>(A, X, ldx, Y, ldy, P, capv, grid, s);
    case 5: return launch_p<5, XT>(A, X, ldx, Y, ldy, P, capv, grid, s);
    case 6: return launch_p<6, XT>(A, X, ldx, Y, ldy, P, capv, grid, s);
    case 7: return launch_p<7, XT>(A, X, ldx, Y, ldy, P, capv, grid, s);
    case 8: return launch_p<8, XT>(A, X, ldx, Y, ldy, P, capv, grid, s);
    case 16: return launch_p<16, XT>(A, X, ldx, Y, ldy, P, capv, grid, s);
    case 32: return launch_p<32, XT>(A, X, ldx, Y, ldy, P, capv, grid, s);
    default: return launch_p<64, XT>(A, X, ldx, Y, ldy, P, capv, grid, s);
    }
}

int launch_staged(const MatView &A, const void *X, int ldx, float *Y, int ldy, int dtype, int K, int P, int capv,
                  int grid, hipStream_t s)
{
    if (dtype == TC_BF16) return launch_k_x<bf16_t>(A, X, ldx, Y, ldy, K, P, capv, grid, s);
    if (dtype == TC_F16) return launch_k_x<f16_t>(A, X, ldx, Y, ldy, K, P, capv, grid, s);
    return launch_k_x<float>(A, X, ldx, Y, ldy, K, P, capv, grid, s);
}

}  // namespace tcd
