// The wide-rows kernel for k > 1 columns (a lane group per row, vector x
// row segments) and its launcher.
#include "spmv_host.h"

#include <type_traits>

namespace tcd {

// Waves a SIMD the compiler must fit the one-row-per-group tiles into
// (scalar registers beyond 96 alone cost these kernels a wave): 8, or 7 at
// 8 columns. The near-window tiles at 8 columns (77-85 vector registers) and
// the wider tiles (RI = 2) keep the compiler's own choice: a bound there
// spills to scratch memory.
constexpr int wide_waves(int K, int RI, bool NEAR)
{
    return RI != 1 || (K == 8 && NEAR) ? 1 : K == 8 ? 7 : 8;
}

template <int K, typename XT, int RI, bool LONG, bool MASK, bool NEAR>
__global__ __launch_bounds__(kWG, wide_waves(K, RI, NEAR)) void spmv_wide_kernel(
    MatView A, const XT *__restrict__ X, int ldx, float *__restrict__ Y, int ldy, int ntiles, int kv)
{
    extern __shared__ float smem[];
    int tb, te;
    block_units(ntiles, &tb, &te);
    for (int t = tb; t < te; t++) wide_tile<K, XT, RI, LONG, MASK, NEAR>(A, X, ldx, Y, ldy, t, smem, kv);
}

// kv < K: a masked chunk (only K = 2, 4, 8 are instantiated masked). A
// matrix whose remainder lies near its rows (A.near) runs the near-window
// instantiation at K <= 8; every other matrix runs the kernel without it.
template <int K, typename XT, int RI>
int launch_wide_t(const MatView &A, const void *X, int ldx, float *Y, int ldy, int grid, int kv, hipStream_t s)
{
    constexpr int TRW = (kWG / (K / wide_cpl(K, (int)sizeof(XT)))) * RI;
    const int ntiles = (A.M + TRW - 1) / TRW;
    if (ntiles == 0) return TC_OK;
    const int G = grid > 0 ? std::min(grid, ntiles) : ntiles;
    // the x window, reused for the long-row sums
    const bool near = K <= 8 && A.near;
    const size_t lds = (A.S > 0 || A.L > 0 || near) ? wide_lds_bytes(K, TRW, near, A.L > 0) : 0;
    const XT *x = (const XT *)X;
    auto go = [&](auto lng, auto msk, auto nr) {
        spmv_wide_kernel<K, XT, RI, decltype(lng)::value, decltype(msk)::value, decltype(nr)::value>
            <<<G, kWG, lds, s>>>(A, x, ldx, Y, ldy, ntiles, decltype(msk)::value ? kv : K);
    };
    using T_ = std::true_type;
    using F_ = std::false_type;
    if constexpr (K <= 8) {
        const bool m = kv < K;
        if (near) {
            if (A.L > 0) m ? go(T_{}, T_{}, T_{}) : go(T_{}, F_{}, T_{});
            else m ? go(F_{}, T_{}, T_{}) : go(F_{}, F_{}, T_{});
        } else {
            if (A.L > 0) m ? go(T_{}, T_{}, F_{}) : go(T_{}, F_{}, F_{});
            else m ? go(F_{}, T_{}, F_{}) : go(F_{}, F_{}, F_{});
        }
    } else {
        if (A.L > 0) go(T_{}, F_{}, F_{});
        else go(F_{}, F_{}, F_{});
    }
    HIPCHECK(hipGetLastError());
    return TC_OK;
}

template <int K, typename XT>
int launch_wide_k(const MatView &A, const void *X, int ldx, float *Y, int ldy, const tc_config &c, int kv,
                  hipStream_t s)
{
    if (c.rows_per_tile == wide_rows(K, 1, (int)sizeof(XT)))
        return launch_wide_t<K, XT, 1>(A, X, ldx, Y, ldy, c.grid, kv, s);
    return launch_wide_t<K, XT, 2>(A, X, ldx, Y, ldy, c.grid, kv, s);
}

template <typename XT>
int launch_wide_x(const MatView &A, const void *X, int ldx, float *Y, int ldy, int K, const tc_config &c, int kv,
                  hipStream_t s)
{
    switch (K) {
    case 2: return launch_wide_k<2, XT>(A, X, ldx, Y, ldy, c, kv, s);
    case 4: return launch_wide_k<4, XT>(A, X, ldx, Y, ldy, c, kv, s);
    case 8: return launch_wide_k<8, XT>(A, X, ldx, Y, ldy, c, kv, s);
    case 16: return launch_wide_k<16, XT>(A, X, ldx, Y, ldy, c, K, s);
    case 32: return launch_wide_k<32, XT>(A, X, ldx, Y, ldy, c, K, s);
    default: return launch_wide_k<64, XT>(A, X, ldx, Y, ldy, c, K, s);
    }
}

// wide rows apply when X and Y rows split into aligned CPL-element vectors
bool wide_applies(int K, const void *X, int ldx, const float *Y, int ldy, int esize, const tc_config &c)
{
    if (!wide_k(K) || (c.flags & kStagedK)) return false;
    if (c.rows_per_tile != wide_rows(K, 1, esize) && c.rows_per_tile != wide_rows(K, 2, esize)) return false;
    const int cpl = wide_cpl(K, esize);
    return ldx % cpl == 0 && ldy % cpl == 0 && (uintptr_t)X % ((uintptr_t)cpl * esize) == 0 &&
           (uintptr_t)Y % ((uintptr_t)cpl * 4) == 0;
}

int launch_wide(const MatView &A, const void *X, int ldx, float *Y, int ldy, int dtype, int K, const tc_config &c,
                hipStream_t s, int kv)
{
    if (kv <= 0 || kv > K) kv = K;
    if (kv < K && K > 8) return tc::fail(TC_ERR_UNSUPPORTED, "masked wide chunks are 2, 4 or 8 columns wide");
    if (dtype == TC_BF16) return launch_wide_x<bf16_t>(A, X, ldx, Y, ldy, K, c, kv, s);
    if (dtype == TC_F16) return launch_wide_x<f16_t>(A, X, ldx, Y, ldy, K, c, kv, s);
    return launch_wide_x<float>(A, X, ldx, Y, ldy, K, c, kv, s);
}

}  // namespace tcd
