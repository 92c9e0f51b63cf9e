// Synthetic matrix and vector generators (host, g++). The reference reads
// downloaded SuiteSparse / GTgraph R-MAT files (README.md:17-20); offline,
// these seeded generators stand in for them. Random values come from a
// counter-based hash (common.h), so the output does not depend on thread
// count or machine.
#include "common.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

// Builds a sorted, duplicate-free CSR from per-row candidate lists produced
// by `emit(row_begin, row_end, push)`; push(r, c) may be called in any order
// for rows of the range. Two passes: count, then fill, then sort+unique.
struct CsrBuilder {
    int n;
    std::vector<long long> cnt;
    explicit CsrBuilder(int n_) : n(n_), cnt((size_t)n_ + 1, 0) {}
};

int finish_csr(int n, std::vector<int> &ptr, std::vector<int> &col, int **rowptr, int **colout,
               long long *nnz)
{
    // sort + unique each row in place, then compact
    std::vector<int> newlen((size_t)n);
    tc::parallel_for(n, [&](long long b, long long e) {
        for (long long r = b; r < e; r++) {
            int *s = col.data() + ptr[r], *t = col.data() + ptr[r + 1];
            std::sort(s, t);
            newlen[r] = (int)(std::unique(s, t) - s);
        }
    });
    long long total = 0;
    for (int r = 0; r < n; r++) total += newlen[r];
    if (total > INT32_MAX) return tc::fail(TC_ERR_UNSUPPORTED, "generator: more than 2^31 entries");
    int *rp = (int *)malloc(((size_t)n + 1) * sizeof(int));
    int *cc = (int *)malloc((size_t)(total ? total : 1) * sizeof(int));
    if (!rp || !cc) {
        free(rp);
        free(cc);
        return tc::fail(TC_ERR_NOMEM, "generator: out of host memory");
    }
    rp[0] = 0;
    for (int r = 0; r < n; r++) rp[r + 1] = rp[r] + newlen[r];
    tc::parallel_for(n, [&](long long b, long long e) {
        for (long long r = b; r < e; r++) memcpy(cc + rp[r], col.data() + ptr[r], (size_t)newlen[r] * sizeof(int));
    });
    *rowptr = rp;
    *colout = cc;
    *nnz = total;
    return TC_OK;
}

// Two-pass scatter: count(r) for every candidate, then place. `gen(pass,
// begin, end, sink)` enumerates candidates of work items [begin,end).
template <class Gen>
int build_from_candidates(int n, long long items, Gen gen, int **rowptr, int **col, long long *nnz)
{
    std::vector<std::atomic<int>> cnt((size_t)n);
    for (auto &c : cnt) c.store(0, std::memory_order_relaxed);
    tc::parallel_for(items, [&](long long b, long long e) {
        gen(b, e, [&](int r, int) { cnt[r].fetch_add(1, std::memory_order_relaxed); });
    });
    std::vector<int> ptr((size_t)n + 1);
    ptr[0] = 0;
    for (int r = 0; r < n; r++) {
        long long nx = (long long)ptr[r] + cnt[r].load();
        if (nx > INT32_MAX) return tc::fail(TC_ERR_UNSUPPORTED, "generator: more than 2^31 candidates");
        ptr[r + 1] = (int)nx;
    }
    std::vector<int> colv((size_t)ptr[n]);
    for (int r = 0; r < n; r++) cnt[r].store(ptr[r], std::memory_order_relaxed);
    tc::parallel_for(items, [&](long long b, long long e) {
        gen(b, e, [&](int r, int c) { colv[cnt[r].fetch_add(1, std::memory_order_relaxed)] = c; });
    });
    return finish_csr(n, ptr, colv, rowptr, col, nnz);
}

}  // namespace

extern "C" int tc_gen_diagonals_random(int n, const int *offsets, int noffsets, long long ndraws,
                                       uint64_t seed, int **rowptr, int **col, long long *nnz)
try {
    if (n <= 0 || noffsets < 0 || (noffsets && !offsets) || ndraws < 0 || !rowptr || !col || !nnz)
        return tc::fail(TC_ERR_INVALID, "tc_gen_diagonals_random: bad argument");
    std::vector<int> offs(offsets, offsets + noffsets);
    const uint64_t key = tc::rand_key(seed, 1);
    // items [0, n) are band rows, [n, n + ndraws) are random draws
    auto gen = [&](long long b, long long e, auto &&sink) {
        for (long long it = b; it < e; it++) {
            if (it < n) {
                for (int o : offs) {
                    long long c = it + o;
                    if (c >= 0 && c < n) sink((int)it, (int)c);
                }
            } else {
                long long d = it - n;
                int r = (int)tc::bounded(tc::rand_at(key, 2 * d), (uint64_t)n);
                int c = (int)tc::bounded(tc::rand_at(key, 2 * d + 1), (uint64_t)n);
                sink(r, c);
            }
        }
    };
    return build_from_candidates(n, (long long)n + ndraws, gen, rowptr, col, nnz);
} catch (...) {
    return tc::exception_status(__func__);
}

extern "C" int tc_gen_rmat(int n, int scale, long long nedges, double a, double b, double c,
                           uint64_t seed, int **rowptr, int **col, long long *nnz)
try {
    if (n <= 0 || scale < 0 || scale > 31 || ((long long)1 << scale) < n || nedges < 0 ||
        a < 0 || b < 0 || c < 0 || a + b + c > 1.0 || !rowptr || !col || !nnz)
        return tc::fail(TC_ERR_INVALID, "tc_gen_rmat: bad argument");
    // quadrant thresholds on 32-bit uniforms: [0,Ta) a, [Ta,Tb) b, [Tb,Tc) c, rest d
    const double two32 = 4294967296.0;
    const uint64_t Ta = (uint64_t)std::llround(a * two32);
    const uint64_t Tb = (uint64_t)std::llround((a + b) * two32);
    const uint64_t Tc = (uint64_t)std::llround((a + b + c) * two32);
    const uint64_t key = tc::rand_key(seed, 2);
    const int words = (scale + 1) / 2;
    auto gen = [&](long long bgn, long long end, auto &&sink) {
        for (long long e = bgn; e < end; e++) {
            uint32_t r = 0, cc = 0;
            for (int w = 0; w < words; w++) {
                uint64_t z = tc::rand_at(key, (uint64_t)e * (uint64_t)words + (uint64_t)w);
                for (int h = 0; h < 2 && 2 * w + h < scale; h++) {
                    uint64_t u = (h == 0) ? (z & 0xffffffffull) : (z >> 32);
                    uint32_t rb = 0, cb = 0;
                    if (u < Ta) {
                    } else if (u < Tb) {
                        cb = 1;
                    } else if (u < Tc) {
                        rb = 1;
                    } else {
                        rb = 1;
                        cb = 1;
                    }
                    r = (r << 1) | rb;
                    cc = (cc << 1) | cb;
                }
            }
            if ((long long)r < n && (long long)cc < n) sink((int)r, (int)cc);
        }
    };
    return build_from_candidates(n, nedges, gen, rowptr, col, nnz);
} catch (...) {
    return tc::exception_status(__func__);
}

extern "C" int tc_gen_stencil5(int gx, int gy, double extra_frac, uint64_t seed, int **rowptr,
                               int **col, long long *nnz)
try {
    if (gx <= 0 || gy <= 0 || (long long)gx * gy > INT32_MAX || extra_frac < 0 || !rowptr || !col || !nnz)
        return tc::fail(TC_ERR_INVALID, "tc_gen_stencil5: bad argument");
    const long long n = (long long)gx * gy;
    // stencil entries: 5n minus the missing neighbours on the grid boundary
    const long long nst = 5 * n - 2 * gx - 2 * gy;
    const long long extra = std::llround(extra_frac * (double)nst);
    const uint64_t key = tc::rand_key(seed, 3);
    auto gen = [&](long long b, long long e, auto &&sink) {
        for (long long it = b; it < e; it++) {
            if (it < n) {
                int i = (int)(it / gy), j = (int)(it % gy);
                if (i > 0) sink((int)it, (int)(it - gy));
                if (j > 0) sink((int)it, (int)(it - 1));
                sink((int)it, (int)it);
                if (j + 1 < gy) sink((int)it, (int)(it + 1));
                if (i + 1 < gx) sink((int)it, (int)(it + gy));
            } else {
                long long d = it - n;
                int r = (int)tc::bounded(tc::rand_at(key, 2 * d), (uint64_t)n);
                int c = (int)tc::bounded(tc::rand_at(key, 2 * d + 1), (uint64_t)n);
                sink(r, c);
            }
        }
    };
    return build_from_candidates((int)n, n + extra, gen, rowptr, col, nnz);
} catch (...) {
    return tc::exception_status(__func__);
}

extern "C" int tc_gen_banded(int n, int w, double density, uint64_t seed, int **rowptr, int **col,
                             long long *nnz)
try {
    if (n <= 0 || w <= 0 || density < 0 || density > 1 || !rowptr || !col || !nnz)
        return tc::fail(TC_ERR_INVALID, "tc_gen_banded: bad argument");
    const int lo = -(w / 2);
    const uint64_t key = tc::rand_key(seed, 4);
    const uint64_t thr = (uint64_t)std::llround(density * 4294967296.0);
    auto gen = [&](long long b, long long e, auto &&sink) {
        for (long long r = b; r < e; r++)
            for (int k = 0; k < w; k++) {
                long long c = r + lo + k;
                if (c < 0 || c >= n) continue;
                if ((tc::rand_at(key, (uint64_t)r * (uint64_t)w + (uint64_t)k) >> 32) < thr)
                    sink((int)r, (int)c);
            }
    };
    return build_from_candidates(n, n, gen, rowptr, col, nnz);
} catch (...) {
    return tc::exception_status(__func__);
}

extern "C" int tc_gen_vector(float *out, long long n, int kind, uint64_t seed)
try {
    if (!out || n < 0 || kind < 0 || kind > 2) return tc::fail(TC_ERR_INVALID, "tc_gen_vector: bad argument");
    const uint64_t key = tc::rand_key(seed, 5 + (uint64_t)kind);
    tc::parallel_for(n, [&](long long b, long long e) {
        for (long long i = b; i < e; i++) {
            if (kind == 0) {
                out[i] = tc::unit_float(tc::rand_at(key, (uint64_t)i));
            } else if (kind == 1) {
                out[i] = (float)tc::bounded(tc::rand_at(key, (uint64_t)i), 8);
            } else {
                uint64_t r1 = tc::rand_at(key, 2 * (uint64_t)i), r2 = tc::rand_at(key, 2 * (uint64_t)i + 1);
                double u1 = ((double)(r1 >> 11) + 1.0) * (1.0 / 9007199254740992.0);
                double u2 = (double)(r2 >> 11) * (1.0 / 9007199254740992.0);
                out[i] = (float)(std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
            }
        }
    });
    return TC_OK;
} catch (...) {
    return tc::exception_status(__func__);
}
