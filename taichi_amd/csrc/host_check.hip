// Host-only check of the library's threaded host code, built with the
// address + undefined-behaviour sanitizers and, separately, the thread
// sanitizer (make sanitize; no GPU involved). It runs the generators, the
// .mtx writer and the parallel reader (TaiChi.cu:120-278 semantics), the
// host partition (TaiChi.cu:313-638) from a CPU thumbnail, the device-layout
// planner (parallel_for over rows, presence bitmaps) and the format cache,
// on matrices that take every threaded path, and checks the results'
// invariants. Exit status 0 when every check passes.
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "partition.h"
#include "spmv_host.h"

namespace {

int g_fails = 0;

#define CHECK(c)                                                                                            \
    do {                                                                                                    \
        if (!(c)) {                                                                                         \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__);                             \
            g_fails++;                                                                                      \
        }                                                                                                   \
    } while (0)

struct Csr {
    int M = 0, N = 0;
    int *rp = nullptr, *col = nullptr;
    long long nnz = 0;
};

// the density thumbnail of a CSR by its definition (morohutils.cu:153-200):
// 255, and 0 at pixel (i/pM, c/pN) of every nonzero inside the image
std::vector<unsigned char> thumbnail(const Csr &A, int *H, int *W)
{
    int pM = 0, pN = 0;
    if (tc_thumb_dims(A.M, A.N, H, W, &pM, &pN) != 0) return {};
    std::vector<unsigned char> img((size_t)*H * *W, 255);
    for (int i = 0; i < A.M; i++)
        for (int j = A.rp[i]; j < A.rp[i + 1]; j++) {
            const int r = i / pM, c = A.col[j] / pN;
            if (r < *H && c < *W) img[(size_t)r * *W + c] = 0;
        }
    return img;
}

void same_format(const TaiChi_new &a, const TaiChi_new &b, int M)
{
    CHECK(a.dia_shapes == b.dia_shapes && a.num_negs == b.num_negs && a.dense_nnz == b.dense_nnz);
    CHECK(a.sparse_nnz == b.sparse_nnz && a.zero_elements == b.zero_elements);
    CHECK(a.row_start == b.row_start && a.row_stop == b.row_stop);
    const int S = a.dia_shapes, D = a.num_negs;
    if (S > 0) {
        CHECK(!std::memcmp(a.xStart, b.xStart, S * sizeof(int)) && !std::memcmp(a.yStart, b.yStart, S * sizeof(int)));
        CHECK(!std::memcmp(a.xStop, b.xStop, S * sizeof(int)));
        CHECK(!std::memcmp(a.nDiasPtr, b.nDiasPtr, (S + 1) * sizeof(int)));
        CHECK(!std::memcmp(a.neg_offsets, b.neg_offsets, D * sizeof(int)));
    }
    CHECK(!std::memcmp(a.csrRow, b.csrRow, ((size_t)M + 1) * sizeof(int)));
    CHECK(!std::memcmp(a.csrCol, b.csrCol, (size_t)a.csrRow[M] * sizeof(int)));
}

void one(const char *name, Csr A, const std::string &dir)
{
    std::fprintf(stderr, "%s: M=%d N=%d nnz=%lld\n", name, A.M, A.N, A.nnz);
    // .mtx round trip (the parallel parse takes files of 4 MiB and more)
    const std::string path = dir + "/" + name + ".mtx";
    CHECK(tc_write_mtx(path.c_str(), A.M, A.N, A.rp, A.col) == TC_OK);
    int m = 0, n = 0;
    unsigned long nnz = 0;
    int *rp = nullptr, *col = nullptr;
    float *val = nullptr;
    CHECK(tc_read_mtx(path.c_str(), &m, &n, &nnz, &rp, &col, &val) == 0);
    CHECK(m == A.M && n == A.N && (long long)nnz == A.nnz);
    if (rp && col) {
        CHECK(!std::memcmp(rp, A.rp, ((size_t)A.M + 1) * sizeof(int)));
        CHECK(!std::memcmp(col, A.col, (size_t)A.nnz * sizeof(int)));
    }
    (void)unlink(path.c_str());

    // host partition from the CPU thumbnail, with csrVal (FLT_MAX markers)
    int H = 0, W = 0;
    std::vector<unsigned char> img = thumbnail(A, &H, &W);
    TaiChi_new T;
    std::memset(&T, 0, sizeof T);
    CHECK(tc_partition_from_image(A.M, A.N, (int)A.nnz, A.rp, A.col, val, img.empty() ? nullptr : img.data(), &T) ==
          TC_OK);
    CHECK((long long)T.dense_nnz + T.sparse_nnz == A.nnz);
    CHECK(T.csrRow[A.M] == T.sparse_nnz + T.zero_elements);
    long long marked = 0;
    for (long long j = 0; j < A.nnz; j++) marked += val[j] > 1e30f;
    CHECK(marked == T.dense_nnz);

    // device layout (threaded planner, presence bitmaps)
    tcd::Layout P;
    CHECK(tcd::plan_layout(A.M, A.N, &T, 0, P) == TC_OK);
    CHECK(P.R + P.holes == T.csrRow[A.M]);
    CHECK(P.gbase.size() == (size_t)P.G + 1 && P.gbase[P.G] + (P.L ? P.lptr[P.L] - P.lptr[0] : 0) == P.R);
    for (int g = 0; g < P.G; g++) CHECK(P.gbase[g] <= P.gbase[g + 1]);
    std::fprintf(stderr, "  S=%d D=%d R=%lld holes=%lld long=%d bitmaps=%zu words\n", P.S, P.D, P.R, P.holes, P.L,
                 P.pres.size());

    // format cache round trip
    const std::string fpath = dir + "/" + name + ".tcfmt";
    CHECK(tc_save_format(fpath.c_str(), A.M, A.N, &T) == TC_OK);
    TaiChi_new U;
    std::memset(&U, 0, sizeof U);
    int M2 = 0, N2 = 0;
    CHECK(tc_load_format(fpath.c_str(), &M2, &N2, &U) == TC_OK);
    CHECK(M2 == A.M && N2 == A.N);
    if (M2 == A.M) same_format(T, U, A.M);
    (void)unlink(fpath.c_str());
    // the loader allocates every array, the CSR included (aliased or not)
    for (void *p : {(void *)U.xStart, (void *)U.yStart, (void *)U.xStop, (void *)U.nDiasPtr, (void *)U.neg_offsets,
                    (void *)U.csrRow, (void *)U.csrCol})
        free(p);

    if (T.dense_nnz == 0) T.csrRow = T.csrCol = nullptr;  // the input CSR (aliased)
    free(T.xStart);
    free(T.yStart);
    free(T.xStop);
    free(T.nDiasPtr);
    free(T.neg_offsets);
    free(T.csrRow);
    free(T.csrCol);
    tc_free_host(rp);
    tc_free_host(col);
    tc_free_host(val);
    tc_free_host(A.rp);
    tc_free_host(A.col);
}

}  // namespace

int main(int argc, char **argv)
{
    char tmpl[] = "/tmp/tc_host_check_XXXXXX";
    const char *d = argc > 1 ? argv[1] : mkdtemp(tmpl);
    if (!d) {
        std::perror("mkdtemp");
        return 2;
    }
    const std::string dir = d;
    Csr A;
    // R-MAT 2^18, edge factor 16: a 5.6 MB .mtx (parallel parse), one shape
    A.M = A.N = 1 << 18;
    CHECK(tc_gen_rmat(A.M, 18, 16ll << 18, 0.499, 0.001, 0.001, 1, &A.rp, &A.col, &A.nnz) == TC_OK);
    one("rmat18", A, dir);
    // a band of 40 diagonals at density 0.7 (presence bitmaps)
    A = Csr();
    A.M = A.N = 200000;
    CHECK(tc_gen_banded(A.M, 40, 0.7, 5, &A.rp, &A.col, &A.nnz) == TC_OK);
    one("band40", A, dir);
    // 512^2 stencil with 2% extra ones (three dense diagonals, zero cells)
    A = Csr();
    A.M = A.N = 512 * 512;
    CHECK(tc_gen_stencil5(512, 512, 0.02, 2, &A.rp, &A.col, &A.nnz) == TC_OK);
    one("stencil512", A, dir);
    // five diagonals plus uniform draws (cfg1)
    A = Csr();
    A.M = A.N = 16384;
    const int offs[5] = {-2, -1, 0, 1, 2};
    CHECK(tc_gen_diagonals_random(A.M, offs, 5, 16384, 0, &A.rp, &A.col, &A.nnz) == TC_OK);
    one("diag5", A, dir);
    if (argc <= 1) (void)rmdir(d);
    std::fprintf(stderr, "%s: %d failed checks\n", argv[0], g_fails);
    return g_fails == 0 ? 0 : 1;
}
