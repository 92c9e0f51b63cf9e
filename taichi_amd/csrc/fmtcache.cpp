// On-disk cache of a partitioned matrix (no reference counterpart: the
// reference re-reads the .mtx text and re-partitions on every run,
// TaiChi.cu:120-278, 313-638). One file holds a TaiChi_new (TaiChi.h:45-60)
// with its sizes and a checksum, so a later run skips the load and the
// partition. Little-endian layout:
//
//   char     magic[8]   "TCFMT01\0"
//   int32    M, N
//   int32    dia_shapes, num_negs, dense_nnz, zero_elements, sparse_nnz,
//            row_start, row_stop
//   int64    R                       remainder entries (csrRow[M])
//   int32    xStart[S], yStart[S], xStop[S], nDiasPtr[S+1] (S > 0),
//            neg_offsets[D], csrRow[M+1], csrCol[R]
//   uint64   FNV-1a 64 of every byte before it
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "common.h"
#include "taichi_amd.h"

namespace {

constexpr char kMagic[8] = {'T', 'C', 'F', 'M', 'T', '0', '1', '\0'};

struct Fnv {
    uint64_t h = 0xcbf29ce484222325ull;
    void add(const void *p, size_t n)
    {
        const unsigned char *b = (const unsigned char *)p;
        for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 0x100000001b3ull;
    }
};

struct Writer {
    FILE *f = nullptr;
    Fnv fnv;
    bool ok = true;
    void put(const void *p, size_t n)
    {
        if (!n) return;
        fnv.add(p, n);
        ok = ok && fwrite(p, 1, n, f) == n;
    }
};

struct Reader {
    const unsigned char *p = nullptr, *end = nullptr;
    Fnv fnv;
    bool get(void *dst, size_t n)
    {
        if ((size_t)(end - p) < n) return false;
        memcpy(dst, p, n);
        fnv.add(p, n);
        p += n;
        return true;
    }
};

int *take_ints(Reader &r, long long n, bool *ok)
{
    if (n <= 0) return nullptr;
    int *a = (int *)malloc((size_t)n * sizeof(int));
    if (!a || !r.get(a, (size_t)n * sizeof(int))) {
        free(a);
        *ok = false;
        return nullptr;
    }
    return a;
}

}  // namespace

extern "C" int tc_save_format(const char *path, int M, int N, const TaiChi_new *t)
{
    if (!path || !t || M < 0 || N < 0 || (M > 0 && !t->csrRow)) return tc::fail(TC_ERR_INVALID, "tc_save_format: bad argument");
    const int S = t->dia_shapes, D = t->num_negs;
    const long long R = M > 0 ? t->csrRow[M] : 0;
    if (S < 0 || D < 0 || R < 0 || (R > 0 && !t->csrCol) ||
        (S > 0 && (!t->xStart || !t->yStart || !t->xStop || !t->nDiasPtr)) || (D > 0 && !t->neg_offsets))
        return tc::fail(TC_ERR_INVALID, "tc_save_format: inconsistent format");
    FILE *f = fopen(path, "wb");
    if (!f) return tc::fail(-1, "cannot create %s", path);
    Writer w;
    w.f = f;
    const int32_t head[9] = {M, N, S, D, t->dense_nnz, t->zero_elements, t->sparse_nnz, t->row_start, t->row_stop};
    const int64_t r64 = R;
    w.put(kMagic, sizeof kMagic);
    w.put(head, sizeof head);
    w.put(&r64, sizeof r64);
    w.put(t->xStart, (size_t)S * 4);
    w.put(t->yStart, (size_t)S * 4);
    w.put(t->xStop, (size_t)S * 4);
    if (S > 0) w.put(t->nDiasPtr, ((size_t)S + 1) * 4);
    w.put(t->neg_offsets, (size_t)D * 4);
    w.put(t->csrRow, ((size_t)M + 1) * 4);
    w.put(t->csrCol, (size_t)R * 4);
    const uint64_t sum = w.fnv.h;
    w.ok = w.ok && fwrite(&sum, 1, sizeof sum, f) == sizeof sum;
    const bool ok = w.ok && fclose(f) == 0;
    return ok ? TC_OK : tc::fail(-1, "write error on %s", path);
}

extern "C" int tc_load_format(const char *path, int *M, int *N, TaiChi_new *t)
{
    if (!path || !M || !N || !t) return tc::fail(TC_ERR_INVALID, "tc_load_format: bad argument");
    FILE *f = fopen(path, "rb");
    if (!f) return tc::fail(-1, "cannot open %s", path);
    std::vector<unsigned char> buf;
    {
        unsigned char tmp[1 << 16];
        size_t got;
        while ((got = fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + got);
        fclose(f);
    }
    Reader r;
    r.p = buf.data();
    r.end = buf.data() + buf.size();
    char magic[8];
    int32_t head[9];
    int64_t R;
    if (!r.get(magic, 8) || memcmp(magic, kMagic, 8) != 0) return tc::fail(TC_ERR_INVALID, "%s: not a format cache file", path);
    if (!r.get(head, sizeof head) || !r.get(&R, sizeof R)) return tc::fail(TC_ERR_INVALID, "%s: truncated header", path);
    const int m = head[0], n = head[1], S = head[2], D = head[3];
    if (m < 0 || n < 0 || S < 0 || D < 0 || R < 0 || R > INT32_MAX) return tc::fail(TC_ERR_INVALID, "%s: bad sizes", path);
    // the payload must be exactly the arrays plus the checksum
    const long long need = 4ll * (3ll * S + (S > 0 ? S + 1 : 0) + D + (m + 1ll) + R) + 8;
    if ((long long)(r.end - r.p) != need) return tc::fail(TC_ERR_INVALID, "%s: size does not match its header", path);
    TaiChi_new out{};
    out.dia_shapes = S;
    out.num_negs = D;
    out.dense_nnz = head[4];
    out.zero_elements = head[5];
    out.sparse_nnz = head[6];
    out.row_start = head[7];
    out.row_stop = head[8];
    bool ok = true;
    out.xStart = take_ints(r, S, &ok);
    out.yStart = take_ints(r, S, &ok);
    out.xStop = take_ints(r, S, &ok);
    out.nDiasPtr = take_ints(r, S > 0 ? S + 1 : 0, &ok);
    out.neg_offsets = take_ints(r, D, &ok);
    out.csrRow = take_ints(r, m + 1ll, &ok);
    out.csrCol = take_ints(r, R, &ok);
    uint64_t want = 0;
    const uint64_t got = r.fnv.h;
    ok = ok && r.get(&want, sizeof want);
    auto drop = [&]() {
        free(out.xStart);
        free(out.yStart);
        free(out.xStop);
        free(out.nDiasPtr);
        free(out.neg_offsets);
        free(out.csrRow);
        free(out.csrCol);
    };
    if (!ok) {
        drop();
        return tc::fail(TC_ERR_NOMEM, "%s: out of host memory", path);
    }
    if (want != got) {
        drop();
        return tc::fail(TC_ERR_INVALID, "%s: checksum mismatch", path);
    }
    // structural checks the device upload relies on
    bool sane = out.csrRow[0] == 0 && out.csrRow[m] == R && (S == 0 || (out.nDiasPtr[0] == 0 && out.nDiasPtr[S] == D));
    for (int i = 0; sane && i < m; i++) sane = out.csrRow[i] <= out.csrRow[i + 1];
    if (!sane) {
        drop();
        return tc::fail(TC_ERR_INVALID, "%s: inconsistent format arrays", path);
    }
    *M = m;
    *N = n;
    *t = out;
    return TC_OK;
}
