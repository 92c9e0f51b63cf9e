// On-disk cache of a partitioned matrix (no reference counterpart: the
// reference re-reads the .mtx text and re-partitions on every run,
// TaiChi.cu:120-278, 313-638). One file holds a TaiChi_new (TaiChi.h:45-60)
// with its sizes and a checksum, so a later run skips the load and the
// partition. Little-endian layout:
//
//   char     magic[8]   "TCFMT02\0"
//   int32    M, N
//   int32    dia_shapes, num_negs, dense_nnz, zero_elements, sparse_nnz,
//            row_start, row_stop
//   int64    R                       remainder entries (csrRow[M])
//   int32    xStart[S], yStart[S], xStop[S], nDiasPtr[S+1] (S > 0),
//            neg_offsets[D], csrRow[M+1], csrCol[R]
//   uint64   checksum of every byte before it: four interleaved FNV-1a 64
//            lanes over little-endian 8-byte words (word i to lane i % 4, a
//            zero-padded last word), folded as ((l0 * P ^ l1) * P ^ l2) * P ^ l3
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "common.h"
#include "taichi_amd.h"

namespace {

constexpr char kMagic[8] = {'T', 'C', 'F', 'M', 'T', '0', '2', '\0'};

// Word-wise FNV-1a over four independent lanes: a stream of bytes is taken as
// 8-byte little-endian words (the last one zero-padded), word i feeding lane
// i % 4. Four lanes keep four multiplies in flight, so the checksum runs at
// memory speed instead of one dependent multiply per byte.
struct Fnv {
    static constexpr uint64_t kP = 0x100000001b3ull;
    uint64_t l[4] = {0xcbf29ce484222325ull, 0xcbf29ce484222325ull ^ 1, 0xcbf29ce484222325ull ^ 2,
                     0xcbf29ce484222325ull ^ 3};
    unsigned char pend[8];
    size_t npend = 0;
    uint64_t nwords = 0;
    void word(uint64_t w)
    {
        uint64_t &h = l[nwords & 3];
        h = (h ^ w) * kP;
        nwords++;
    }
    void add(const void *p, size_t n)
    {
        const unsigned char *b = (const unsigned char *)p;
        while (n && npend) {  // complete a word started by the previous call
            pend[npend++] = *b++;
            n--;
            if (npend == 8) {
                uint64_t w;
                memcpy(&w, pend, 8);
                word(w);
                npend = 0;
            }
        }
        size_t i = 0;
        for (; i + 32 <= n; i += 32) {
            uint64_t w[4];
            memcpy(w, b + i, 32);
            for (int j = 0; j < 4; j++) word(w[j]);
        }
        for (; i + 8 <= n; i += 8) {
            uint64_t w;
            memcpy(&w, b + i, 8);
            word(w);
        }
        for (; i < n; i++) pend[npend++] = b[i];
    }
    uint64_t digest()
    {
        Fnv t = *this;
        if (t.npend) {
            uint64_t w = 0;
            memcpy(&w, t.pend, t.npend);
            t.word(w);
        }
        return ((t.l[0] * kP ^ t.l[1]) * kP ^ t.l[2]) * kP ^ t.l[3];
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
try {
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
    const uint64_t sum = w.fnv.digest();
    w.ok = w.ok && fwrite(&sum, 1, sizeof sum, f) == sizeof sum;
    const bool ok = w.ok && fclose(f) == 0;
    return ok ? TC_OK : tc::fail(-1, "write error on %s", path);
} catch (...) {
    return tc::exception_status(__func__);
}

extern "C" int tc_load_format(const char *path, int *M, int *N, TaiChi_new *t)
try {
    if (!path || !M || !N || !t) return tc::fail(TC_ERR_INVALID, "tc_load_format: bad argument");
    // the file mapped read-only with its pages populated (a file that cannot
    // be mapped is read into memory)
    const int fd = ::open(path, O_RDONLY);
    if (fd < 0) return tc::fail(-1, "cannot open %s", path);
    struct stat st;
    if (fstat(fd, &st) != 0) {
        ::close(fd);
        return tc::fail(-1, "cannot stat %s", path);
    }
    const size_t fsz = (size_t)st.st_size;
    void *map = fsz ? mmap(nullptr, fsz, PROT_READ, MAP_PRIVATE | MAP_POPULATE, fd, 0) : MAP_FAILED;
    std::vector<unsigned char> buf;
    if (map == MAP_FAILED) {
        map = nullptr;
        unsigned char tmp[1 << 16];
        ssize_t got;
        while ((got = ::read(fd, tmp, sizeof tmp)) > 0) buf.insert(buf.end(), tmp, tmp + got);
    }
    ::close(fd);
    struct Unmap {
        void *p;
        size_t n;
        ~Unmap()
        {
            if (p) munmap(p, n);
        }
    } unmap{map, fsz};
    Reader r;
    r.p = map ? (const unsigned char *)map : buf.data();
    r.end = r.p + (map ? fsz : buf.size());
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
    const uint64_t got = r.fnv.digest();
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
} catch (...) {
    return tc::exception_status(__func__);
}
