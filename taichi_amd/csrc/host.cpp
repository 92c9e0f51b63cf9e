// Status strings, error reporting, thread pool helper and the Matrix Market
// reader/writer of libtaichi_amd.so (host code, built with g++).
#include "common.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <exception>
#include <new>
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace tc {

static thread_local std::string g_last_error;

int fail(int status, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return status;
}

void clear_error() { g_last_error.clear(); }

int exception_status(const char *fn)
{
    try {
        throw;
    } catch (const std::bad_alloc &) {
        return fail(TC_ERR_NOMEM, "%s: out of host memory", fn);
    } catch (const std::exception &e) {
        return fail(TC_ERR_INVALID, "%s: %s", fn, e.what());
    } catch (...) {
        return fail(TC_ERR_INVALID, "%s: unexpected error", fn);
    }
}

void parallel_for(long long n, const std::function<void(long long, long long)> &fn, int max_threads,
                  long long grain)
{
    int hw = (int)std::thread::hardware_concurrency();
    if (hw <= 0) hw = 1;
    int nt = max_threads > 0 ? max_threads : std::min(hw, 16);
    if (n < grain || nt <= 1) {
        if (n > 0) fn(0, n);
        return;
    }
    nt = (int)std::min<long long>(nt, n);
    std::vector<std::thread> th;
    th.reserve(nt);
    for (int t = 0; t < nt; t++) {
        long long b = n * t / nt, e = n * (t + 1) / nt;
        th.emplace_back([&fn, b, e] { fn(b, e); });
    }
    for (auto &t : th) t.join();
}

}  // namespace tc

extern "C" {

const char *tc_status_string(int s)
{
    switch (s) {
    case TC_OK: return "ok";
    case -1: return "cannot open file";
    case -2: return "could not process Matrix Market banner";
    case -3: return "complex matrices are not supported";
    case -4: return "bad size line or data section";
    case TC_ERR_INVALID: return "invalid argument";
    case TC_ERR_HIP: return "HIP runtime error";
    case TC_ERR_NOMEM: return "out of memory";
    case TC_ERR_UNSUPPORTED: return "unsupported";
    case TC_ERR_CAPACITY: return "over capacity";
    case TC_ERR_NODEVICE: return "no HIP device";
    case TC_ERR_STATE: return "wrong state";
    default: return "unknown status";
    }
}

const char *tc_last_error(void) { return tc::g_last_error.c_str(); }

int tc_abi_version(void) { return TC_ABI_VERSION; }

void tc_free_host(void *p) { free(p); }

}  // extern "C"

// ---------------------------------------------------------------------------
// Matrix Market reader. Same results and return codes as readMtx
// (TaiChi.cu:120-278) with mm_read_banner / mm_read_mtx_crd_size
// (mmio.cpp:96-217): row order = file order, duplicates kept, pattern -> 1.0,
// integer -> double -> float, symmetric/hermitian off-diagonals mirrored.
// The data section is parsed from one buffer instead of fscanf per entry.
// ---------------------------------------------------------------------------
namespace {

struct Cursor {
    const char *p, *end;
    void skip_ws()
    {
        while (p < end && isspace((unsigned char)*p)) ++p;
    }
    bool read_long(long long *v)
    {
        skip_ws();
        if (p >= end) return false;
        char *q;
        errno = 0;
        long long r = strtoll(p, &q, 10);
        if (q == p) return false;
        p = q;
        *v = r;
        return true;
    }
    bool read_double(double *v)
    {
        skip_ws();
        if (p >= end) return false;
        char *q;
        double r = strtod(p, &q);
        if (q == p) return false;
        p = q;
        *v = r;
        return true;
    }
};

// Banner fields as mm_read_banner classifies them. Returns 0 or an MM_* code.
// A text file in memory with a NUL after its last byte.
struct MappedText {
    const char *data = nullptr;
    size_t size = 0;
    void *map = nullptr;
    size_t map_len = 0;
    std::vector<char> buf;

    bool open(const char *path)
    {
        const int fd = ::open(path, O_RDONLY);
        if (fd < 0) return false;
        struct stat st;
        const long page = sysconf(_SC_PAGESIZE);
        if (fstat(fd, &st) == 0 && S_ISREG(st.st_mode) && st.st_size > 0 && page > 0 && st.st_size % page != 0) {
            void *m = mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_PRIVATE | MAP_POPULATE, fd, 0);
            if (m != MAP_FAILED) {
                ::close(fd);
                map = m;
                map_len = (size_t)st.st_size;
                data = (const char *)m;
                size = map_len;
                return true;
            }
        }
        char tmp[1 << 16];
        ssize_t got;
        while ((got = ::read(fd, tmp, sizeof tmp)) > 0) buf.insert(buf.end(), tmp, tmp + got);
        ::close(fd);
        if (got < 0) return false;
        size = buf.size();
        buf.push_back('\0');
        data = buf.data();
        return true;
    }
    ~MappedText()
    {
        if (map) munmap(map, map_len);
    }
};

int parse_banner(const char *line, char tc[4])
{
    char banner[64], mtx[64], crd[64], dtype[64], storage[64];
    tc[0] = tc[1] = tc[2] = ' ';
    tc[3] = 'G';
    if (sscanf(line, "%63s %63s %63s %63s %63s", banner, mtx, crd, dtype, storage) != 5) return 12;
    auto lower = [](char *s) {
        for (; *s; ++s) *s = (char)tolower((unsigned char)*s);
    };
    lower(mtx);
    lower(crd);
    lower(dtype);
    lower(storage);
    if (strncmp(banner, "%%MatrixMarket", 14) != 0) return 14;
    if (strcmp(mtx, "matrix") != 0) return 15;
    tc[0] = 'M';
    if (!strcmp(crd, "coordinate")) tc[1] = 'C';
    else if (!strcmp(crd, "array")) tc[1] = 'A';
    else return 15;
    if (!strcmp(dtype, "real")) tc[2] = 'R';
    else if (!strcmp(dtype, "complex")) tc[2] = 'C';
    else if (!strcmp(dtype, "pattern")) tc[2] = 'P';
    else if (!strcmp(dtype, "integer")) tc[2] = 'I';
    else return 15;
    if (!strcmp(storage, "general")) tc[3] = 'G';
    else if (!strcmp(storage, "symmetric")) tc[3] = 'S';
    else if (!strcmp(storage, "hermitian")) tc[3] = 'H';
    else if (!strcmp(storage, "skew-symmetric")) tc[3] = 'K';
    else return 15;
    return 0;
}

// Whole-line integer of the data section: optional sign, digits.
inline bool line_int(const char *&q, const char *e, long long *v)
{
    while (q < e && (*q == ' ' || *q == '\t')) ++q;
    bool neg = false;
    if (q < e && (*q == '-' || *q == '+')) neg = *q++ == '-';
    if (q >= e || *q < '0' || *q > '9') return false;
    long long r = 0;
    while (q < e && *q >= '0' && *q <= '9') {
        if (r > (INT64_MAX - 9) / 10) return false;
        r = r * 10 + (*q++ - '0');
    }
    *v = neg ? -r : r;
    return true;
}

// Parallel parse of a line-structured data section: every non-blank line of
// [p, end) holds exactly one entry (row, col and, unless pattern, a value).
// Each thread parses a newline-aligned chunk; the chunks are joined in file
// order and the first NZ entries kept, as the serial reader would read them.
// Returns false -- and the caller parses serially, which gives the reference
// semantics and error codes for any layout -- on a short file, an entry out
// of range, fewer than NZ entries, or any line that is not exactly one entry.
bool parse_lines_parallel(const char *p, const char *end, long long M, long long N, long long NZ, char kind,
                          std::vector<int> &ri, std::vector<int> &ci, std::vector<float> &vi)
{
    if (end - p < (1 << 22)) return false;  // small files: the serial parse is as fast
    const int T = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    std::vector<const char *> bnd((size_t)T + 1);
    bnd[0] = p;
    bnd[T] = end;
    for (int t = 1; t < T; t++) {
        const char *q = p + (end - p) * t / T;
        if (q < bnd[t - 1]) q = bnd[t - 1];
        const char *nl = (const char *)memchr(q, '\n', end - q);
        bnd[t] = nl ? nl + 1 : end;
    }
    struct Part {
        std::vector<int> r, c;
        std::vector<float> v;
        bool ok = true;
    };
    std::vector<Part> parts((size_t)T);
    tc::parallel_for(T, [&](long long tb, long long te) {
        for (long long t = tb; t < te; t++) {
            Part &pt = parts[t];
            const char *q = bnd[t], *e = bnd[t + 1];
            pt.r.reserve((size_t)((e - q) / 8));
            pt.c.reserve((size_t)((e - q) / 8));
            while (q < e && pt.ok) {
                const char *le = (const char *)memchr(q, '\n', e - q);
                if (!le) le = e;
                const char *x = q;
                while (x < le && isspace((unsigned char)*x)) ++x;
                if (x < le) {
                    long long a, b, iv;
                    double fv = 1.0;
                    bool good = line_int(x, le, &a) && line_int(x, le, &b);
                    if (good && kind == 'I') {
                        good = line_int(x, le, &iv);
                        fv = (double)(int)iv;
                    } else if (good && kind == 'R') {
                        char *y;
                        fv = strtod(x, &y);
                        good = y != x && y <= le;
                        x = y;
                    }
                    while (good && x < le && isspace((unsigned char)*x)) ++x;
                    good = good && x == le && a >= 1 && a <= M && b >= 1 && b <= N;
                    if (!good) {
                        pt.ok = false;
                        break;
                    }
                    pt.r.push_back((int)(a - 1));
                    pt.c.push_back((int)(b - 1));
                    if (kind != 'P') pt.v.push_back((float)fv);
                }
                q = le + 1;
            }
        }
    }, T, 1);
    long long total = 0;
    for (const Part &pt : parts) {
        if (!pt.ok) return false;
        total += (long long)pt.r.size();
    }
    if (total < NZ) return false;
    ri.resize((size_t)NZ);
    ci.resize((size_t)NZ);
    vi.resize((size_t)NZ);
    std::vector<long long> off((size_t)T + 1, 0);
    for (int t = 0; t < T; t++) off[t + 1] = off[t] + (long long)parts[t].r.size();
    tc::parallel_for(T, [&](long long tb, long long te) {  // join in file order, first NZ entries
        for (long long t = tb; t < te; t++) {
            const Part &pt = parts[t];
            const long long o = off[t], take = std::min<long long>((long long)pt.r.size(), NZ - o);
            if (take <= 0) continue;
            std::copy(pt.r.begin(), pt.r.begin() + take, ri.begin() + o);
            std::copy(pt.c.begin(), pt.c.begin() + take, ci.begin() + o);
            if (kind != 'P') std::copy(pt.v.begin(), pt.v.begin() + take, vi.begin() + o);
            else std::fill(vi.begin() + o, vi.begin() + o + take, 1.0f);
        }
    }, T, 1);
    return true;
}

}  // namespace

extern "C" int tc_read_mtx(const char *path, int *m, int *n, unsigned long *nnzA, int **rowptr,
                           int **colidx, float **val)
try {
    if (!path || !m || !n || !nnzA || !rowptr || !colidx || !val)
        return tc::fail(TC_ERR_INVALID, "tc_read_mtx: null argument");
    // The file, NUL-terminated for the text scanners. A regular file whose
    // size is not a multiple of the page size is mapped read-only with its
    // pages populated up front (the bytes after EOF in the last page read as
    // 0); anything else is read into memory.
    MappedText text;
    if (!text.open(path)) return tc::fail(-1, "cannot open %s", path);
    const char *p = text.data, *end = text.data + text.size;
    // banner: first line (fgets of at most 1024 chars, mmio.cpp:109)
    const char *nl = (const char *)memchr(p, '\n', end - p);
    std::string first(p, nl ? nl : end);
    if (first.size() > 1024) first.resize(1024);
    char tc[4];
    if (first.empty() && !nl) return tc::fail(-2, "empty file");
    if (parse_banner(first.c_str(), tc) != 0) return tc::fail(-2, "bad Matrix Market banner");
    if (tc[2] == 'C') return tc::fail(-3, "complex matrices are not supported");
    p = nl ? nl + 1 : end;
    // size line: skip lines starting with '%' (mmio.cpp:199-203)
    for (;;) {
        if (p >= end) return tc::fail(-4, "premature end before the size line");
        const char *e = (const char *)memchr(p, '\n', end - p);
        if (*p != '%') break;
        p = e ? e + 1 : end;
    }
    Cursor cur{p, end};
    long long M, N, NZ;
    if (!cur.read_long(&M) || !cur.read_long(&N) || !cur.read_long(&NZ))
        return tc::fail(-4, "bad size line");
    // 'array' files: the reference misreads their two-number size line; rejected.
    if (tc[1] != 'C' || M < 0 || N < 0 || NZ < 0 || M > INT32_MAX || N > INT32_MAX)
        return tc::fail(-4, "unsupported size line or storage");
    const bool isReal = tc[2] == 'R', isInt = tc[2] == 'I';
    const bool isSym = tc[3] == 'S' || tc[3] == 'H';
    std::vector<int> ri, ci;
    std::vector<float> vi;
    std::vector<long long> cnt((size_t)M + 1, 0);
    const bool parallel = parse_lines_parallel(cur.p, end, M, N, NZ, tc[2], ri, ci, vi);
    bool sorted = true;  // entries already in row order (file order = CSR order)
    if (parallel) {
        for (long long i = 0; i < NZ; i++) {
            cnt[ri[i]]++;
            if (isSym && ri[i] != ci[i]) cnt[ci[i]]++;
            if (i > 0 && ri[i] < ri[i - 1]) sorted = false;
        }
    } else {
        ri.resize((size_t)NZ);
        ci.resize((size_t)NZ);
        vi.resize((size_t)NZ);
    }
    for (long long i = 0; i < NZ && !parallel; i++) {
        long long a, b;
        double fv = 1.0;
        if (!cur.read_long(&a) || !cur.read_long(&b))
            return tc::fail(-4, "data section ends at entry %lld", i);
        if (isReal) {
            if (!cur.read_double(&fv)) return tc::fail(-4, "missing value at entry %lld", i);
        } else if (isInt) {
            long long iv;
            if (!cur.read_long(&iv)) return tc::fail(-4, "missing value at entry %lld", i);
            fv = (double)(int)iv;
        }
        a--;
        b--;
        if (a < 0 || a >= M || b < 0 || b >= N)
            return tc::fail(-4, "entry %lld out of range", i);
        ri[i] = (int)a;
        ci[i] = (int)b;
        vi[i] = (float)fv;
        if (i > 0 && ri[i] < ri[i - 1]) sorted = false;
        cnt[a]++;
        if (isSym && a != b) cnt[b]++;
    }
    long long total = 0;
    for (long long r = 0; r < M; r++) total += cnt[r];
    if (total > INT32_MAX) return tc::fail(-4, "too many entries");
    int *rp = (int *)malloc(((size_t)M + 1) * sizeof(int));
    int *cc = (int *)malloc((size_t)(total ? total : 1) * sizeof(int));
    float *vv = (float *)malloc((size_t)(total ? total : 1) * sizeof(float));
    if (!rp || !cc || !vv) {
        free(rp);
        free(cc);
        free(vv);
        return tc::fail(TC_ERR_NOMEM, "tc_read_mtx: out of host memory");
    }
    rp[0] = 0;
    for (long long r = 0; r < M; r++) rp[r + 1] = rp[r] + (int)cnt[r];
    if (sorted && !isSym) {
        // the CSR is the entries in file order (the scatter below would
        // place each entry at its own position)
        tc::parallel_for(NZ, [&](long long b, long long e) {
            memcpy(cc + b, ci.data() + b, (size_t)(e - b) * sizeof(int));
            memcpy(vv + b, vi.data() + b, (size_t)(e - b) * sizeof(float));
        });
        *m = (int)M;
        *n = (int)N;
        *nnzA = (unsigned long)total;
        *rowptr = rp;
        *colidx = cc;
        *val = vv;
        return 0;
    }
    std::vector<int> fillp(rp, rp + M);
    for (long long i = 0; i < NZ; i++) {
        int a = ri[i], b = ci[i];
        int o = fillp[a]++;
        cc[o] = b;
        vv[o] = vi[i];
        if (isSym && a != b) {
            o = fillp[b]++;
            cc[o] = a;
            vv[o] = vi[i];
        }
    }
    *m = (int)M;
    *n = (int)N;
    *nnzA = (unsigned long)total;
    *rowptr = rp;
    *colidx = cc;
    *val = vv;
    return 0;
} catch (...) {
    return tc::exception_status(__func__);
}

extern "C" int tc_write_mtx(const char *path, int M, int N, const int *rowptr, const int *col)
try {
    if (!path || M < 0 || N < 0 || !rowptr || (rowptr[M] > 0 && !col))
        return tc::fail(TC_ERR_INVALID, "tc_write_mtx: bad argument");
    FILE *f = fopen(path, "w");
    if (!f) return tc::fail(-1, "cannot create %s", path);
    fprintf(f, "%%%%MatrixMarket matrix coordinate pattern general\n%d %d %d\n", M, N, rowptr[M]);
    for (int r = 0; r < M; r++)
        for (int j = rowptr[r]; j < rowptr[r + 1]; j++) fprintf(f, "%d %d\n", r + 1, col[j] + 1);
    int bad = ferror(f);
    fclose(f);
    return bad ? tc::fail(-1, "write error on %s", path) : TC_OK;
} catch (...) {
    return tc::exception_status(__func__);
}
