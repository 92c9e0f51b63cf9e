/*
 * taichi_amd.h -- C ABI of the MI355X-native TaiChi binary-SpMV library.
 *
 * Plain C: pointers, sizes and status codes, no C++ or torch types. Device
 * pointers are HIP device addresses; "stream" arguments are hipStream_t
 * passed as void* (NULL = the null stream). Every function returns TC_OK (0)
 * or a negative status; tc_last_error() returns the calling thread's last
 * message. The Matrix Market reader keeps the reference's own return codes
 * -1..-4 (TaiChi.cu:130-160).
 *
 * Entry points and the reference interface each one replaces:
 *   tc_read_mtx ............ readMtx            TaiChi.h:79-80, TaiChi.cu:120-278
 *   tc_thumb_dims .......... thumbnail sizing   morohutils.cu:191-200
 *   tc_thumbnail ........... getSdf_GPU1        morohutils.cu:153-180, :204-229
 *   tc_sdf_from_image ...... getSdf morphology  morohutils.cu:231-282
 *   tc_partition_from_image  partition_shapes   TaiChi.h:83-84 (given the image)
 *   tc_partition_shapes .... partition_shapes   TaiChi.h:83-84, TaiChi.cu:313-638
 *   tc_partition_device .... partition_dia_shape's passes, TaiChi.cu:677-879 (device)
 *   tc_morph_device ........ getSdf's morphology   morohutils.cu:231-251 (device)
 *   tc_cal_shape_size ...... cal_shape_size     TaiChi.h:87-88, TaiChi.cu:280-311
 *   tc_cal_start_end ....... cal_start_end      TaiChi.h:103, TaiChi.cu:881-916
 *   tc_free_taichi_new ..... free_taichi_new    TaiChi.h:106, TaiChi.cu:77-118
 *   tc_taichi_spmv ......... taichi_SpMV        TaiChi.h:95, TaiChi.cu:1331-1456
 *   tc_matrix_create/tc_spmv the upload + kernel launch inside taichi_SpMV
 *                            (TaiChi.cu:1356-1417), split so the format stays
 *                            resident in HBM across calls
 * Entry points with no reference counterpart (multi-RHS, batched persistent
 * launch, graphs, autotuner, serving, generators) say so where declared.
 */
#ifndef TAICHI_AMD_H
#define TAICHI_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TC_ABI_VERSION 2

/* ---- status codes (the reader's -1..-4 are the reference's own) ---- */
#define TC_OK 0
#define TC_ERR_INVALID (-101)      /* bad argument or shape */
#define TC_ERR_HIP (-102)          /* HIP runtime error */
#define TC_ERR_NOMEM (-103)        /* host or device allocation failed */
#define TC_ERR_UNSUPPORTED (-104)  /* valid but not supported (e.g. k > 64) */
#define TC_ERR_CAPACITY (-105)     /* admission refused: over the byte cap */
#define TC_ERR_NODEVICE (-106)     /* no HIP device visible */
#define TC_ERR_STATE (-107)        /* object in the wrong state */

/* storage type of x (y is always FP32, accumulation is always FP32) */
#define TC_F32 0
#define TC_BF16 1
#define TC_F16 2

/* The reference's persistent format, same field order (TaiChi.h:45-60).
   csrCol holds (col << 1) | sign when dense_nnz > 0 (sign 1 = +x, 0 = -x),
   plain column indices when dense_nnz == 0 (then csrRow/csrCol alias the
   input CSR, as at TaiChi.cu:575-578). */
struct TaiChi_new {
    int dia_shapes;     /* S */
    int num_negs;       /* D */
    int dense_nnz;
    int zero_elements;  /* Z */
    int sparse_nnz;     /* P */
    int row_start;
    int row_stop;
    int *xStart;        /* [S] */
    int *yStart;        /* [S] */
    int *xStop;         /* [S] */
    int *nDiasPtr;      /* [S+1] */
    int *neg_offsets;   /* [D] */
    int *csrRow;        /* [M+1] */
    int *csrCol;        /* [Z+P] */
};
typedef struct TaiChi_new TaiChi_new;

const char *tc_status_string(int status);
const char *tc_last_error(void);
int tc_abi_version(void);
/* number of visible HIP devices, or a negative status */
int tc_device_count(void);

/* ---- a1: Matrix Market ---- */
/* readMtx: arrays are malloc'd; free with tc_free_host. Returns 0, -1 (open),
   -2 (banner), -3 (complex), -4 (size line or data). */
int tc_read_mtx(const char *path, int *m, int *n, unsigned long *nnz,
                int **rowptr, int **col, float **val);
void tc_free_host(void *p);
/* writes a 'coordinate pattern general' file in CSR order (no reference
   counterpart: fixtures and format round trips) */
int tc_write_mtx(const char *path, int M, int N, const int *rowptr, const int *col);

/* ---- a2: density thumbnail ---- */
/* returns 0, 1 when M <= 512 (no thumbnail), 2 when the width rounds to 0 */
int tc_thumb_dims(int M, int N, int *H, int *W, int *pM, int *pN);
/* kernel (a): d_img[H*W] = 0 where a nonzero lands, 255 elsewhere.
   d_rowptr[M+1], d_col[nnz] are device arrays. */
int tc_thumbnail(int M, int N, const int *d_rowptr, const int *d_col,
                 unsigned char *d_img, void *stream);

/* ---- a3..a6: partition ---- */
/* getSdf's image pipeline: writes up to cap shapes as x1,y1,x2,y2 into pts
   (4 ints each) and their count into *num_shapes */
int tc_sdf_from_image(const unsigned char *img, int H, int W, int *pts, int cap,
                      int *num_shapes);
/* findContours(RETR_CCOMP, CHAIN_APPROX_NONE) of a binary image, flattened as
   [n, then per contour: len, is_hole, x0, y0, x1, y1, ...] in OpenCV's order
   (no reference counterpart: exposes the step inside getSdf for checking).
   *needed receives the int count; TC_ERR_CAPACITY when cap is too small. */
int tc_contours(const unsigned char *img, int H, int W, int *out, int cap, int *needed);
/* getSdf's morphology alone (morohutils.cu:231-251: 255 - mat, imfill, open
   with the 30x1 and 1x30 lines, saturated subtraction, open with eye(5)):
   thin[H*W] is the binary image whose contours give the shapes. */
int tc_morph_from_image(const unsigned char *img, int H, int W, unsigned char *thin);
/* The same morphology on the device (d_img, d_thin device arrays of H*W);
   work: device scratch of tc_morph_work_bytes(H, W) bytes. Synchronises the
   stream. */
size_t tc_morph_work_bytes(int H, int W);
int tc_morph_device(const unsigned char *d_img, int H, int W, unsigned char *d_thin,
                    void *work, void *stream);
void tc_cal_shape_size(int M, int N, int mF, int x1, int y1, int a1, int b1,
                       int *x, int *y, int *a, int *b, float *gain);
void tc_cal_start_end(int M_sub, int N_sub, int width, int off, int n,
                      int *start, int *end);
/* partition + format build from a given thumbnail (img may be NULL when
   M <= 512). val may be NULL (binary input); when given it is mutated like the
   reference's (FLT_MAX on claimed entries). out must be zeroed. */
int tc_partition_from_image(int M, int N, int nnz, int *rowptr, int *col,
                            float *val, const unsigned char *img, TaiChi_new *out);
/* partition_shapes: uploads the CSR and builds the format on the device
   (tc_partition_device); with val given, or when the device build returns
   TC_ERR_UNSUPPORTED, runs tc_thumbnail on the device and then
   tc_partition_from_image. path is accepted for signature parity (unused). */
int tc_partition_shapes(const char *path, int M, int N, int nnz, int *rowptr,
                        int *col, float *val, TaiChi_new *out);
/* Device build of partition_shapes for a binary CSR (the histogram, claim and
   zero-row passes of partition_dia_shape, TaiChi.cu:723-734, 809-827, 850-871,
   and the merge :480-612, as kernels). d_rowptr/d_col are device copies of
   rowptr/col; the format's arrays come back in host memory, identical to
   tc_partition_from_image's (with no dense diagonal they alias rowptr/col).
   TC_ERR_UNSUPPORTED when an entry is duplicated on a dense diagonal (the
   reference's zero-row walk then depends on the duplicate) or the format
   would pass 2^31 entries. */
int tc_partition_device(int M, int N, int nnz, int *rowptr, int *col,
                        const int *d_rowptr, const int *d_col, TaiChi_new *out,
                        void *stream);
/* zeroed malloc'd struct, for callers that cannot malloc (FFI) */
TaiChi_new *tc_taichi_new_alloc(void);
/* free_taichi_new: frees the arrays (csrRow/csrCol only when dense_nnz > 0)
   and the struct itself */
int tc_free_taichi_new(TaiChi_new *t);

/* ---- format cache (no reference counterpart; the reference re-reads the
   .mtx text and re-partitions on every run, TaiChi.cu:120-278, 313-638) ----
   tc_save_format writes a partitioned TaiChi_new with its sizes and an
   FNV-1a checksum; tc_load_format validates magic, sizes, checksum and the
   row pointers, then fills *t with malloc'd arrays: free them with
   tc_free_taichi_new on a heap struct (tc_taichi_new_alloc) -- which, like
   free_taichi_new, leaves csrRow/csrCol alone when dense_nnz == 0: free
   those two with tc_free_host. Bad files return TC_ERR_INVALID. */
int tc_save_format(const char *path, int M, int N, const TaiChi_new *t);
int tc_load_format(const char *path, int *M, int *N, TaiChi_new *t);

/* ---- a7: one-shot host SpMV (taichi_SpMV) ---- */
/* uploads the format and x, runs `warmup` + `runs` launches (10 + 500 in the
   reference), downloads y; optional timings in ms (NULL to skip) */
int tc_taichi_spmv(int M, int N, const TaiChi_new *t, const float *x_host,
                   float *y_host, int warmup, int runs, float *transfer_ms,
                   float *spmv_ms);

/* ---- resident matrices: the hot path ---- */
typedef struct tc_matrix tc_matrix;

typedef struct tc_matrix_stats {
    int M, N;
    int dia_shapes, num_negs;
    int long_rows;          /* rows handled by a wavefront each */
    int offset_bytes;       /* 1 or 2: width of the per-row local offsets */
    long long nnz;          /* nnz of A as given at create (duplicates removed by the caller) */
    long long remainder;    /* Z + P entries stored in cols */
    long long format_bytes; /* bytes of every device array of the format */
    long long device_bytes; /* bytes allocated on the device for this matrix */
    long long holes;        /* Z entries (cancelled zero cells of dense diagonals)
                               held as cleared bits of a shape's term-presence
                               bitmap instead of in cols */
    int bitmap_shapes;      /* shapes with a term-presence bitmap */
    int near_window;        /* 1: most remainder entries (24 or more a row) lie
                               within 64 rows of their row, so the k > 1 tiles
                               (k <= 8) read them from x rows staged in LDS */
} tc_matrix_stats;

/* upload a TaiChi_new (host) into the device layout. nnz is the count used
   for GNNZ/s; long_threshold <= 0 picks the default (64). */
int tc_matrix_create(int M, int N, const TaiChi_new *t, long long nnz,
                     int long_threshold, void *stream, tc_matrix **out);
int tc_matrix_destroy(tc_matrix *A);
/* device bytes tc_matrix_create would allocate for this format (host-only
   computation, no device needed; used for admission by footprint) */
int tc_matrix_footprint(int M, int N, const TaiChi_new *t, int long_threshold, long long *bytes);
int tc_matrix_get_stats(const tc_matrix *A, tc_matrix_stats *out);
/* algorithmic HBM bytes of one SpMV: format arrays once, x once (k columns of
   the given storage type), y written once */
long long tc_spmv_bytes(const tc_matrix *A, int k, int dtype);

/* Y[M x k] = A * X[N x k]; X row-major with leading dimension ldx (elements),
   Y row-major with ldy. k in 1..64. Asynchronous on `stream`. */
int tc_spmv(tc_matrix *A, const void *X, int ldx, int dtype, float *Y, int ldy,
            int k, void *stream);

/* launch configuration (no reference counterpart; the reference's TPV tier,
   TaiChi.cu:1390-1394, plays this role): rows per workgroup tile (k = 1:
   256, 512, 1024, 2048 or 4096 = 1..16 rows per lane; k > 1, wide rows:
   256*c/k*R for R in 1, 2 rows per lane group, c the columns a lane owns --
   4, or 8 for BF16/FP16 x at k >= 8 (2 for k = 2); k > 1, staged tiles
   (flag bit 2): 256*P/k for P in 1, 2, 4), LDS value capacity, persistent
   grid size (0 = one workgroup per tile), kernel choice. Configurations are
   kept per (k, x type). Results do not depend on the configuration. */
typedef struct tc_config {
    int rows_per_tile;   /* power of two, 1..2048; 0 = default for k */
    int lds_values;      /* LDS staging capacity in floats (512..8192); 0 = default */
    int grid;            /* 0 = one workgroup per tile */
    int flags;           /* bit 0: reserved (ignored). bit 1: at k = 1 use the
                            one-row-per-lane kernel, not quad rows (quad rows
                            apply to 512/1024-row tiles of matrices with at
                            most one shape of one offset). bit 2: k > 1 on the
                            staged tile kernel. bit 3 (in the config of width
                            4 or 8): no masked chunks -- 3, 5, 6, 7 columns
                            then run as narrower chunks (or the staged
                            kernel) instead of one masked pass of that
                            width */
} tc_config;
int tc_matrix_set_config(tc_matrix *A, int k, int dtype, const tc_config *cfg);
int tc_matrix_get_config(const tc_matrix *A, int k, int dtype, tc_config *cfg);
/* autotuner (no reference counterpart): times candidate configs on the given
   buffers, keeps the fastest for (k, dtype); writes it and its time */
int tc_autotune(tc_matrix *A, const void *X, int ldx, int dtype, float *Y, int ldy,
                int k, void *stream, tc_config *best, float *best_us);

/* ---- batched persistent launch (no reference counterpart) ---- */
typedef struct tc_request {
    tc_matrix *A;
    const void *X;
    int ldx;
    int dtype;
    float *Y;
    int ldy;
    int k;      /* 1..64 */
} tc_request;
/* batched SpMV over n requests on any mix of matrices: the requests are cut
   into column chunks of 8, 4, 2 and 1 columns and each (chunk width, x type)
   group is one kernel launch over all its tiles; each y is bit-identical to
   what tc_spmv computes for that request alone. Asynchronous on `stream`:
   the descriptors go through buffers the library owns (8 slots a device,
   pinned host and device memory, each reused once its batch has run); the
   call returns without waiting for the stream. */
int tc_spmv_batch(const tc_request *reqs, int n, void *stream);
/* tc_spmv_batch without any allocation. No counterpart in the reference: it
   is for callers that cannot allocate (code under HIP graph capture, tight
   serving loops, harnesses that forbid hipMalloc). The descriptor table lives
   in a fixed 1 MiB array of device memory that each GPU the library runs on
   holds until the process ends, staged through a static host buffer; a batch
   whose descriptors need more (about 5,000 column chunks of at most 8
   columns) is refused with TC_ERR_CAPACITY before anything is enqueued. All
   calls on one GPU share that array and it keeps the last call's
   descriptors: calls that overlap on one GPU (on different streams, before
   the previous batch has run) race; calls on one stream are ordered. Same
   results as tc_spmv_batch. */
int tc_spmv_batch_static(const tc_request *reqs, int n, void *stream);

/* ---- serving (no reference counterpart): resident matrices, admission by
   HBM footprint, continuous batching into persistent launches ---- */
typedef struct tc_server tc_server;
typedef struct tc_server_config {
    long long hbm_cap_bytes;  /* admission cap on the resident formats' bytes */
    int max_batch;            /* requests per launch */
    int streams;              /* launch slots that may run concurrently */
    int max_inflight;         /* requests on the GPU at once */
    int launch;               /* how a batch is launched on its slot's stream:
                                 0 and 1: as one batched launch group (the
                                 requests coalesced, tc_spmv_batch's kernels),
                                 a batch of one request as tc_spmv;
                                 2: one by one (tc_spmv). Same results either
                                 way. */
} tc_server_config;
typedef struct tc_server_stats {
    long long resident_bytes;
    int matrices, pending, inflight;
    long long completed, batches;
    double mean_batch;
} tc_server_stats;
int tc_server_create(const tc_server_config *cfg, tc_server **out);
/* drains pending work, then frees everything, resident matrices included */
int tc_server_destroy(tc_server *s);
/* uploads a format if its footprint fits under the cap (TC_ERR_CAPACITY
   otherwise, nothing allocated); *id names it in requests */
int tc_server_add_matrix(tc_server *s, int M, int N, const TaiChi_new *t, long long nnz, int *id);
/* TC_ERR_STATE while requests on it are pending or in flight */
int tc_server_remove_matrix(tc_server *s, int id);
/* Y[M x k] = A_id * X[N x k] (k in 1..64); X and Y are device buffers owned
   by the caller and left untouched by others until the ticket completes */
int tc_server_submit(tc_server *s, int matrix_id, const void *X, int ldx, int dtype, float *Y,
                     int ldy, int k, unsigned long long *ticket);
/* blocks until the ticket completes; latency = submit -> completion seen */
int tc_server_wait(tc_server *s, unsigned long long ticket, double *latency_us);
int tc_server_get_stats(tc_server *s, tc_server_stats *out);

/* load generation over resident matrices: each request picks a template
   uniformly (seeded) and writes y into ybufs[i % nybufs] (each large enough
   for any request, row stride ldy; ldy = 0 gives each request's y the row
   stride of its own x, ldx) */
typedef struct tc_load_item {
    int matrix_id;
    const void *X;
    int ldx;
    int dtype;
    int k;
} tc_load_item;
typedef struct tc_load_spec {
    const tc_load_item *items;
    int nitems;
    float *const *ybufs;
    int nybufs;
    int ldy;
    double rate;          /* > 0: Poisson arrivals at this rate (req/s); else closed loop */
    int concurrency;      /* closed loop: requests kept outstanding */
    long long count;      /* requests to issue */
    uint64_t seed;
} tc_load_spec;
typedef struct tc_load_result {
    long long completed;
    double seconds, req_per_s, p50_us, p99_us, mean_us, max_us, mean_batch;
} tc_load_result;
/* served: requests go through the server's continuous batching */
int tc_server_run_load(tc_server *s, const tc_load_spec *spec, tc_load_result *out);
/* plain baseline: one tc_spmv launch per request, each waited for */
int tc_server_run_plain(tc_server *s, const tc_load_spec *spec, tc_load_result *out);

/* ---- graphs (no reference counterpart) ---- */
typedef struct tc_graph tc_graph;
/* capture `iters` back-to-back tc_spmv launches into a hipGraph */
int tc_graph_capture(tc_matrix *A, const void *X, int ldx, int dtype, float *Y,
                     int ldy, int k, int iters, void *stream, tc_graph **out);
int tc_graph_launch(tc_graph *g, void *stream);
int tc_graph_destroy(tc_graph *g);

/* ---- iterative driver (no reference counterpart; the step after SpMV) ----
   PageRank by power iteration on a resident square binary matrix: A[i][j] =
   1 means j links to i; inv_deg[j] = 1/outdeg(j) or 0 for a page without
   out-links, whose rank is spread uniformly. From r = 1/n, `iters` steps of
   r <- (1 - alpha)/n + alpha (A (r .* inv_deg) + dangling/n), all on the
   device (each step: tc_spmv, a fused update, a fixed-order reduction).
   inv_deg, r (n floats, written) and work (tc_pagerank_work_bytes(n)) are
   device buffers. tc_pagerank enqueues on `stream`; tc_pagerank_capture
   records the same sequence once as a graph for tc_graph_launch (replays
   give identical ranks). */
size_t tc_pagerank_work_bytes(int n);
int tc_pagerank(tc_matrix *A, const float *inv_deg, float alpha, int iters, float *r, void *work,
                void *stream);
int tc_pagerank_capture(tc_matrix *A, const float *inv_deg, float alpha, int iters, float *r,
                        void *work, void *stream, tc_graph **out);

/* ---- synthetic generators (no reference counterpart; the reference reads
   downloaded datasets, README.md:17-20). Deterministic for a seed on any
   thread count. CSR arrays are malloc'd (tc_free_host); rows sorted, no
   duplicates. ---- */
/* ones on the given diagonal offsets plus ndraws uniform (r,c) draws;
   repeats and band hits dropped */
int tc_gen_diagonals_random(int n, const int *offsets, int noffsets, long long ndraws,
                            uint64_t seed, int **rowptr, int **col, long long *nnz);
/* R-MAT at 2^scale with quadrant probabilities a,b,c (d = 1-a-b-c); endpoints
   >= n are rejected (not redrawn); self loops kept; duplicates removed */
int tc_gen_rmat(int n, int scale, long long nedges, double a, double b, double c,
                uint64_t seed, int **rowptr, int **col, long long *nnz);
/* 5-point stencil on a gx x gy grid plus round(extra_frac * nnz) uniform draws */
int tc_gen_stencil5(int gx, int gy, double extra_frac, uint64_t seed,
                    int **rowptr, int **col, long long *nnz);
/* band of w diagonals, offsets -(w/2) .. w-1-(w/2), each cell 1 w.p. density */
int tc_gen_banded(int n, int w, double density, uint64_t seed,
                  int **rowptr, int **col, long long *nnz);
/* vectors: kind 0 = uniform [0,1), 1 = uniform integers {0..7},
   2 = N(0,1) (Box-Muller, float-rounded) */
int tc_gen_vector(float *out, long long n, int kind, uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif
