// Serving runtime: many TaiChi matrices resident in HBM, SpMV requests on
// any of them continuously batched into persistent-kernel launches. No
// reference counterpart (the reference runs one SpMV per process,
// TaiChi.cu:1331-1456).
//
//   * Admission by footprint: a matrix is admitted only if the device bytes
//     its format needs (tc_matrix_footprint, computed before any
//     allocation) fit under the configured HBM cap.
//   * Continuous batching: a dispatcher thread takes whatever requests are
//     pending (up to max_batch, and up to max_inflight requests on the GPU)
//     as soon as one of `streams` launch slots is free and enqueues them on
//     that slot's stream as one batched launch group (one kernel per column
//     chunk class, every request's tiles in it), whose work descriptors go
//     through the slot's pinned buffer. Requests that arrive while batches
//     run wait only for the next free slot, and batches on different streams
//     overlap on the device.
//   * A completion thread polls the oldest in-flight batch's event and
//     completes its requests; latency is arrival -> completion observed.
//   * Results do not depend on batching: the batched kernel computes every
//     row in the same canonical order as tc_spmv.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch.h"
#include "common.h"

namespace {

int64_t now_ns()
{
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct ReqRec {
    unsigned long long ticket = 0;
    int matrix_id = 0;
    tc_request r{};
    int64_t t_arrive = 0;
    int64_t t_done = 0;
    int status = TC_OK;
    bool done = false;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    void *host_desc = nullptr;
    void *dev_desc = nullptr;
    size_t desc_cap = 0;  // bytes of each descriptor buffer
    std::vector<ReqRec *> reqs;
    bool busy = false;
};

struct Mat {
    tc_matrix *A = nullptr;
    long long bytes = 0;
    int outstanding = 0;
};

}  // namespace

struct tc_server {
    tc_server_config cfg{};
    int device = 0;
    std::mutex mu;
    std::condition_variable cv_disp, cv_comp, cv_wait;
    std::deque<ReqRec *> pending;
    std::vector<Slot> slots;
    std::deque<int> inflight_order;
    int inflight_reqs = 0;
    std::unordered_map<int, Mat> mats;
    std::unordered_map<unsigned long long, ReqRec *> live;
    long long resident = 0;
    int next_id = 1;
    unsigned long long next_ticket = 1;
    long long n_done = 0, n_batches = 0, n_batched_reqs = 0;
    bool stop = false;
    bool disp_done = false;  // the dispatcher has exited (nothing more to enqueue)
    std::thread disp, comp;
};

namespace {

void dispatcher(tc_server *S)
{
    (void)hipSetDevice(S->device);
    std::vector<tc_request> reqs;
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        int si = -1;
        S->cv_disp.wait(lk, [&] {
            if (S->stop && S->pending.empty()) return true;
            if (S->pending.empty() || S->inflight_reqs >= S->cfg.max_inflight) return false;
            for (size_t i = 0; i < S->slots.size(); i++)
                if (!S->slots[i].busy) return true;
            return false;
        });
        if (S->stop && S->pending.empty()) break;
        for (size_t i = 0; i < S->slots.size(); i++)
            if (!S->slots[i].busy) {
                si = (int)i;
                break;
            }
        Slot &sl = S->slots[si];
        const int take = std::min<int>({(int)S->pending.size(), S->cfg.max_batch,
                                        S->cfg.max_inflight - S->inflight_reqs});
        sl.reqs.assign(S->pending.begin(), S->pending.begin() + take);
        S->pending.erase(S->pending.begin(), S->pending.begin() + take);
        sl.busy = true;
        S->inflight_reqs += take;
        lk.unlock();
        reqs.resize(take);
        for (int i = 0; i < take; i++) reqs[i] = sl.reqs[i]->r;
        // the whole batch is one batched launch group (coalesced: one launch
        // per column-chunk class, every request's tiles in it); launch mode 2
        // runs the requests one by one on the slot's stream instead. On the
        // 64-matrix serving mix both give the same throughput (85.8k against
        // 85.4k req/s), and the coalesced form launches a few kernels a
        // batch instead of one a request.
        // A batch of one request (most of them at light load) launches it
        // as tc_spmv: no descriptor copy ahead of its kernel.
        int rc = TC_OK;
        int nsmall = 0;
        const bool alone = S->cfg.launch == 2 || take == 1;
        for (int i = 0; i < take && rc == TC_OK; i++) {
            const tc_request &q = reqs[i];
            if (alone) rc = tc_spmv(q.A, q.X, q.ldx, q.dtype, q.Y, q.ldy, q.k, sl.stream);
            else reqs[nsmall++] = q;
        }
        if (rc == TC_OK && nsmall > 0) rc = tc::batch_enqueue(reqs.data(), nsmall, sl.host_desc, sl.dev_desc, sl.stream, sl.desc_cap);
        if (rc == TC_OK && hipEventRecord(sl.ev, sl.stream) != hipSuccess) rc = TC_ERR_HIP;
        lk.lock();
        if (rc != TC_OK)
            for (ReqRec *q : sl.reqs) q->status = rc;
        S->n_batches++;
        S->n_batched_reqs += take;
        S->inflight_order.push_back(si);
        S->cv_comp.notify_all();
    }
    S->disp_done = true;
    S->cv_comp.notify_all();
}

void completer(tc_server *S)
{
    (void)hipSetDevice(S->device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        S->cv_comp.wait(lk, [&] { return !S->inflight_order.empty() || S->disp_done; });
        if (S->inflight_order.empty()) {
            if (S->disp_done) break;
            continue;
        }
        const int si = S->inflight_order.front();
        Slot &sl = S->slots[si];
        lk.unlock();
        // poll: completion is observed within microseconds of the kernel end
        if (sl.reqs.empty() || sl.reqs[0]->status == TC_OK) {
            for (;;) {
                hipError_t e = hipEventQuery(sl.ev);
                if (e != hipErrorNotReady) break;
                std::this_thread::yield();
            }
        }
        const int64_t t = now_ns();
        lk.lock();
        for (ReqRec *q : sl.reqs) {
            q->t_done = t;
            q->done = true;
            S->mats[q->matrix_id].outstanding--;
        }
        S->n_done += (long long)sl.reqs.size();
        S->inflight_reqs -= (int)sl.reqs.size();
        sl.reqs.clear();
        sl.busy = false;
        S->inflight_order.pop_front();
        S->cv_disp.notify_all();
        S->cv_wait.notify_all();
    }
}

int submit_at(tc_server *S, int matrix_id, const void *X, int ldx, int dtype, float *Y, int ldy, int k,
              int64_t t_arrive, unsigned long long *ticket)
{
    std::lock_guard<std::mutex> g(S->mu);
    auto it = S->mats.find(matrix_id);
    if (it == S->mats.end()) return tc::fail(TC_ERR_INVALID, "tc_server_submit: unknown matrix %d", matrix_id);
    if (k < 1 || k > 64 || ldx < k || ldy < k || !X || !Y)
        return tc::fail(TC_ERR_INVALID, "tc_server_submit: bad request (k=%d)", k);
    if (dtype != TC_F32 && dtype != TC_BF16 && dtype != TC_F16) return tc::fail(TC_ERR_INVALID, "bad dtype");
    if (S->stop) return tc::fail(TC_ERR_STATE, "tc_server_submit: server stopping");
    ReqRec *q = new ReqRec();
    q->ticket = S->next_ticket++;
    q->matrix_id = matrix_id;
    q->r = tc_request{it->second.A, X, ldx, dtype, Y, ldy, k};
    q->t_arrive = t_arrive;
    it->second.outstanding++;
    S->pending.push_back(q);
    S->live[q->ticket] = q;
    if (ticket) *ticket = q->ticket;
    S->cv_disp.notify_one();
    return TC_OK;
}

// waits for a ticket, removes it, returns its status and latency (ns)
int wait_ticket(tc_server *S, unsigned long long ticket, int64_t *lat_ns)
{
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->live.find(ticket);
    if (it == S->live.end()) return tc::fail(TC_ERR_INVALID, "unknown ticket %llu", ticket);
    ReqRec *q = it->second;
    S->cv_wait.wait(lk, [&] { return q->done; });
    if (lat_ns) *lat_ns = q->t_done - q->t_arrive;
    const int st = q->status;
    S->live.erase(ticket);  // by key: inserts while waiting may have rehashed
    delete q;
    return st;
}

double pct(std::vector<double> v, double p)
{
    if (v.empty()) return 0;
    std::sort(v.begin(), v.end());
    const double idx = p * (double)(v.size() - 1);
    const size_t lo = (size_t)std::floor(idx), hi = std::min(v.size() - 1, lo + 1);
    return v[lo] + (v[hi] - v[lo]) * (idx - (double)lo);
}

void fill_result(const std::vector<double> &lat_us, double secs, long long n_batches, long long n_batched,
                 tc_load_result *res)
{
    res->completed = (long long)lat_us.size();
    res->seconds = secs;
    res->req_per_s = secs > 0 ? (double)lat_us.size() / secs : 0;
    res->p50_us = pct(lat_us, 0.50);
    res->p99_us = pct(lat_us, 0.99);
    double s = 0, mx = 0;
    for (double v : lat_us) {
        s += v;
        mx = std::max(mx, v);
    }
    res->mean_us = lat_us.empty() ? 0 : s / (double)lat_us.size();
    res->max_us = mx;
    res->mean_batch = n_batches > 0 ? (double)n_batched / (double)n_batches : 0;
}

}  // namespace

#define HIPCHECK(expr)                                                                                      \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess) return tc::fail(TC_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));          \
    } while (0)

extern "C" {

int tc_server_create(const tc_server_config *cfg, tc_server **out)
try {
    if (!cfg || !out || cfg->max_batch < 1 || cfg->streams < 1 || cfg->max_inflight < 1 || cfg->hbm_cap_bytes < 0 ||
        cfg->launch < 0 || cfg->launch > 2)
        return tc::fail(TC_ERR_INVALID, "tc_server_create: bad config");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return tc::fail(TC_ERR_NODEVICE, "no HIP device");
    tc_server *S = new tc_server();
    S->cfg = *cfg;
    (void)hipGetDevice(&S->device);
    S->slots.resize(cfg->streams);
    const size_t db = tc::batch_desc_bytes(cfg->max_batch);
    for (Slot &sl : S->slots) {
        sl.desc_cap = db;
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc(&sl.host_desc, db, hipHostMallocDefault) != hipSuccess ||
            hipMalloc(&sl.dev_desc, db) != hipSuccess) {
            tc_server_destroy(S);
            return tc::fail(TC_ERR_HIP, "tc_server_create: stream/event/buffer allocation failed");
        }
    }
    S->disp = std::thread(dispatcher, S);
    S->comp = std::thread(completer, S);
    *out = S;
    return TC_OK;
} catch (...) {
    return tc::exception_status(__func__);
}

int tc_server_destroy(tc_server *S)
{
    if (!S) return TC_OK;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
        S->cv_disp.notify_all();
        S->cv_comp.notify_all();
    }
    if (S->disp.joinable()) S->disp.join();
    if (S->comp.joinable()) S->comp.join();
    for (Slot &sl : S->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.host_desc) (void)hipHostFree(sl.host_desc);
        if (sl.dev_desc) (void)hipFree(sl.dev_desc);
    }
    for (auto &kv : S->mats) tc_matrix_destroy(kv.second.A);
    for (auto &kv : S->live) delete kv.second;
    delete S;
    return TC_OK;
}

int tc_server_add_matrix(tc_server *S, int M, int N, const TaiChi_new *fmt, long long nnz, int *id)
try {
    if (!S || !fmt || !id) return tc::fail(TC_ERR_INVALID, "tc_server_add_matrix: bad argument");
    long long bytes = 0;
    int rc = tc_matrix_footprint(M, N, fmt, 0, &bytes);
    if (rc != TC_OK) return rc;
    {
        std::lock_guard<std::mutex> g(S->mu);
        if (S->resident + bytes > S->cfg.hbm_cap_bytes)
            return tc::fail(TC_ERR_CAPACITY, "tc_server_add_matrix: %lld + %lld bytes over the cap of %lld",
                            S->resident, bytes, S->cfg.hbm_cap_bytes);
        S->resident += bytes;  // reserved before the upload
    }
    tc_matrix *A = nullptr;
    rc = tc_matrix_create(M, N, fmt, nnz, 0, nullptr, &A);
    std::lock_guard<std::mutex> g(S->mu);
    if (rc != TC_OK) {
        S->resident -= bytes;
        return rc;
    }
    const int mid = S->next_id++;
    S->mats[mid] = Mat{A, bytes, 0};
    *id = mid;
    return TC_OK;
} catch (...) {
    return tc::exception_status(__func__);
}

int tc_server_remove_matrix(tc_server *S, int id)
try {
    if (!S) return tc::fail(TC_ERR_INVALID, "tc_server_remove_matrix: null server");
    tc_matrix *A = nullptr;
    {
        std::lock_guard<std::mutex> g(S->mu);
        auto it = S->mats.find(id);
        if (it == S->mats.end()) return tc::fail(TC_ERR_INVALID, "unknown matrix %d", id);
        if (it->second.outstanding > 0)
            return tc::fail(TC_ERR_STATE, "matrix %d has %d outstanding requests", id, it->second.outstanding);
        A = it->second.A;
        S->resident -= it->second.bytes;
        S->mats.erase(it);
    }
    return tc_matrix_destroy(A);
} catch (...) {
    return tc::exception_status(__func__);
}

int tc_server_submit(tc_server *S, int matrix_id, const void *X, int ldx, int dtype, float *Y, int ldy, int k,
                     unsigned long long *ticket)
try {
    if (!S) return tc::fail(TC_ERR_INVALID, "tc_server_submit: null server");
    return submit_at(S, matrix_id, X, ldx, dtype, Y, ldy, k, now_ns(), ticket);
} catch (...) {
    return tc::exception_status(__func__);
}

int tc_server_wait(tc_server *S, unsigned long long ticket, double *latency_us)
try {
    if (!S) return tc::fail(TC_ERR_INVALID, "tc_server_wait: null server");
    int64_t ns = 0;
    int rc = wait_ticket(S, ticket, &ns);
    if (latency_us) *latency_us = (double)ns * 1e-3;
    return rc;
} catch (...) {
    return tc::exception_status(__func__);
}

int tc_server_get_stats(tc_server *S, tc_server_stats *out)
try {
    if (!S || !out) return tc::fail(TC_ERR_INVALID, "tc_server_get_stats: bad argument");
    std::lock_guard<std::mutex> g(S->mu);
    out->resident_bytes = S->resident;
    out->matrices = (int)S->mats.size();
    out->pending = (int)S->pending.size();
    out->inflight = S->inflight_reqs;
    out->completed = S->n_done;
    out->batches = S->n_batches;
    out->mean_batch = S->n_batches ? (double)S->n_batched_reqs / (double)S->n_batches : 0.0;
    return TC_OK;
} catch (...) {
    return tc::exception_status(__func__);
}

int tc_server_run_load(tc_server *S, const tc_load_spec *sp, tc_load_result *res)
try {
    if (!S || !sp || !res || !sp->items || sp->nitems < 1 || !sp->ybufs || sp->nybufs < 1 || sp->count < 1)
        return tc::fail(TC_ERR_INVALID, "tc_server_run_load: bad spec");
    const bool open = sp->rate > 0;
    const long long n = sp->count;
    const int window = open ? sp->nybufs : std::max(1, std::min(sp->concurrency, sp->nybufs));
    const uint64_t key = tc::rand_key(sp->seed, 11);
    std::vector<unsigned long long> tickets((size_t)n, 0);
    std::vector<double> lat_us;
    lat_us.reserve((size_t)n);
    long long b0, r0;
    {
        std::lock_guard<std::mutex> g(S->mu);
        b0 = S->n_batches;
        r0 = S->n_batched_reqs;
    }
    const int64_t t0 = now_ns();
    double t_next = 0;  // seconds after t0 of the next arrival
    long long waited = 0;
    int rc = TC_OK;
    for (long long i = 0; i < n && rc == TC_OK; i++) {
        // the request that used this output buffer (or window slot) must be done
        if (i >= window) {
            int64_t ns = 0;
            rc = wait_ticket(S, tickets[(size_t)(i - window)], &ns);
            lat_us.push_back((double)ns * 1e-3);
            waited = i - window + 1;
            if (rc != TC_OK) break;
        }
        int64_t t_arr = now_ns();
        if (open) {
            const double u = ((double)(tc::rand_at(key, 2 * (uint64_t)i) >> 11) + 1.0) * (1.0 / 9007199254740992.0);
            t_next += -std::log(u) / sp->rate;
            t_arr = t0 + (int64_t)(t_next * 1e9);
            while (now_ns() < t_arr) std::this_thread::yield();
        }
        const tc_load_item &it = sp->items[tc::bounded(tc::rand_at(key, 2 * (uint64_t)i + 1), (uint64_t)sp->nitems)];
        rc = submit_at(S, it.matrix_id, it.X, it.ldx, it.dtype, sp->ybufs[i % sp->nybufs], sp->ldy > 0 ? sp->ldy : it.ldx,
                       it.k, t_arr,
                       &tickets[(size_t)i]);
    }
    for (long long i = waited; i < n && tickets[(size_t)i]; i++) {
        int64_t ns = 0;
        int r2 = wait_ticket(S, tickets[(size_t)i], &ns);
        if (rc == TC_OK) rc = r2;
        lat_us.push_back((double)ns * 1e-3);
    }
    const double secs = (double)(now_ns() - t0) * 1e-9;
    long long b1, r1;
    {
        std::lock_guard<std::mutex> g(S->mu);
        b1 = S->n_batches;
        r1 = S->n_batched_reqs;
    }
    fill_result(lat_us, secs, b1 - b0, r1 - r0, res);
    return rc;
} catch (...) {
    return tc::exception_status(__func__);
}

int tc_server_run_plain(tc_server *S, const tc_load_spec *sp, tc_load_result *res)
try {
    if (!S || !sp || !res || !sp->items || sp->nitems < 1 || !sp->ybufs || sp->nybufs < 1 || sp->count < 1)
        return tc::fail(TC_ERR_INVALID, "tc_server_run_plain: bad spec");
    hipStream_t s;
    HIPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    const uint64_t key = tc::rand_key(sp->seed, 11);
    std::vector<double> lat_us;
    lat_us.reserve((size_t)sp->count);
    const int64_t t0 = now_ns();
    int rc = TC_OK;
    for (long long i = 0; i < sp->count && rc == TC_OK; i++) {
        const tc_load_item &it = sp->items[tc::bounded(tc::rand_at(key, 2 * (uint64_t)i + 1), (uint64_t)sp->nitems)];
        tc_matrix *A = nullptr;
        {
            std::lock_guard<std::mutex> g(S->mu);
            auto m = S->mats.find(it.matrix_id);
            if (m == S->mats.end()) {
                rc = tc::fail(TC_ERR_INVALID, "unknown matrix %d", it.matrix_id);
                break;
            }
            A = m->second.A;
        }
        const int64_t ta = now_ns();
        rc = tc_spmv(A, it.X, it.ldx, it.dtype, sp->ybufs[i % sp->nybufs], sp->ldy > 0 ? sp->ldy : it.ldx, it.k, s);
        if (rc == TC_OK && hipStreamSynchronize(s) != hipSuccess) rc = TC_ERR_HIP;
        lat_us.push_back((double)(now_ns() - ta) * 1e-3);
    }
    const double secs = (double)(now_ns() - t0) * 1e-9;
    (void)hipStreamDestroy(s);
    fill_result(lat_us, secs, (long long)lat_us.size(), (long long)lat_us.size(), res);
    return rc;
} catch (...) {
    return tc::exception_status(__func__);
}

}  // extern "C"
