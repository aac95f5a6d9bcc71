// Served pms: a request loop that reduces arrivals of different sizes
// together on several device workers, with continuous batching at iteration
// granularity.
//
// The reference has no serving mode. Design:
//   * submit() validates a request's 1-based COO (so that a malformed one is
//     refused alone), records its distinct column lengths and its shape class
//     (input columns of at most 4 entries, or longer), copies it and queues
//     it; it never blocks on the GPU. A full queue refuses with OPENPH_EBUSY.
//     A COO the caller lends is not copied: submit queues it at once for
//     intake threads that validate several such requests in parallel and
//     pass them on; the worker reads it in place. Results go to the
//     request's own arrays or straight into buffers the caller gave.
//   * W worker threads each own a device workspace (Context), sized at start
//     for its byte budget, and a HIP stream. A worker keeps one running batch
//     (LiveBatch): a block-diagonal matrix whose members are reduced by the
//     same pms iterations. Between chunks of two iterations the worker
//       - retires every member whose lows no longer repeat (its own
//         convergence, from per-member duplicate flags of Phase I): its
//         result is final, it is copied out and its waiter released, and its
//         columns become inert;
//       - admits queued requests of the batch's shape class while they fit
//         the budget: columns, entries, pool space and at most
//         OPENPH_MAX_DIMS pseudo-dimensions over the batch. They are
//         appended to the running batch (set-up and Phase 0 on their own
//         columns) and take part from the next iteration on.
//     A worker with no running batch starts one with the oldest queued
//     request. Batches of different workers run concurrently on the device.
//     The host work of a worker overlaps its device work: while a chunk
//     runs, the outputs of the members the last chunk retired are delivered
//     and the next arrivals' inputs are copied to pinned memory; they are
//     uploaded and set up after the chunk, before the next one.
//   * A batch that fails (pool budget, iteration cap, runtime error) runs
//     each of its members again alone, so that one request cannot make
//     another fail; a request too large for an empty batch runs alone too.
//   * wait() blocks until a request is done and returns its outputs with its
//     submit / start / finish times.
// Results are unique for a matrix, and a member's columns never interact
// with another's, so a request's result is the same alone and in any batch
// (tests/test_gpu_server.py and tests/test_gpu_workloads.py check it).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "host.h"

namespace oph {

static double now_s() {
    using clk = std::chrono::steady_clock;
    static const clk::time_point t0 = clk::now();
    return std::chrono::duration<double>(clk::now() - t0).count();
}

// An allocator that leaves new elements uninitialised: a request's COO is
// written by the threads that copy it in, so its pages are first touched
// (and faulted in) by several threads instead of one zero-filling pass.
template <typename T>
struct NoInit : std::allocator<T> {
    template <typename U>
    struct rebind {
        using other = NoInit<U>;
    };
    NoInit() = default;
    template <typename U>
    NoInit(const NoInit<U> &) noexcept {}
    template <typename U>
    void construct(U *p) noexcept {
        ::new ((void *)p) U;
    }
    template <typename U, typename... A>
    void construct(U *p, A &&...a) {
        ::new ((void *)p) U(std::forward<A>(a)...);
    }
};

// Pinned host memory shared by the queued requests' COO. Submit (copied
// requests) and the intake threads (lent ones) place a request's rows and
// cols there as they validate it, so that its upload is a DMA with no host
// copy on the worker; the block is given back when the request finishes.
// First fit over a free list of 256-byte aligned blocks.
class Arena {
  public:
    static constexpr size_t kNone = ~(size_t)0;
    ~Arena() {
        if (base_) (void)hipHostFree(base_);
    }
    bool init(size_t bytes) {
        if (!bytes) return false;
        if (hipHostMalloc((void **)&base_, bytes, hipHostMallocDefault) != hipSuccess) {
            base_ = nullptr;
            return false;
        }
        cap_ = bytes;
        free_[0] = bytes;
        return true;
    }
    unsigned char *at(size_t off) const { return base_ + off; }
    size_t alloc(size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        std::lock_guard<std::mutex> g(mu_);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t off = it->first, left = it->second - bytes;
            free_.erase(it);
            if (left) free_[off + bytes] = left;
            return off;
        }
        return kNone;
    }
    void release(size_t off, size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        std::lock_guard<std::mutex> g(mu_);
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && off + bytes == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

  private:
    std::mutex mu_;
    unsigned char *base_ = nullptr;
    size_t cap_ = 0;
    std::map<size_t, size_t> free_;  // offset -> bytes
};

struct Request {
    int64_t id = 0;
    int m = 0, nnz = 0;
    // the matrix: given as 1-based COO (kind 0), or built on the worker's
    // stream from a point cloud (kind 1, VR 2-skeleton) or a grid (kind 2,
    // cubical lower-star)
    int kind = 0;
    std::vector<int, NoInit<int>> rows, cols;
    std::vector<double> points;
    std::vector<float> field;
    int n = 0, pdim = 0, s0 = 0, s1 = 0, s2 = 0;
    // where the input is read at staging: the copies above, or the caller's
    // arrays (borrowed); where the outputs go: low/ess below, or the
    // caller's buffers
    const int *in_rows = nullptr, *in_cols = nullptr;
    size_t arena_off = Arena::kNone, arena_bytes = 0;  // in_rows/in_cols in the arena
    const double *in_points = nullptr;
    const float *in_field = nullptr;
    int *out_low = nullptr, *out_ess = nullptr;
    std::vector<int> lens;  // distinct input column lengths
    int cls = 0;            // 0: input columns of at most 4 entries, 1: longer
    size_t layout = 0;      // pool entries of its initial layout (an upper bound)
    std::vector<int> low, ess;
    double t_submit = 0, t_start = 0, t_done = 0;
    int status = OPENPH_OK;
    int batch_size = 0, worker = -1, num_iters = 0;
    int join_iter = 0;      // the batch's iteration count when it joined
    bool done = false, claimed = false;
};

// A retired member whose outputs are being copied to the worker's pinned
// buffers at offset a.
struct Outgoing {
    Request *r;
    int a;
};

struct Worker {
    int index = 0;
    Context ctx;
    LiveBatch lb;
    std::vector<Request *> members;  // member k of the running batch
    std::vector<char> retired;
    int admitted = 0;  // requests that have joined the running batch
    // seconds spent per phase (reported with OPENPH_DEBUG=1): waiting for
    // work, staging and appending, iterations, retiring, requests run alone
    double t_wait = 0, t_stage = 0, t_iter = 0, t_retire = 0, t_alone = 0;
    double t_host = 0, t_up = 0, t_append = 0;  // parts of t_stage
    long long n_rounds = 0;
    long long n_chunks = 0;
    size_t m_cap = 0, nnz_cap = 0, pool_cap = 0;
    hipStream_t stream = nullptr;
    int *h_rows = nullptr, *h_cols = nullptr, *h_low = nullptr, *h_ess = nullptr;
    size_t c_rows = 0, c_cols = 0, c_low = 0, c_ess = 0;
    int *d_low = nullptr, *d_ess = nullptr;
    size_t c_dlow = 0, c_dess = 0;
    // inputs of the requests built on the device: pinned staging, device copy
    unsigned char *h_in = nullptr, *d_in = nullptr;
    size_t c_hin = 0, c_din = 0;
    // ev_in: the last upload from the pinned input buffers is done;
    // ev_out: the outputs of the members in `outbox` are on the host
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    std::vector<Outgoing> outbox;
    std::thread thread;

    ~Worker() {
        int *pinned[] = {h_rows, h_cols, h_low, h_ess};
        for (int *p : pinned)
            if (p) (void)hipHostFree(p);
        if (h_in) (void)hipHostFree(h_in);
        if (d_low) (void)hipFree(d_low);
        if (d_ess) (void)hipFree(d_ess);
        if (d_in) (void)hipFree(d_in);
        if (ev_in) (void)hipEventDestroy(ev_in);
        if (ev_out) (void)hipEventDestroy(ev_out);
    }
};

}  // namespace oph

struct openph_server {
    openph_server_options opt{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<oph::Request *> queue;
    std::map<int64_t, std::unique_ptr<oph::Request>> requests;
    std::vector<std::unique_ptr<oph::Worker>> workers;
    // lent COO requests waiting for validation by the intake threads, which
    // then queue them (or finish them with their error)
    std::deque<oph::Request *> intake;
    std::condition_variable cv_intake;
    std::vector<std::thread> intakers;
    bool stop_intake = false;
    oph::Arena arena;
    int64_t next_id = 1;
    bool stop = false;
    int ready = 0;  // workers whose workspace is reserved
    // statistics
    int64_t batches = 0, completed = 0;
    long long columns_done = 0;
    // seconds spent inside the submit calls (reported with OPENPH_DEBUG=1)
    double t_submit = 0;
    long long n_submit = 0;
};

namespace oph {

constexpr int kChunkIters = 2;     // iterations between two admissions / retirements
constexpr int kIntakeThreads = 4;  // threads validating lent COO requests

template <typename T>
static hipError_t grow_pinned(T **p, size_t &cap, size_t need) {
    if (need <= cap && *p) return hipSuccess;
    if (*p) {
        (void)hipHostFree(*p);
        *p = nullptr;
    }
    size_t n = std::max<size_t>(need, 1);
    hipError_t e = hipHostMalloc((void **)p, n * sizeof(T), hipHostMallocDefault);
    if (e != hipSuccess) {
        cap = 0;
        *p = nullptr;
        return e;
    }
    cap = n;
    return hipSuccess;
}

static size_t in_bytes(const Request *r) {
    if (r->kind == 1) return ((size_t)r->n * r->pdim * sizeof(double) + 255) & ~(size_t)255;
    if (r->kind == 2) return ((size_t)r->s0 * r->s1 * r->s2 * sizeof(float) + 255) & ~(size_t)255;
    return 0;
}

// memcpy, split over up to 8 threads when it is large (host copies of a
// request's COO and results run at several GB/s per thread).
static void pcopy(void *dst, const void *src, size_t bytes) {
    constexpr size_t kPer = (size_t)2 << 20;
    const int nt = (int)std::min<size_t>(8, std::max<size_t>(1, bytes / kPer));
    if (nt == 1) {
        if (bytes) std::memcpy(dst, src, bytes);
        return;
    }
    const size_t part = (bytes / nt + 63) & ~(size_t)63;
    std::vector<std::thread> th;
    for (int k = 1; k < nt; k++) {
        const size_t a = std::min(bytes, part * k), b = std::min(bytes, part * (k + 1));
        if (b > a)
            th.emplace_back([=] { std::memcpy((char *)dst + a, (const char *)src + a, b - a); });
    }
    std::memcpy(dst, src, std::min(bytes, part));
    for (std::thread &t : th) t.join();
}

// A finished request's outputs (1-based low, ess) from pinned host memory
// into the caller's buffers or the request's own.
static void deliver(Request *r, const int *low1, const int *ess) {
    const size_t M = (size_t)r->m;
    if (r->out_low) {
        pcopy(r->out_low, low1, sizeof(int) * M);
        pcopy(r->out_ess, ess, sizeof(int) * M);
    } else {
        r->low.assign(low1, low1 + M);
        r->ess.assign(ess, ess + M);
    }
}

// A request whose COO waits in the arena, large enough to go up on its own.
static bool direct(const Request *r) {
    return r->kind == 0 && r->arena_off != Arena::kNone && r->nnz >= (1 << 14);
}

// Host half of staging: the COO of `reqs` (copied at submit or lent by the
// caller) and the inputs of the requests built on the device into the
// worker's pinned buffers, once the last upload from them is done; a COO
// that waits in the arena is left there (it goes up from there).
static int stage_host(Worker &w, const std::vector<Request *> &reqs) {
    struct Clock {
        double &acc, t0 = now_s();
        ~Clock() { acc += now_s() - t0; }
    } clock{w.t_host};
    size_t N = 0, inb = 0;
    for (Request *r : reqs) {
        N += r->nnz;
        inb += in_bytes(r);
    }
    OPH_CHECK(hipEventSynchronize(w.ev_in));
    OPH_CHECK(grow_pinned(&w.h_rows, w.c_rows, N + 1));
    OPH_CHECK(grow_pinned(&w.h_cols, w.c_cols, N + 1));
    if (inb) OPH_CHECK(grow_pinned(&w.h_in, w.c_hin, inb));
    size_t o = 0, b = 0;
    for (Request *r : reqs) {
        if (r->kind == 0 && r->nnz && !direct(r)) {
            pcopy(w.h_rows + o, r->in_rows, sizeof(int) * r->nnz);
            pcopy(w.h_cols + o, r->in_cols, sizeof(int) * r->nnz);
        }
        if (r->kind == 1)
            std::memcpy(w.h_in + b, r->in_points, (size_t)r->n * r->pdim * sizeof(double));
        if (r->kind == 2)
            std::memcpy(w.h_in + b, r->in_field, (size_t)r->s0 * r->s1 * r->s2 * sizeof(float));
        o += r->nnz;
        b += in_bytes(r);
    }
    return OPENPH_OK;
}

// Device half: the staged inputs go up and the local 1-based COO of `reqs`,
// concatenated, is written into the device buffers rows1/cols1 (the given
// COO as one copy, the built matrices over their own ranges after it), all
// queued on the worker's stream.
static int stage_device(Worker &w, const std::vector<Request *> &reqs, int *rows1, int *cols1) {
    Context &c = w.ctx;
    hipStream_t s = w.stream;
    size_t N = 0, inb = 0;
    for (Request *r : reqs) {
        N += r->nnz;
        inb += in_bytes(r);
    }
    // runs of staged COO go up from the worker's buffers, each arena COO
    // from the arena, the built matrices' ranges are left to their builders
    size_t run = 0, o = 0;
    auto flush = [&](size_t end) -> hipError_t {
        hipError_t e = hipSuccess;
        if (end > run) {
            e = hipMemcpyAsync(rows1 + run, w.h_rows + run, sizeof(int) * (end - run),
                               hipMemcpyHostToDevice, s);
            if (e == hipSuccess)
                e = hipMemcpyAsync(cols1 + run, w.h_cols + run, sizeof(int) * (end - run),
                                   hipMemcpyHostToDevice, s);
        }
        return e;
    };
    for (Request *r : reqs) {
        if (r->kind != 0 || direct(r)) {
            OPH_CHECK(flush(o));
            if (direct(r)) {
                OPH_CHECK(hipMemcpyAsync(rows1 + o, r->in_rows, sizeof(int) * r->nnz,
                                         hipMemcpyHostToDevice, s));
                OPH_CHECK(hipMemcpyAsync(cols1 + o, r->in_cols, sizeof(int) * r->nnz,
                                         hipMemcpyHostToDevice, s));
            }
            run = o + r->nnz;
        }
        o += r->nnz;
    }
    OPH_CHECK(flush(o));
    (void)N;
    if (inb) {
        OPH_CHECK(grow(&w.d_in, w.c_din, inb));
        OPH_CHECK(hipMemcpyAsync(w.d_in, w.h_in, inb, hipMemcpyHostToDevice, s));
    }
    OPH_CHECK(hipEventRecord(w.ev_in, s));
    size_t ro = 0, b = 0;
    for (Request *r : reqs) {
        int rc = OPENPH_OK;
        if (r->kind == 1)
            rc = build_vr(c, reinterpret_cast<const double *>(w.d_in + b), r->n, r->pdim,
                          rows1 + ro, cols1 + ro, nullptr, nullptr, s);
        else if (r->kind == 2)
            rc = build_cubical(c, reinterpret_cast<const float *>(w.d_in + b), r->s0, r->s1,
                               r->s2, rows1 + ro, cols1 + ro, nullptr, nullptr, s);
        if (rc) return rc;
        b += in_bytes(r);
        ro += r->nnz;
    }
    return OPENPH_OK;
}

// Releases finished requests to their waiters.
static void finish(openph_server *srv, const std::vector<Request *> &done) {
    if (done.empty()) return;
    {
        std::lock_guard<std::mutex> g(srv->mu);
        for (Request *r : done) {
            r->done = true;
            r->rows.clear();
            r->rows.shrink_to_fit();
            r->cols.clear();
            r->cols.shrink_to_fit();
            r->points.clear();
            r->points.shrink_to_fit();
            r->field.clear();
            r->field.shrink_to_fit();
            r->in_rows = r->in_cols = nullptr;
            r->in_points = nullptr;
            r->in_field = nullptr;
            if (r->arena_off != Arena::kNone) srv->arena.release(r->arena_off, r->arena_bytes);
            r->arena_off = Arena::kNone;
            srv->columns_done += r->m;
        }
        srv->completed += (int64_t)done.size();
    }
    srv->cv_done.notify_all();
}

static void fail_all(openph_server *srv, std::vector<Request *> &reqs, int rc) {
    const double t = now_s();
    for (Request *r : reqs) {
        r->status = rc;
        r->t_done = t;
    }
    finish(srv, reqs);
    reqs.clear();
}

// Members retired by the last chunk: their outputs are on their way to the
// pinned host buffers (w.ev_out marks the copies' end); they are delivered
// and released.
static void flush_out(openph_server *srv, Worker &w) {
    if (w.outbox.empty()) return;
    const double t0 = now_s();
    const int rc = hipEventSynchronize(w.ev_out) == hipSuccess ? OPENPH_OK : OPENPH_EDEVICE;
    std::vector<Request *> done;
    for (const Outgoing &o : w.outbox) {
        if (!rc) deliver(o.r, w.h_low + o.a, w.h_ess + o.a);
        o.r->status = rc;
        o.r->t_done = now_s();
        done.push_back(o.r);
    }
    w.outbox.clear();
    finish(srv, done);
    w.t_retire += now_s() - t0;
}

// One request reduced alone through the batch path (it may grow the
// workspace): requests too large for the running batches, and the members
// of a batch that failed.
static void run_alone(openph_server *srv, Worker &w, Request *r) {
    Context &c = w.ctx;
    hipStream_t s = w.stream;
    flush_out(srv, w);  // before the host buffers are reused
    w.lb.running = false;
    r->t_start = r->t_start > 0 ? r->t_start : now_s();
    int rc = OPENPH_OK;
    const size_t M = (size_t)r->m, N = (size_t)r->nnz;
    if ((rc = (int)grow(&c.h_rows, c.c_hrows, N + 1)) != hipSuccess ||
        (rc = (int)grow(&c.h_cols, c.c_hcols, N + 1)) != hipSuccess ||
        (rc = (int)grow(&w.d_low, w.c_dlow, M + 1)) != hipSuccess ||
        (rc = (int)grow(&w.d_ess, w.c_dess, M + 1)) != hipSuccess ||
        (rc = (int)grow_pinned(&w.h_low, w.c_low, M + 1)) != hipSuccess ||
        (rc = (int)grow_pinned(&w.h_ess, w.c_ess, M + 1)) != hipSuccess) {
        rc = OPENPH_ENOMEM;
    }
    if (!rc) rc = stage_host(w, {r});
    if (!rc) rc = stage_device(w, {r}, c.h_rows, c.h_cols);
    openph_stats st;
    std::memset(&st, 0, sizeof(st));
    if (!rc) {
        RunCfg cfg = cfg_from(&srv->opt.reduce, r->nnz);
        rc = run_batch(c, 1, &r->m, &r->nnz, c.h_rows, c.h_cols, w.d_low, w.d_ess, s, cfg, &st);
    }
    if (!rc && M) {
        if (hipMemcpyAsync(w.h_low, w.d_low, sizeof(int) * M, hipMemcpyDeviceToHost, s) ||
            hipMemcpyAsync(w.h_ess, w.d_ess, sizeof(int) * M, hipMemcpyDeviceToHost, s))
            rc = OPENPH_EDEVICE;
    }
    if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = OPENPH_EDEVICE;
    if (!rc) deliver(r, w.h_low, w.h_ess);
    r->status = rc;
    r->num_iters = st.num_iters;
    r->batch_size = 1;
    r->worker = w.index;
    r->t_done = now_s();
    finish(srv, {r});
    {
        std::lock_guard<std::mutex> g(srv->mu);
        srv->batches++;
    }
}

// The running batch failed: each of its members that is still running goes
// again alone.
static void fail_batch(openph_server *srv, Worker &w) {
    (void)hipStreamSynchronize(w.stream);
    std::vector<Request *> left;
    for (size_t k = 0; k < w.members.size(); k++)
        if (!w.retired[k]) left.push_back(w.members[k]);
    w.members.clear();
    w.retired.clear();
    w.lb.running = false;
    for (Request *r : left) run_alone(srv, w, r);
}

// Can r join a batch of class cls holding `cols` columns, `pool` pool
// entries and the pseudo-dimension slots `slots` (lengths), with `nnz` COO
// entries already staged in this round?
static bool fits(const Worker &w, const Request *r, int cls, size_t cols, size_t nnz, size_t pool,
                 const std::vector<int> &slots, int members, int max_batch) {
    if (r->cls != cls || members + 1 > max_batch) return false;
    if (cols + r->m > w.m_cap || nnz + r->nnz > w.nnz_cap) return false;
    if (pool + r->layout > w.pool_cap / 4 * 3) return false;
    int n = (int)slots.size();
    for (int L : r->lens)
        if (std::find(slots.begin(), slots.end(), L) == slots.end()) n++;
    return n <= OPENPH_MAX_DIMS;
}

// Takes queued requests for this worker into `round`. With no running batch
// the oldest request starts one (or, too large for an empty batch, is
// returned in `alone`); requests of the batch's class follow while they fit
// (`pool_now`: the running batch's pool use) and the round holds less than
// a fair share of the queued columns (their sum over the number of
// workers), so that a burst is spread over the workers instead of filling
// the first one to wake. `block` waits for work; returns false when the
// server stops and the queue is empty.
static bool admit(openph_server *srv, Worker &w, bool block, size_t pool_now,
                  std::vector<Request *> &round, Request *&alone) {
    LiveBatch &lb = w.lb;
    const int max_batch = srv->opt.max_batch;
    std::unique_lock<std::mutex> g(srv->mu);
    if (block) {
        srv->cv_work.wait(g, [&] { return srv->stop || !srv->queue.empty(); });
        if (srv->queue.empty()) return false;
    }
    std::vector<int> slots;
    for (size_t L = 0; L < lb.slot_of_len.size(); L++)
        if (lb.running && lb.slot_of_len[L] >= 0) slots.push_back((int)L);
    int cls = lb.running ? lb.cls : -1;
    size_t cols = lb.running ? (size_t)lb.mstart.back() : 0, nnz = 0;
    size_t pool = lb.running ? pool_now : 0;
    int members = lb.running ? (int)w.members.size() : 0;
    size_t queued = 0, taken = 0;
    for (const Request *r : srv->queue) queued += (size_t)r->m;
    const size_t share = queued / srv->workers.size() + 1;
    for (auto it = srv->queue.begin(); it != srv->queue.end() && taken < share;) {
        Request *r = *it;
        if (cls < 0) {
            cls = r->cls;
            if (!fits(w, r, cls, 0, 0, 0, {}, 0, max_batch)) {
                alone = r;
                srv->queue.erase(it);
                break;
            }
        }
        if (fits(w, r, cls, cols, nnz, pool, slots, members, max_batch)) {
            round.push_back(r);
            taken += (size_t)r->m;
            cols += r->m;
            nnz += r->nnz;
            pool += r->layout;
            members++;
            for (int L : r->lens)
                if (std::find(slots.begin(), slots.end(), L) == slots.end()) slots.push_back(L);
            it = srv->queue.erase(it);
        } else {
            ++it;
        }
    }
    if (!lb.running && !round.empty()) srv->batches++;
    return true;
}

// The staged round joins the running batch: uploads, builds, set-up and
// Phase 0 of its columns are queued; it takes part from the next chunk on.
static int join(Worker &w, std::vector<Request *> &round) {
    Context &c = w.ctx;
    const double t0 = now_s();
    int rc = stage_device(w, round, c.h_rows, c.h_cols);
    const double t1 = now_s();
    w.t_up += t1 - t0;
    w.n_rounds++;
    if (!rc) {
        std::vector<int> ms, nnzs;
        std::vector<std::vector<int>> lens;
        int N = 0;
        for (Request *r : round) {
            ms.push_back(r->m);
            nnzs.push_back(r->nnz);
            lens.push_back(r->lens);
            N += r->nnz;
        }
        rc = live_append(c, w.lb, (int)round.size(), ms.data(), nnzs.data(), lens.data(),
                         c.h_rows, c.h_cols, N, w.stream);
    }
    w.t_append += now_s() - t1;
    const double t = now_s();
    w.admitted += (int)round.size();
    for (Request *r : round) {
        r->t_start = t;
        r->join_iter = c.h_ctrl->iter;
        r->worker = w.index;
        w.members.push_back(r);
        w.retired.push_back(0);
    }
    round.clear();
    return rc;
}

static void worker_main(openph_server *srv, Worker *w) {
    (void)hipSetDevice(srv->opt.device);
    w->ctx.device = srv->opt.device;
    if (hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&w->ev_in, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&w->ev_out, hipEventDisableTiming) != hipSuccess)
        w->stream = nullptr;
    const size_t budget = (size_t)srv->opt.worker_bytes;
    const int slot0 = srv->opt.reduce.slot_elems > 0 ? srv->opt.reduce.slot_elems : 32;
    const int max_batch = srv->opt.max_batch;
    {
        // Size the workspace for a full budget up front, taking the typical
        // nnz = 3 m of the boundary matrices served here: per column 84 B of
        // state, 48 B of local and global COO, 8 B of outputs and two pools
        // of 3 + slot0 + 24 entries (layout and headroom).
        const size_t per_col = 84 + 48 + 8 + 2 * 4 * (3 + (size_t)slot0 + 24);
        w->m_cap = std::min<size_t>(budget / per_col, 0x7ffffff0ull);
        w->nnz_cap = std::min<size_t>(3 * w->m_cap, 0x7ffffff0ull);
        w->pool_cap = std::min<size_t>(w->m_cap * (3 + (size_t)slot0 + 24), 0xFFFFFFFFull);
        Context &c = w->ctx;
        (void)reserve(c, w->m_cap, w->nnz_cap, w->pool_cap);
        (void)grow(&c.d.pool2, c.cap_pool2, w->pool_cap);
        (void)reserve_builders(c, std::min<size_t>(w->m_cap, (size_t)8 << 20), 256);
        (void)grow_pinned(&w->h_rows, w->c_rows, w->nnz_cap + 1);
        (void)grow_pinned(&w->h_cols, w->c_cols, w->nnz_cap + 1);
        (void)grow_pinned(&w->h_low, w->c_low, w->m_cap + 1);
        (void)grow_pinned(&w->h_ess, w->c_ess, w->m_cap + 1);
        (void)grow(&w->d_low, w->c_dlow, w->m_cap + 1);
        (void)grow(&w->d_ess, w->c_dess, w->m_cap + 1);
        // inputs of built requests (a point cloud or a grid of side 160):
        // growing them later would wait for the whole device
        (void)grow_pinned(&w->h_in, w->c_hin, (size_t)16 << 20);
        (void)grow(&w->d_in, w->c_din, (size_t)16 << 20);
        std::lock_guard<std::mutex> g(srv->mu);
        srv->ready++;
    }
    srv->cv_done.notify_all();
    Context &c = w->ctx;
    LiveBatch &lb = w->lb;
    const RunCfg cfg = [&] {
        RunCfg r = cfg_from(&srv->opt.reduce, 0);
        r.pool_max = std::min(r.pool_max, w->pool_cap);
        return r;
    }();
    const int max_iters = cfg.max_iters;
    hipStream_t s = w->stream;
    // the next requests to join: taken from the queue, their inputs staged
    // in the pinned buffers (`ready`) while the last chunk ran
    std::vector<Request *> next;
    bool ready = false;
    size_t pool_now = 0;  // the running batch's pool use after the last chunk
    while (true) {
        if (!lb.running) {
            flush_out(srv, *w);
            if (next.empty()) {
                Request *alone = nullptr;
                const double tw0 = now_s();
                const bool go = admit(srv, *w, true, 0, next, alone);
                w->t_wait += now_s() - tw0;
                if (!go) break;
                if (!w->stream) {
                    if (alone) next.push_back(alone);
                    fail_all(srv, next, OPENPH_EDEVICE);
                    continue;
                }
                if (alone) {
                    const double ta = now_s();
                    run_alone(srv, *w, alone);
                    w->t_alone += now_s() - ta;
                    continue;
                }
                if (next.empty()) continue;
                ready = false;
            }
            const double t0 = now_s();
            int rc = ready ? OPENPH_OK : stage_host(*w, next);
            if (rc) {
                fail_all(srv, next, rc);
                continue;
            }
            rc = live_start(c, lb, cfg, next[0]->cls, max_batch, w->pool_cap, s);
            w->admitted = 0;
            pool_now = 0;
            for (Request *r : next) pool_now += r->layout;
            if (!rc) {
                rc = join(*w, next);
            } else {
                for (Request *r : next) {
                    w->members.push_back(r);
                    w->retired.push_back(0);
                }
                next.clear();
            }
            ready = false;
            w->t_stage += now_s() - t0;
            if (rc) {
                fail_batch(srv, *w);
                continue;
            }
        }
        // One chunk of iterations on the device; meanwhile the host delivers
        // the members retired by the last one and stages the next arrivals.
        const double ti = now_s();
        int rc = live_chunk_launch(c, lb, kChunkIters, s);
        if (!rc) {
            flush_out(srv, *w);
            if (next.empty()) {
                Request *none = nullptr;
                (void)admit(srv, *w, false, pool_now, next, none);
                if (!next.empty()) {
                    const double t0 = now_s();
                    const int rs = stage_host(*w, next);
                    w->t_stage += now_s() - t0;
                    if (rs) fail_all(srv, next, rs);
                    ready = !rs;
                }
            }
            rc = live_chunk_finish(c, lb, cfg, s);
        }
        const double tr = now_s();
        w->t_iter += tr - ti;
        w->n_chunks++;
        if (rc) {
            fail_batch(srv, *w);
            ready = false;  // the members run alone through the same buffers
            continue;
        }
        const Ctrl &h = *c.h_ctrl;
        pool_now = (size_t)h.pool_top;
        const bool all = h.halt == kHaltConverged;
        // retirements: members whose lows no longer repeat (all of them once
        // the batch has converged); members past the iteration cap fail alone
        std::vector<Request *> over;
        for (size_t k = 0; k < w->members.size(); k++) {
            if (w->retired[k]) continue;
            Request *r = w->members[k];
            const bool conv = all || (h.iter > r->join_iter && lb.h_mdup[k] == 0);
            if (!conv && h.iter - r->join_iter <= max_iters) continue;
            const int a = lb.mstart[k];
            if ((rc = live_retire(c, lb, (int)k, w->d_low + a, w->d_ess + a, s))) break;
            w->retired[k] = 1;
            r->num_iters = h.iter - r->join_iter;
            r->batch_size = w->admitted;
            if (!conv) {
                r->status = OPENPH_EITER;
                r->t_done = now_s();
                over.push_back(r);
                continue;
            }
            const int n = lb.mstart[k + 1] - a;
            if (n && (hipMemcpyAsync(w->h_low + a, w->d_low + a, sizeof(int) * n,
                                     hipMemcpyDeviceToHost, s) ||
                      hipMemcpyAsync(w->h_ess + a, w->d_ess + a, sizeof(int) * n,
                                     hipMemcpyDeviceToHost, s))) {
                rc = OPENPH_EDEVICE;
                break;
            }
            w->outbox.push_back({r, a});
        }
        finish(srv, over);
        if (!rc && hipEventRecord(w->ev_out, s) != hipSuccess) rc = OPENPH_EDEVICE;
        if (rc) {
            // the retired members' outputs are not trusted: they run again
            for (const Outgoing &o : w->outbox)
                for (size_t k = 0; k < w->members.size(); k++)
                    if (w->members[k] == o.r) w->retired[k] = 0;
            w->outbox.clear();
            fail_batch(srv, *w);
            ready = false;
            continue;
        }
        w->t_retire += now_s() - tr;
        bool any = false;
        for (char x : w->retired) any |= !x;
        if (!any || all) {
            // a converged batch has retired every member
            lb.running = false;
            w->members.clear();
            w->retired.clear();
        } else {
            live_trim(c, lb, w->retired, s);
            w->members.resize(lb.mstart.size() - 1);
            w->retired.resize(lb.mstart.size() - 1);
        }
        if (lb.running && !next.empty() && ready) {
            const double t0 = now_s();
            for (Request *r : next) pool_now += r->layout;
            rc = join(*w, next);
            ready = false;
            w->t_stage += now_s() - t0;
            if (rc) fail_batch(srv, *w);
        }
    }
    flush_out(srv, *w);
    if (std::getenv("OPENPH_DEBUG"))
        std::fprintf(stderr,
                     "openph server worker %d: wait %.3f s, stage %.3f s (host copies %.3f, "
                     "uploads %.3f, appends %.3f; %lld rounds), iterations %.3f s (%lld chunks), "
                     "retire %.3f s, alone %.3f s\n",
                     w->index, w->t_wait, w->t_stage, w->t_host, w->t_up, w->t_append,
                     w->n_rounds, w->t_iter, w->n_chunks, w->t_retire, w->t_alone);
    if (w->stream) (void)hipStreamDestroy(w->stream);
}

static void intake_main(openph_server *srv);

}  // namespace oph

using namespace oph;

extern "C" {

int openph_server_create(const openph_server_options *opt, openph_server **out) {
    if (!out) return OPENPH_EINVAL;
    *out = nullptr;
    openph_server_options o;
    std::memset(&o, 0, sizeof(o));
    if (opt) o = *opt;
    if (o.workers <= 0) o.workers = 4;
    if (o.workers > 16) o.workers = 16;
    if (o.worker_bytes <= 0) o.worker_bytes = (int64_t)8 << 30;
    if (o.max_batch <= 0) o.max_batch = 1024;
    if (o.max_queue <= 0) o.max_queue = 1 << 16;
    o.reduce.device = o.device;
    int rc = set_device(o.device);
    if (rc) return rc;
    if (o.stage_bytes == 0) o.stage_bytes = (int64_t)2 << 30;
    openph_server *srv = new openph_server();
    srv->opt = o;
    if (o.stage_bytes > 0) (void)srv->arena.init((size_t)o.stage_bytes);
    for (int i = 0; i < o.workers; i++) {
        auto w = std::make_unique<Worker>();
        w->index = i;
        srv->workers.push_back(std::move(w));
    }
    for (auto &w : srv->workers) w->thread = std::thread(worker_main, srv, w.get());
    for (int i = 0; i < kIntakeThreads; i++) srv->intakers.emplace_back(intake_main, srv);
    {
        // return once every worker has its workspace
        std::unique_lock<std::mutex> g(srv->mu);
        srv->cv_done.wait(g, [&] { return srv->ready == (int)srv->workers.size(); });
    }
    *out = srv;
    return OPENPH_OK;
}

}  // extern "C"

namespace oph {

// Queues a request; its size, column lengths and class decide which batch
// it may join.
static void classify(const openph_server *srv, Request *r) {
    const int slot0 = srv->opt.reduce.slot_elems > 0 ? srv->opt.reduce.slot_elems : 32;
    r->layout = (size_t)r->nnz + (size_t)r->m * (size_t)slot0;
    int maxlen = 0;
    for (int L : r->lens) maxlen = std::max(maxlen, L);
    r->cls = maxlen <= 4 ? 0 : 1;
}

// Queues a request: validated ones for the workers, lent COO still to be
// validated for the intake threads.
static int enqueue(openph_server *srv, std::unique_ptr<Request> r, int64_t *ticket,
                   bool validated = true) {
    if (validated) classify(srv, r.get());
    {
        std::lock_guard<std::mutex> g(srv->mu);
        if ((int)(srv->queue.size() + srv->intake.size()) >= srv->opt.max_queue) {
            if (r->arena_off != Arena::kNone) srv->arena.release(r->arena_off, r->arena_bytes);
            return OPENPH_EBUSY;
        }
        r->id = srv->next_id++;
        r->t_submit = now_s();
        *ticket = r->id;
        (validated ? srv->queue : srv->intake).push_back(r.get());
        srv->requests[r->id] = std::move(r);
    }
    (validated ? srv->cv_work : srv->cv_intake).notify_one();
    return OPENPH_OK;
}

// Validation (the device set-up's rules), distinct column lengths and the
// copy of one stretch [e0, e1) of a request's COO whose ends are column
// boundaries. Returns 0 or an OPENPH_E* code.
static int intake(const int *rows1, const int *cols1, int m, int e0, int e1, int *rows_out,
                  int *cols_out, std::vector<int> &lens, int &ncols) {
    int e = e0;
    ncols = 0;
    while (e < e1) {
        const int c = cols1[e];
        if (c > m || c < 1) return OPENPH_EINVAL;
        if (e > e0 && c < cols1[e - 1]) return OPENPH_EUNSORTED;
        int f = e;
        int prev = 0;
        while (f < e1 && cols1[f] == c) {
            const int rr = rows1[f];
            if (rr < 1 || rr >= c) return OPENPH_EINVAL;
            if (f > e && rr <= prev) return OPENPH_EUNSORTED;
            prev = rr;
            f++;
        }
        const int L = f - e;
        if (std::find(lens.begin(), lens.end(), L) == lens.end()) {
            if ((int)lens.size() >= OPENPH_MAX_DIMS) return OPENPH_EDIMS;
            lens.push_back(L);
        }
        ncols++;
        e = f;
    }
    if (rows_out) {
        std::memcpy(rows_out + e0, rows1 + e0, sizeof(int) * (size_t)(e1 - e0));
        std::memcpy(cols_out + e0, cols1 + e0, sizeof(int) * (size_t)(e1 - e0));
    }
    return OPENPH_OK;
}

// Adds the duration of one submit call to the server's total.
struct SubmitClock {
    openph_server *srv;
    double t0;
    explicit SubmitClock(openph_server *s) : srv(s), t0(now_s()) {}
    ~SubmitClock() {
        if (!srv) return;
        const double dt = now_s() - t0;
        std::lock_guard<std::mutex> g(srv->mu);
        srv->t_submit += dt;
        srv->n_submit++;
    }
};

// Validates a COO (and copies it into rows_out/cols_out unless they are
// NULL) with up to max_threads threads; `lens_out` receives its distinct
// column lengths. A large request is taken in stretches, each ending at a
// column boundary.
static int check_coo(const int *rows1, const int *cols1, int m, int nnz, int *rows_out,
                     int *cols_out, int max_threads, std::vector<int> &lens_out) {
    const int nt = std::min(max_threads, std::max(1, nnz >> 18));
    std::vector<int> cut((size_t)nt + 1, nnz);
    cut[0] = 0;
    for (int k = 1; k < nt; k++) {
        int e = (int)((long long)nnz * k / nt);
        if (e < cut[k - 1]) e = cut[k - 1];
        while (e > 0 && e < nnz && cols1[e] == cols1[e - 1]) e++;
        cut[k] = e;
    }
    std::vector<std::vector<int>> lens((size_t)nt);
    std::vector<int> rc((size_t)nt, 0), ncols((size_t)nt, 0);
    {
        std::vector<std::thread> th;
        for (int k = 1; k < nt; k++)
            th.emplace_back([&, k] {
                rc[k] = intake(rows1, cols1, m, cut[k], cut[k + 1], rows_out, cols_out, lens[k],
                               ncols[k]);
            });
        rc[0] = intake(rows1, cols1, m, cut[0], cut[1], rows_out, cols_out, lens[0], ncols[0]);
        for (std::thread &t : th) t.join();
    }
    int cols_seen = 0;
    for (int k = 0; k < nt; k++) {
        if (rc[k]) return rc[k];
        // across a stretch boundary columns must keep ascending
        if (k > 0 && cut[k] > 0 && cut[k] < nnz && cols1[cut[k]] < cols1[cut[k] - 1])
            return OPENPH_EUNSORTED;
        cols_seen += ncols[k];
        for (int L : lens[k])
            if (std::find(lens_out.begin(), lens_out.end(), L) == lens_out.end())
                lens_out.push_back(L);
    }
    if (cols_seen < m && std::find(lens_out.begin(), lens_out.end(), 0) == lens_out.end())
        lens_out.push_back(0);
    if ((int)lens_out.size() > OPENPH_MAX_DIMS) return OPENPH_EDIMS;
    return OPENPH_OK;
}

// Gives request r a block of the arena for its rows and cols (in_rows and
// in_cols then point there); false when the arena is absent or full.
static bool place(openph_server *srv, Request *r) {
    if (r->nnz <= 0) return false;
    const size_t bytes = 2 * sizeof(int) * ((size_t)r->nnz + 64);
    const size_t off = srv->arena.alloc(bytes);
    if (off == Arena::kNone) return false;
    r->arena_off = off;
    r->arena_bytes = bytes;
    int *p = reinterpret_cast<int *>(srv->arena.at(off));
    r->in_rows = p;
    r->in_cols = p + (((size_t)r->nnz + 63) & ~(size_t)63);
    return true;
}

// Queues a COO request. A copied one is validated here, so that a malformed
// request is refused alone instead of failing the batch it would join (same
// rules as the device set-up), with its distinct column lengths
// (pseudo-dimension + 1): a batch holds at most OPENPH_MAX_DIMS of them over
// all its members. A lent one is validated the same way by an intake thread
// after submit has returned, and a malformed one is finished with its error.
static int submit_coo(openph_server *srv, const int *rows1, const int *cols1, int m, int nnz,
                      bool borrow, int *low1, int *ess, int64_t *ticket) {
    if (!srv || !ticket || m < 0 || nnz < 0 || (nnz > 0 && (!rows1 || !cols1)))
        return OPENPH_EINVAL;
    auto r = std::make_unique<Request>();
    r->m = m;
    r->nnz = nnz;
    r->out_low = low1;
    r->out_ess = ess;
    if (borrow) {
        r->in_rows = rows1;
        r->in_cols = cols1;
        return enqueue(srv, std::move(r), ticket, false);
    }
    int *rows_out = nullptr, *cols_out = nullptr;
    if (place(srv, r.get())) {
        rows_out = const_cast<int *>(r->in_rows);
        cols_out = const_cast<int *>(r->in_cols);
    } else {
        r->rows.resize((size_t)nnz);
        r->cols.resize((size_t)nnz);
        r->in_rows = rows_out = r->rows.data();
        r->in_cols = cols_out = r->cols.data();
    }
    const int rc = check_coo(rows1, cols1, m, nnz, rows_out, cols_out, 8, r->lens);
    if (rc) {
        if (r->arena_off != Arena::kNone) srv->arena.release(r->arena_off, r->arena_bytes);
        return rc;
    }
    return enqueue(srv, std::move(r), ticket);
}

// An intake thread: takes lent COO requests in arrival order (several
// threads validate several requests at once) and queues them for the
// workers.
static void intake_main(openph_server *srv) {
    while (true) {
        Request *r = nullptr;
        {
            std::unique_lock<std::mutex> g(srv->mu);
            srv->cv_intake.wait(g, [&] { return srv->stop_intake || !srv->intake.empty(); });
            if (srv->intake.empty()) return;
            r = srv->intake.front();
            srv->intake.pop_front();
        }
        // validated, and copied into the arena on the way when it has room
        const int *rows1 = r->in_rows, *cols1 = r->in_cols;
        int *rows_out = nullptr, *cols_out = nullptr;
        if (place(srv, r)) {
            rows_out = const_cast<int *>(r->in_rows);
            cols_out = const_cast<int *>(r->in_cols);
        }
        const int rc = check_coo(rows1, cols1, r->m, r->nnz, rows_out, cols_out, 4, r->lens);
        if (rc) {
            r->status = rc;
            r->t_done = now_s();
            finish(srv, {r});
            continue;
        }
        classify(srv, r);
        {
            std::lock_guard<std::mutex> g(srv->mu);
            srv->queue.push_back(r);
        }
        srv->cv_work.notify_one();
    }
}

static int submit_vr(openph_server *srv, const double *points, int n, int dim, bool borrow,
                     int *low1, int *ess, int64_t *ticket) {
    if (!srv || !ticket || n < 0 || dim < 1 || dim > 7 || (n > 0 && !points)) return OPENPH_EINVAL;
    int64_t m = 0, nnz = 0;
    int rc = openph_vr2_size(n, &m, &nnz);
    if (rc) return rc;
    if (m > 0x7fffffffLL || nnz > 0x7fffffffLL) return OPENPH_EINVAL;
    for (long long k = 0; k < (long long)n * dim; k++)
        if (!std::isfinite(points[k])) return OPENPH_EINVAL;
    auto r = std::make_unique<Request>();
    r->kind = 1;
    r->m = (int)m;
    r->nnz = (int)nnz;
    r->n = n;
    r->pdim = dim;
    if (borrow) {
        r->in_points = points;
    } else {
        r->points.assign(points, points + (size_t)n * dim);
        r->in_points = r->points.data();
    }
    r->out_low = low1;
    r->out_ess = ess;
    r->lens = {0, 2, 3};  // vertices, edges, triangles
    return enqueue(srv, std::move(r), ticket);
}

static int submit_cubical(openph_server *srv, const float *field, int s0, int s1, int s2,
                          bool borrow, int *low1, int *ess, int64_t *ticket) {
    if (!srv || !ticket || !field) return OPENPH_EINVAL;
    int64_t m = 0, nnz = 0;
    int rc = openph_cubical_size(s0, s1, s2, &m, &nnz);
    if (rc) return rc;
    if (m >= (1LL << 28) || nnz > 0x7fffffffLL) return OPENPH_EINVAL;
    const size_t cells = (size_t)s0 * s1 * s2;
    for (size_t k = 0; k < cells; k++)
        if (std::isnan(field[k])) return OPENPH_EINVAL;
    auto r = std::make_unique<Request>();
    r->kind = 2;
    r->m = (int)m;
    r->nnz = (int)nnz;
    r->s0 = s0;
    r->s1 = s1;
    r->s2 = s2;
    if (borrow) {
        r->in_field = field;
    } else {
        r->field.assign(field, field + cells);
        r->in_field = r->field.data();
    }
    r->out_low = low1;
    r->out_ess = ess;
    r->lens = {0, 2, 4, 6};  // vertices, edges, squares, cubes
    return enqueue(srv, std::move(r), ticket);
}

}  // namespace oph

extern "C" {

int openph_server_submit(openph_server *srv, const int *rows1, const int *cols1, int m, int nnz,
                         int64_t *ticket) {
    SubmitClock clock(srv);
    return submit_coo(srv, rows1, cols1, m, nnz, false, nullptr, nullptr, ticket);
}

int openph_server_submit_vr(openph_server *srv, const double *points, int n, int dim,
                            int64_t *ticket) {
    SubmitClock clock(srv);
    return submit_vr(srv, points, n, dim, false, nullptr, nullptr, ticket);
}

int openph_server_submit_cubical(openph_server *srv, const float *field, int s0, int s1, int s2,
                                 int64_t *ticket) {
    SubmitClock clock(srv);
    return submit_cubical(srv, field, s0, s1, s2, false, nullptr, nullptr, ticket);
}

int openph_server_submit_request(openph_server *srv, const openph_request *q, int64_t *ticket) {
    if (!q || (!q->low1) != (!q->ess) || (q->flags & ~OPENPH_BORROW)) return OPENPH_EINVAL;
    const bool borrow = (q->flags & OPENPH_BORROW) != 0;
    SubmitClock clock(srv);
    int rc = OPENPH_EINVAL;
    switch (q->kind) {
        case OPENPH_REQ_COO:
            rc = submit_coo(srv, q->rows1, q->cols1, q->m, q->nnz, borrow, q->low1, q->ess,
                            ticket);
            break;
        case OPENPH_REQ_VR:
            rc = submit_vr(srv, q->points, q->n, q->dim, borrow, q->low1, q->ess, ticket);
            break;
        case OPENPH_REQ_CUBICAL:
            rc = submit_cubical(srv, q->field, q->s0, q->s1, q->s2, borrow, q->low1, q->ess,
                                ticket);
            break;
        default:
            break;
    }
    return rc;
}

int openph_server_wait(openph_server *srv, int64_t ticket, int *low1, int *ess,
                       openph_request_info *info) {
    if (!srv) return OPENPH_EINVAL;
    std::unique_ptr<Request> r;
    {
        std::unique_lock<std::mutex> g(srv->mu);
        auto it = srv->requests.find(ticket);
        // a ticket is waited for once: a second waiter is refused
        if (it == srv->requests.end() || it->second->claimed) return OPENPH_EINVAL;
        Request *p = it->second.get();
        p->claimed = true;
        srv->cv_done.wait(g, [&] { return p->done; });
        it = srv->requests.find(ticket);  // the map may have changed meanwhile
        r = std::move(it->second);
        srv->requests.erase(it);
    }
    if (info) {
        info->submit_s = r->t_submit;
        info->start_s = r->t_start;
        info->done_s = r->t_done;
        info->batch_size = r->batch_size;
        info->worker = r->worker;
        info->num_iters = r->num_iters;
        info->status = r->status;
        info->m = r->m;
    }
    if (r->status) return r->status;
    const int *l = r->out_low ? r->out_low : r->low.data();
    const int *e = r->out_ess ? r->out_ess : r->ess.data();
    if (r->m && low1 && low1 != l) std::memcpy(low1, l, sizeof(int) * r->m);
    if (r->m && ess && ess != e) std::memcpy(ess, e, sizeof(int) * r->m);
    return OPENPH_OK;
}

int openph_server_stats(openph_server *srv, int64_t *queued, int64_t *batches,
                        int64_t *completed) {
    if (!srv) return OPENPH_EINVAL;
    std::lock_guard<std::mutex> g(srv->mu);
    if (queued) *queued = (int64_t)(srv->queue.size() + srv->intake.size());
    if (batches) *batches = srv->batches;
    if (completed) *completed = srv->completed;
    return OPENPH_OK;
}

int openph_server_destroy(openph_server *srv) {
    if (!srv) return OPENPH_EINVAL;
    // the intake threads queue what they hold, then the workers drain the queue
    {
        std::lock_guard<std::mutex> g(srv->mu);
        srv->stop_intake = true;
    }
    srv->cv_intake.notify_all();
    for (std::thread &t : srv->intakers)
        if (t.joinable()) t.join();
    {
        std::lock_guard<std::mutex> g(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    for (auto &w : srv->workers)
        if (w->thread.joinable()) w->thread.join();
    if (std::getenv("OPENPH_DEBUG"))
        std::fprintf(stderr, "openph server: %lld requests submitted in %.3f s\n", srv->n_submit,
                     srv->t_submit);
    (void)hipSetDevice(srv->opt.device);  // the workspaces are freed on this thread
    delete srv;
    return OPENPH_OK;
}

}  // extern "C"
