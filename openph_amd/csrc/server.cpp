// Served pms: a request loop that reduces arrivals of different sizes
// together on several device workers, with continuous batching at iteration
// granularity.
//
// The reference has no serving mode. Design:
//   * submit() validates a request's 1-based COO (so that a malformed one is
//     refused alone), records its distinct column lengths and its shape class
//     (input columns of at most 4 entries, or longer), copies it and queues
//     it; it never blocks on the GPU. A full queue refuses with OPENPH_EBUSY.
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
    std::thread thread;

    ~Worker() {
        int *pinned[] = {h_rows, h_cols, h_low, h_ess};
        for (int *p : pinned)
            if (p) (void)hipHostFree(p);
        if (h_in) (void)hipHostFree(h_in);
        if (d_low) (void)hipFree(d_low);
        if (d_ess) (void)hipFree(d_ess);
        if (d_in) (void)hipFree(d_in);
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
    int64_t next_id = 1;
    bool stop = false;
    int ready = 0;  // workers whose workspace is reserved
    // statistics
    int64_t batches = 0, completed = 0;
    long long columns_done = 0;
};

namespace oph {

constexpr int kChunkIters = 2;  // iterations between two admissions / retirements

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
    if (r->kind == 1) return ((r->points.size() * sizeof(double)) + 255) & ~(size_t)255;
    if (r->kind == 2) return ((r->field.size() * sizeof(float)) + 255) & ~(size_t)255;
    return 0;
}

// The local 1-based COO of `reqs`, concatenated, into the device buffers
// rows1/cols1 (given COO copied in, the others built on the device).
static int stage(Worker &w, const std::vector<Request *> &reqs, int *rows1, int *cols1) {
    Context &c = w.ctx;
    hipStream_t s = w.stream;
    size_t N = 0, inb = 0;
    for (Request *r : reqs) {
        N += r->nnz;
        inb += in_bytes(r);
    }
    OPH_CHECK(grow_pinned(&w.h_rows, w.c_rows, N + 1));
    OPH_CHECK(grow_pinned(&w.h_cols, w.c_cols, N + 1));
    size_t o = 0;
    for (Request *r : reqs) {
        if (r->kind == 0 && r->nnz) {
            std::memcpy(w.h_rows + o, r->rows.data(), sizeof(int) * r->nnz);
            std::memcpy(w.h_cols + o, r->cols.data(), sizeof(int) * r->nnz);
        }
        o += r->nnz;
    }
    // the given COO goes in as one copy; the built matrices are then written
    // over their own ranges on the same stream
    if (N) {
        OPH_CHECK(hipMemcpyAsync(rows1, w.h_rows, sizeof(int) * N, hipMemcpyHostToDevice, s));
        OPH_CHECK(hipMemcpyAsync(cols1, w.h_cols, sizeof(int) * N, hipMemcpyHostToDevice, s));
    }
    if (inb) {
        OPH_CHECK(grow_pinned(&w.h_in, w.c_hin, inb));
        OPH_CHECK(grow(&w.d_in, w.c_din, inb));
        size_t b = 0;
        for (Request *r : reqs) {
            if (r->kind == 1) std::memcpy(w.h_in + b, r->points.data(), r->points.size() * sizeof(double));
            if (r->kind == 2) std::memcpy(w.h_in + b, r->field.data(), r->field.size() * sizeof(float));
            b += in_bytes(r);
        }
        OPH_CHECK(hipMemcpyAsync(w.d_in, w.h_in, inb, hipMemcpyHostToDevice, s));
        size_t ro = 0;
        b = 0;
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
            srv->columns_done += r->m;
        }
        srv->completed += (int64_t)done.size();
    }
    srv->cv_done.notify_all();
}

// One request reduced alone through the batch path (it may grow the
// workspace): requests too large for the running batches, and the members
// of a batch that failed.
static void run_alone(openph_server *srv, Worker &w, Request *r) {
    Context &c = w.ctx;
    hipStream_t s = w.stream;
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
    if (!rc) rc = stage(w, {r}, c.h_rows, c.h_cols);
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
    if (!rc) {
        r->low.assign(w.h_low, w.h_low + M);
        r->ess.assign(w.h_ess, w.h_ess + M);
    }
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

static void worker_main(openph_server *srv, Worker *w) {
    (void)hipSetDevice(srv->opt.device);
    w->ctx.device = srv->opt.device;
    if (hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking) != hipSuccess) w->stream = nullptr;
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
    while (true) {
        // admission: the oldest request starts a batch; requests of the
        // running batch's class join it while they fit
        std::vector<Request *> round;
        Request *alone = nullptr;
        const double tw0 = now_s();
        {
            std::unique_lock<std::mutex> g(srv->mu);
            if (!lb.running) {
                srv->cv_work.wait(g, [&] { return srv->stop || !srv->queue.empty(); });
                if (srv->queue.empty()) break;  // stopping
            }
            std::vector<int> slots;
            for (size_t L = 0; L < lb.slot_of_len.size(); L++)
                if (lb.running && lb.slot_of_len[L] >= 0) slots.push_back((int)L);
            int cls = lb.running ? lb.cls : -1;
            size_t cols = lb.running ? (size_t)lb.mstart.back() : 0, nnz = 0;
            size_t pool = lb.running ? (size_t)c.h_ctrl->pool_top : 0;
            int members = lb.running ? (int)w->members.size() : 0;
            for (auto it = srv->queue.begin(); it != srv->queue.end() && w->stream;) {
                Request *r = *it;
                if (cls < 0) {
                    cls = r->cls;
                    if (!fits(*w, r, cls, 0, 0, 0, {}, 0, max_batch)) {
                        alone = r;  // too large for a batch: runs alone
                        srv->queue.erase(it);
                        break;
                    }
                }
                if (fits(*w, r, cls, cols, nnz, pool, slots, members, max_batch)) {
                    round.push_back(r);
                    cols += r->m;
                    nnz += r->nnz;
                    pool += r->layout;
                    members++;
                    for (int L : r->lens)
                        if (std::find(slots.begin(), slots.end(), L) == slots.end())
                            slots.push_back(L);
                    it = srv->queue.erase(it);
                } else {
                    ++it;
                }
            }
            if (!round.empty() || alone) srv->batches++;
        }
        w->t_wait += now_s() - tw0;
        if (!w->stream) {
            for (Request *r : round) r->status = OPENPH_EDEVICE;
            if (alone) alone->status = OPENPH_EDEVICE;
            if (alone) round.push_back(alone);
            finish(srv, round);
            continue;
        }
        if (alone) {
            const double ta = now_s();
            run_alone(srv, *w, alone);
            w->t_alone += now_s() - ta;
            continue;
        }
        if (!round.empty()) {
            const double t0 = now_s();
            int rc = OPENPH_OK;
            if (!lb.running) {
                rc = live_start(c, lb, cfg, round[0]->cls, max_batch, w->pool_cap, s);
                w->admitted = 0;
            }
            w->admitted += (int)round.size();
            if (!rc) rc = stage(*w, round, c.h_rows, c.h_cols);
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
                rc = live_append(c, lb, (int)round.size(), ms.data(), nnzs.data(), lens.data(),
                                 c.h_rows, c.h_cols, N, s);
            }
            w->t_stage += now_s() - t0;
            for (Request *r : round) {
                r->t_start = t0;
                r->join_iter = c.h_ctrl->iter;
                r->worker = w->index;
                w->members.push_back(r);
                w->retired.push_back(0);
            }
            if (rc) {
                (void)hipStreamSynchronize(s);
                fail_batch(srv, *w);
                continue;
            }
        }
        if (!lb.running) continue;
        // iterations, then retirements
        const double ti = now_s();
        int rc = live_chunk(c, lb, cfg, kChunkIters, s);
        const double tr = now_s();
        w->t_iter += tr - ti;
        w->n_chunks++;
        const Ctrl &h = *c.h_ctrl;
        if (rc) {
            fail_batch(srv, *w);
            continue;
        }
        const bool all = h.halt == kHaltConverged;
        std::vector<int> ret;
        std::vector<Request *> over;  // members past the iteration cap
        for (size_t k = 0; k < w->members.size(); k++) {
            if (w->retired[k]) continue;
            const Request *r = w->members[k];
            if (all || (h.iter > r->join_iter && lb.h_mdup[k] == 0)) ret.push_back((int)k);
            else if (h.iter - r->join_iter > max_iters) over.push_back(w->members[k]);
        }
        if (!over.empty()) {
            // the cap is per request: a member past it fails alone, the rest go on
            std::vector<Request *> done;
            for (size_t k = 0; k < w->members.size(); k++)
                for (Request *r : over)
                    if (w->members[k] == r && !w->retired[k]) {
                        (void)live_retire(c, lb, (int)k, w->d_low + lb.mstart[k],
                                          w->d_ess + lb.mstart[k], s);
                        w->retired[k] = 1;
                        r->status = OPENPH_EITER;
                        r->t_done = now_s();
                        done.push_back(r);
                    }
            finish(srv, done);
        }
        if (!ret.empty()) {
            for (int k : ret)
                if ((rc = live_retire(c, lb, k, w->d_low + lb.mstart[k], w->d_ess + lb.mstart[k], s)))
                    break;
            for (int k : ret) {
                const int a = lb.mstart[k], n = lb.mstart[k + 1] - a;
                if (rc || !n) continue;
                if (hipMemcpyAsync(w->h_low + a, w->d_low + a, sizeof(int) * n,
                                   hipMemcpyDeviceToHost, s) ||
                    hipMemcpyAsync(w->h_ess + a, w->d_ess + a, sizeof(int) * n,
                                   hipMemcpyDeviceToHost, s))
                    rc = OPENPH_EDEVICE;
            }
            if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = OPENPH_EDEVICE;
            if (rc) {
                fail_batch(srv, *w);
                continue;
            }
            const double t1 = now_s();
            std::vector<Request *> done;
            for (int k : ret) {
                Request *r = w->members[k];
                const int a = lb.mstart[k];
                r->low.assign(w->h_low + a, w->h_low + a + r->m);
                r->ess.assign(w->h_ess + a, w->h_ess + a + r->m);
                r->status = OPENPH_OK;
                r->num_iters = h.iter - r->join_iter;
                r->batch_size = w->admitted;
                r->t_done = t1;
                w->retired[k] = 1;
                done.push_back(r);
            }
            finish(srv, done);
        }
        w->t_retire += now_s() - tr;
        bool any = false;
        for (char x : w->retired) any |= !x;
        if (!any || all) {
            lb.running = false;
            w->members.clear();
            w->retired.clear();
        } else {
            live_trim(c, lb, w->retired, s);
            w->members.resize(lb.mstart.size() - 1);
            w->retired.resize(lb.mstart.size() - 1);
        }
    }
    if (std::getenv("OPENPH_DEBUG"))
        std::fprintf(stderr,
                     "openph server worker %d: wait %.3f s, stage %.3f s, iterations %.3f s "
                     "(%lld chunks), retire %.3f s, alone %.3f s\n",
                     w->index, w->t_wait, w->t_stage, w->t_iter, w->n_chunks, w->t_retire,
                     w->t_alone);
    if (w->stream) (void)hipStreamDestroy(w->stream);
}

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
    openph_server *srv = new openph_server();
    srv->opt = o;
    for (int i = 0; i < o.workers; i++) {
        auto w = std::make_unique<Worker>();
        w->index = i;
        srv->workers.push_back(std::move(w));
    }
    for (auto &w : srv->workers) w->thread = std::thread(worker_main, srv, w.get());
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
static int enqueue(openph_server *srv, std::unique_ptr<Request> r, int64_t *ticket) {
    const int slot0 = srv->opt.reduce.slot_elems > 0 ? srv->opt.reduce.slot_elems : 32;
    r->layout = (size_t)r->nnz + (size_t)r->m * (size_t)slot0;
    int maxlen = 0;
    for (int L : r->lens) maxlen = std::max(maxlen, L);
    r->cls = maxlen <= 4 ? 0 : 1;
    {
        std::lock_guard<std::mutex> g(srv->mu);
        if ((int)srv->queue.size() >= srv->opt.max_queue) return OPENPH_EBUSY;
        r->id = srv->next_id++;
        r->t_submit = now_s();
        *ticket = r->id;
        srv->queue.push_back(r.get());
        srv->requests[r->id] = std::move(r);
    }
    srv->cv_work.notify_one();
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
    std::memcpy(rows_out + e0, rows1 + e0, sizeof(int) * (size_t)(e1 - e0));
    std::memcpy(cols_out + e0, cols1 + e0, sizeof(int) * (size_t)(e1 - e0));
    return OPENPH_OK;
}

}  // namespace oph

extern "C" {

int openph_server_submit(openph_server *srv, const int *rows1, const int *cols1, int m, int nnz,
                         int64_t *ticket) {
    if (!srv || !ticket || m < 0 || nnz < 0 || (nnz > 0 && (!rows1 || !cols1)))
        return OPENPH_EINVAL;
    auto r = std::make_unique<Request>();
    r->m = m;
    r->nnz = nnz;
    r->rows.resize((size_t)nnz);
    r->cols.resize((size_t)nnz);
    // Validate here, so that a malformed request is refused alone instead of
    // failing the batch it would join (same rules as the device set-up), and
    // record its distinct column lengths (pseudo-dimension + 1): a batch
    // holds at most OPENPH_MAX_DIMS of them over all its members. A large
    // request is taken in stretches by several threads, each stretch ending
    // at a column boundary.
    const int nt = nnz >= (1 << 21) ? 8 : 1;
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
                rc[k] = intake(rows1, cols1, m, cut[k], cut[k + 1], r->rows.data(),
                               r->cols.data(), lens[k], ncols[k]);
            });
        rc[0] = intake(rows1, cols1, m, cut[0], cut[1], r->rows.data(), r->cols.data(), lens[0],
                       ncols[0]);
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
            if (std::find(r->lens.begin(), r->lens.end(), L) == r->lens.end()) r->lens.push_back(L);
    }
    if (cols_seen < m && std::find(r->lens.begin(), r->lens.end(), 0) == r->lens.end())
        r->lens.push_back(0);
    if ((int)r->lens.size() > OPENPH_MAX_DIMS) return OPENPH_EDIMS;
    return enqueue(srv, std::move(r), ticket);
}

int openph_server_submit_vr(openph_server *srv, const double *points, int n, int dim,
                            int64_t *ticket) {
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
    r->points.assign(points, points + (size_t)n * dim);
    r->lens = {0, 2, 3};  // vertices, edges, triangles
    return enqueue(srv, std::move(r), ticket);
}

int openph_server_submit_cubical(openph_server *srv, const float *field, int s0, int s1, int s2,
                                 int64_t *ticket) {
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
    r->field.assign(field, field + cells);
    r->lens = {0, 2, 4, 6};  // vertices, edges, squares, cubes
    return enqueue(srv, std::move(r), ticket);
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
    if (r->m && low1) std::memcpy(low1, r->low.data(), sizeof(int) * r->m);
    if (r->m && ess) std::memcpy(ess, r->ess.data(), sizeof(int) * r->m);
    return OPENPH_OK;
}

int openph_server_stats(openph_server *srv, int64_t *queued, int64_t *batches,
                        int64_t *completed) {
    if (!srv) return OPENPH_EINVAL;
    std::lock_guard<std::mutex> g(srv->mu);
    if (queued) *queued = (int64_t)srv->queue.size();
    if (batches) *batches = srv->batches;
    if (completed) *completed = srv->completed;
    return OPENPH_OK;
}

int openph_server_destroy(openph_server *srv) {
    if (!srv) return OPENPH_EINVAL;
    {
        std::lock_guard<std::mutex> g(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    for (auto &w : srv->workers)
        if (w->thread.joinable()) w->thread.join();
    (void)hipSetDevice(srv->opt.device);  // the workspaces are freed on this thread
    delete srv;
    return OPENPH_OK;
}

}  // extern "C"
