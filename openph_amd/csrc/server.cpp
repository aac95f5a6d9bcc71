// Served pms: a request loop that batches arrivals of different sizes onto
// several device workers.
//
// The reference has no serving mode. Design:
//   * submit() copies a request's 1-based COO into the server and queues it;
//     it never blocks on the GPU. A full queue refuses with OPENPH_EBUSY.
//   * W worker threads each own a device workspace (Context) and a HIP
//     stream. A free worker takes queued requests in arrival order as long
//     as their pool footprint fits its byte budget (admission by footprint;
//     the first request is always taken, so an oversized request runs alone)
//     and reduces them as one block-diagonal batch. Batches of different
//     workers run concurrently on the device; requests that arrive while
//     batches run join the next batch a worker forms (continuous batching).
//   * The matrices of a batch stay in HBM for the whole reduction; only the
//     COO goes in and low/ess come out, through pinned staging buffers.
//   * wait() blocks until a request is done and returns its outputs with its
//     submit / start / finish times.
// Results are unique for a matrix, so a request's result is the same alone
// and in any batch (tests/test_gpu_server.py checks it).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
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

struct Request {
    int64_t id = 0;
    int m = 0, nnz = 0;
    // the matrix: given as 1-based COO (kind 0), or built on the worker's
    // stream from a point cloud (kind 1, VR 2-skeleton) or a grid (kind 2,
    // cubical lower-star)
    int kind = 0;
    std::vector<int> rows, cols;
    std::vector<double> points;
    std::vector<float> field;
    int n = 0, pdim = 0, s0 = 0, s1 = 0, s2 = 0;
    std::vector<int> low, ess;
    size_t footprint = 0;
    double t_submit = 0, t_start = 0, t_done = 0;
    int status = OPENPH_OK;
    int batch_size = 0, worker = -1, num_iters = 0;
    bool done = false;
};

struct Worker {
    int index = 0;
    Context ctx;
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

// Device bytes a request needs inside a batch: COO staging, the column pool
// (slot-granular layout plus the default headroom) and the per-column state.
static size_t footprint(int m, int nnz, int slot0) {
    size_t pool = (size_t)nnz + (size_t)m * (size_t)slot0 + (size_t)nnz * 8;
    return 4 * (2 * (size_t)nnz + pool) + 84 * (size_t)m;
}

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

static int run_worker_batch(openph_server *srv, Worker &w, std::vector<Request *> &batch) {
    Context &c = w.ctx;
    const int count = (int)batch.size();
    std::vector<int> ms(count), nnzs(count);
    size_t N = 0, M = 0;
    for (int k = 0; k < count; k++) {
        ms[k] = batch[k]->m;
        nnzs[k] = batch[k]->nnz;
        N += batch[k]->nnz;
        M += batch[k]->m;
    }
    if (N > 0x7fffffffull || M > 0x7fffffffull) return OPENPH_EINVAL;
    OPH_CHECK(grow_pinned(&w.h_rows, w.c_rows, N + 1));
    OPH_CHECK(grow_pinned(&w.h_cols, w.c_cols, N + 1));
    OPH_CHECK(grow_pinned(&w.h_low, w.c_low, M + 1));
    OPH_CHECK(grow_pinned(&w.h_ess, w.c_ess, M + 1));
    OPH_CHECK(grow(&c.h_rows, c.c_hrows, N + 1));
    OPH_CHECK(grow(&c.h_cols, c.c_hcols, N + 1));
    OPH_CHECK(grow(&w.d_low, w.c_dlow, M + 1));
    OPH_CHECK(grow(&w.d_ess, w.c_dess, M + 1));
    size_t o = 0, in_bytes = 0;
    for (Request *r : batch) {
        if (r->kind == 0 && r->nnz) {
            std::memcpy(w.h_rows + o, r->rows.data(), sizeof(int) * r->nnz);
            std::memcpy(w.h_cols + o, r->cols.data(), sizeof(int) * r->nnz);
        }
        if (r->kind == 1) in_bytes += ((r->points.size() * sizeof(double)) + 255) & ~(size_t)255;
        if (r->kind == 2) in_bytes += ((r->field.size() * sizeof(float)) + 255) & ~(size_t)255;
        o += r->nnz;
    }
    hipStream_t s = w.stream;
    // the given COO goes in as one copy; the built matrices are then written
    // over their own ranges on the same stream
    if (N) {
        OPH_CHECK(hipMemcpyAsync(c.h_rows, w.h_rows, sizeof(int) * N, hipMemcpyHostToDevice, s));
        OPH_CHECK(hipMemcpyAsync(c.h_cols, w.h_cols, sizeof(int) * N, hipMemcpyHostToDevice, s));
    }
    if (in_bytes) {
        OPH_CHECK(grow_pinned(&w.h_in, w.c_hin, in_bytes));
        OPH_CHECK(grow(&w.d_in, w.c_din, in_bytes));
        size_t b = 0;
        for (Request *r : batch) {
            if (r->kind == 1) {
                std::memcpy(w.h_in + b, r->points.data(), r->points.size() * sizeof(double));
                b += ((r->points.size() * sizeof(double)) + 255) & ~(size_t)255;
            } else if (r->kind == 2) {
                std::memcpy(w.h_in + b, r->field.data(), r->field.size() * sizeof(float));
                b += ((r->field.size() * sizeof(float)) + 255) & ~(size_t)255;
            }
        }
        OPH_CHECK(hipMemcpyAsync(w.d_in, w.h_in, in_bytes, hipMemcpyHostToDevice, s));
        size_t ro = 0;
        b = 0;
        for (Request *r : batch) {
            int rc = OPENPH_OK;
            if (r->kind == 1) {
                rc = build_vr(c, reinterpret_cast<const double *>(w.d_in + b), r->n, r->pdim,
                              c.h_rows + ro, c.h_cols + ro, nullptr, nullptr, s);
                b += ((r->points.size() * sizeof(double)) + 255) & ~(size_t)255;
            } else if (r->kind == 2) {
                rc = build_cubical(c, reinterpret_cast<const float *>(w.d_in + b), r->s0, r->s1,
                                   r->s2, c.h_rows + ro, c.h_cols + ro, nullptr, nullptr, s);
                b += ((r->field.size() * sizeof(float)) + 255) & ~(size_t)255;
            }
            if (rc) {
                (void)hipStreamSynchronize(s);
                return rc;
            }
            ro += r->nnz;
        }
    }
    RunCfg cfg = cfg_from(&srv->opt.reduce, (int)N);
    openph_stats st;
    std::memset(&st, 0, sizeof(st));
    int rc = run_batch(c, count, ms.data(), nnzs.data(), c.h_rows, c.h_cols, w.d_low, w.d_ess, s,
                       cfg, &st);
    if (rc) {
        (void)hipStreamSynchronize(s);
        return rc;
    }
    if (M) {
        OPH_CHECK(hipMemcpyAsync(w.h_low, w.d_low, sizeof(int) * M, hipMemcpyDeviceToHost, s));
        OPH_CHECK(hipMemcpyAsync(w.h_ess, w.d_ess, sizeof(int) * M, hipMemcpyDeviceToHost, s));
    }
    OPH_CHECK(hipStreamSynchronize(s));
    size_t mo = 0;
    for (Request *r : batch) {
        r->low.assign(w.h_low + mo, w.h_low + mo + r->m);
        r->ess.assign(w.h_ess + mo, w.h_ess + mo + r->m);
        r->num_iters = st.num_iters;
        mo += r->m;
    }
    return OPENPH_OK;
}

static void worker_main(openph_server *srv, Worker *w) {
    (void)hipSetDevice(srv->opt.device);
    w->ctx.device = srv->opt.device;
    if (hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking) != hipSuccess) w->stream = nullptr;
    const size_t budget = (size_t)srv->opt.worker_bytes;
    const int slot0 = srv->opt.reduce.slot_elems > 0 ? srv->opt.reduce.slot_elems : 32;
    {
        // Size the workspace for a full budget up front, taking the typical
        // nnz = 3 m of the boundary matrices served here (footprint() per
        // column, about 44 x 3 + 4 x slot0 + 84 bytes).
        const size_t per_col = 44 * 3 + 4 * (size_t)slot0 + 84;
        const size_t m_cap = budget / per_col, nnz_cap = 3 * m_cap;
        const size_t pool_cap = nnz_cap + m_cap * (size_t)slot0 + 8 * nnz_cap;
        (void)reserve(w->ctx, m_cap, nnz_cap, std::min<size_t>(pool_cap, 0xFFFFFFFFull));
        (void)grow_pinned(&w->h_rows, w->c_rows, nnz_cap + 1);
        (void)grow_pinned(&w->h_cols, w->c_cols, nnz_cap + 1);
        (void)grow_pinned(&w->h_low, w->c_low, m_cap + 1);
        (void)grow_pinned(&w->h_ess, w->c_ess, m_cap + 1);
        (void)grow(&w->d_low, w->c_dlow, m_cap + 1);
        (void)grow(&w->d_ess, w->c_dess, m_cap + 1);
        std::lock_guard<std::mutex> g(srv->mu);
        srv->ready++;
    }
    srv->cv_done.notify_all();
    while (true) {
        std::vector<Request *> batch;
        {
            std::unique_lock<std::mutex> g(srv->mu);
            srv->cv_work.wait(g, [&] { return srv->stop || !srv->queue.empty(); });
            if (srv->stop && srv->queue.empty()) break;
            size_t used = 0;
            while (!srv->queue.empty() && (int)batch.size() < srv->opt.max_batch) {
                Request *r = srv->queue.front();
                if (!batch.empty() && used + r->footprint > budget) break;
                used += r->footprint;
                batch.push_back(r);
                srv->queue.pop_front();
            }
        }
        const double t0 = now_s();
        int rc = w->stream ? run_worker_batch(srv, *w, batch) : OPENPH_EDEVICE;
        const double t1 = now_s();
        {
            std::lock_guard<std::mutex> g(srv->mu);
            for (Request *r : batch) {
                r->t_start = t0;
                r->t_done = t1;
                r->status = rc;
                r->batch_size = (int)batch.size();
                r->worker = w->index;
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
            srv->batches++;
            srv->completed += (int64_t)batch.size();
        }
        srv->cv_done.notify_all();
    }
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

// Queues a request; its footprint decides which batch it may join.
static int enqueue(openph_server *srv, std::unique_ptr<Request> r, int64_t *ticket) {
    const int slot0 = srv->opt.reduce.slot_elems > 0 ? srv->opt.reduce.slot_elems : 32;
    r->footprint = footprint(r->m, r->nnz, slot0);
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

}  // namespace oph

extern "C" {

int openph_server_submit(openph_server *srv, const int *rows1, const int *cols1, int m, int nnz,
                         int64_t *ticket) {
    if (!srv || !ticket || m < 0 || nnz < 0 || (nnz > 0 && (!rows1 || !cols1)))
        return OPENPH_EINVAL;
    // Validate here, so that a malformed request is refused alone instead of
    // failing the batch it would join (same rules as the device setup).
    for (int e = 0; e < nnz; e++) {
        const int c = cols1[e], rr = rows1[e];
        if (rr < 1 || c > m || rr >= c) return OPENPH_EINVAL;
        if (e > 0) {
            if (c < cols1[e - 1]) return OPENPH_EUNSORTED;
            if (c == cols1[e - 1] && rr <= rows1[e - 1]) return OPENPH_EUNSORTED;
        }
    }
    auto r = std::make_unique<Request>();
    r->m = m;
    r->nnz = nnz;
    r->rows.assign(rows1, rows1 + nnz);
    r->cols.assign(cols1, cols1 + nnz);
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
    return enqueue(srv, std::move(r), ticket);
}

int openph_server_wait(openph_server *srv, int64_t ticket, int *low1, int *ess,
                       openph_request_info *info) {
    if (!srv) return OPENPH_EINVAL;
    std::unique_ptr<Request> r;
    {
        std::unique_lock<std::mutex> g(srv->mu);
        auto it = srv->requests.find(ticket);
        if (it == srv->requests.end()) return OPENPH_EINVAL;
        Request *p = it->second.get();
        srv->cv_done.wait(g, [&] { return p->done; });
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
