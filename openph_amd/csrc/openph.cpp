// Host orchestration and C ABI of the gfx950 pms reduction.
//
// Mirrors the reference's host path (src/cuda/openph.cu:29-237 and the pms
// branch of src/cuda/algorithms/algorithm_factory.cu:124-146) with these
// differences, all deliberate:
//   * device workspace is owned by a per-device context and reused across
//     calls instead of ~25 cudaMalloc/cudaFree per call;
//   * the COO -> column store, low, left/beta and pseudo-dimension setup run
//     on the device (the reference does dimensions and beta on the host,
//     dimensions.cu:54-104, beta.cu:17-26);
//   * columns live in an adaptive pool, so col_width is not needed;
//   * the loop reads back one small control block per iteration instead of
//     a device symbol plus five cuBLAS scalars (pms.cu:92, tracker.cu:92-134);
//   * errors are reported, never exit().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "host.h"

namespace oph {

static Context *context_for(int device) {
    static std::mutex gmu;
    static std::vector<Context *> ctxs;
    std::lock_guard<std::mutex> g(gmu);
    if (device < 0) return nullptr;
    if ((size_t)device >= ctxs.size()) ctxs.resize(device + 1, nullptr);
    if (!ctxs[device]) {
        ctxs[device] = new Context();
        ctxs[device]->device = device;
    }
    return ctxs[device];
}

static int ensure(Context &c, int m, int nnz, size_t pool_elems) {
    Dev &d = c.d;
    size_t mm = (size_t)m + 1;
    OPH_CHECK(grow(&d.low, c.c_low, mm));
    OPH_CHECK(grow(&d.arglow, c.c_arglow, mm));
    OPH_CHECK(grow(&d.classes, c.c_classes, mm));
    OPH_CHECK(grow(&d.ess, c.c_ess, mm));
    OPH_CHECK(grow(&d.left, c.c_left, mm));
    OPH_CHECK(grow(&d.beta, c.c_beta, mm));
    OPH_CHECK(grow(&d.first, c.c_first, mm));
    OPH_CHECK(grow(&d.clear, c.c_clear, mm));
    OPH_CHECK(grow(&d.newpiv, c.c_newpiv, mm));
    OPH_CHECK(grow(&d.slot, c.c_slot, mm));
    OPH_CHECK(grow(&d.len, c.c_len, mm));
    OPH_CHECK(grow(&d.off, c.c_off, mm));
    OPH_CHECK(grow(&d.off2, c.c_off2, mm));
    OPH_CHECK(grow(&d.cap, c.c_cap, mm));
    OPH_CHECK(grow(&d.work, c.c_work, mm));
    OPH_CHECK(grow(&d.pivdesc, c.c_pivdesc, mm));
    OPH_CHECK(grow(&d.longw, c.c_longw, mm));
    OPH_CHECK(grow(&d.slotmap, c.cap_slotmap, mm + 1));
    size_t nt = (size_t)(m + kTile - 1) / kTile + 1;
    OPH_CHECK(grow(&d.tile_max, c.c_tmax, nt * kMaxSlots));
    OPH_CHECK(grow(&d.tile_pre, c.c_tpre, nt * kMaxSlots));
    OPH_CHECK(grow(&d.scan_tmp, c.cap_scan, nt + 1));
    OPH_CHECK(grow(&d.ctrl, c.c_ctrl, 1));
    if (pool_elems < (size_t)nnz) pool_elems = (size_t)nnz;
    OPH_CHECK(grow(&d.pool, c.cap_pool, pool_elems));
    (void)c.cap_pool2;
    return OPENPH_OK;
}

int reserve(Context &c, size_t m, size_t nnz, size_t pool_elems) {
    if (m > 0x7ffffffeull || nnz > 0x7ffffffeull) return OPENPH_EINVAL;
    int rc = ensure(c, (int)m, (int)nnz, pool_elems);
    if (rc) return rc;
    OPH_CHECK(grow(&c.h_rows, c.c_hrows, nnz + 1));
    OPH_CHECK(grow(&c.h_cols, c.c_hcols, nnz + 1));
    OPH_CHECK(grow(&c.b_rows, c.c_brows, nnz + 1));
    OPH_CHECK(grow(&c.b_cols, c.c_bcols, nnz + 1));
    OPH_CHECK(grow(&c.b_err, c.c_berr, 1));
    return OPENPH_OK;
}

RunCfg cfg_from(const openph_options *opt, int nnz) {
    RunCfg r;
    (void)nnz;
    r.pool_elems = 0;  // automatic: the initial layout plus headroom
    if (opt) {
        if (opt->max_iters > 0) r.max_iters = opt->max_iters;
        if (opt->pool_max_elems > 0) r.pool_max = std::min<size_t>((size_t)opt->pool_max_elems, 0xFFFFFFFFull);
        if (opt->pool_elems > 0) r.pool_elems = (size_t)opt->pool_elems;
        r.lanes = opt->lanes_per_column;
        r.wlds = opt->group_lds;
        r.track = opt->track != 0;
        if (opt->slot_elems > 0) r.slot0 = std::min(opt->slot_elems, 1 << 16);
    }
    if (r.pool_elems > r.pool_max) r.pool_elems = r.pool_max;
    if (r.lanes != 0) {
        if (r.lanes != 1 && !r.wlds) r.wlds = phase2_default_lds(r.lanes);
        if (!phase2_shape_ok(r.lanes, r.wlds)) r.lanes = -1;  // rejected by run_pms
    }
    return r;
}

static int err_from_bits(int bits) {
    if (bits & (kErrRange | kErrState)) return OPENPH_EINVAL;
    if (bits & (kErrOrder | kErrRows)) return OPENPH_EUNSORTED;
    if (bits & kErrSlots) return OPENPH_EDIMS;
    return OPENPH_OK;
}

// Pool compaction, growing the pool when the live data fills more than half
// of it. Returns OPENPH_ENOMEM when the pool budget cannot hold the live data
// with room to make progress.
static int compact(Context &c, const RunCfg &cfg, hipStream_t s, int *ncompact, size_t &lcap) {
    Dev &d = c.d;
    launch_compact_scan(d, s);
    OPH_CHECK(hipMemcpyAsync(&c.h_word[0], &d.ctrl->pool_top, sizeof(unsigned long long),
                             hipMemcpyDeviceToHost, s));
    OPH_CHECK(hipStreamSynchronize(s));
    const unsigned long long live = c.h_word[0];
    size_t cap = lcap;
    size_t want = cap;
    if (live * 2 > cap) want = std::max<size_t>(cap * 2, live * 2);
    if (want > cfg.pool_max) want = cfg.pool_max;
    if (live + live / 16 + 1024 > want) return OPENPH_ENOMEM;
    OPH_CHECK(grow(&d.pool2, c.cap_pool2, want));
    launch_compact_copy(d, s);
    std::swap(c.cap_pool, c.cap_pool2);
    // the kernels now use pool2's buffer, logically `want` entries long
    lcap = want;
    c.h_word[1] = lcap;
    OPH_CHECK(hipMemcpyAsync(&d.ctrl->pool_cap, &c.h_word[1], sizeof(unsigned long long),
                             hipMemcpyHostToDevice, s));
    OPH_CHECK(hipStreamSynchronize(s));  // h_word[1] may be rewritten by the next compaction
    (*ncompact)++;
    return OPENPH_OK;
}

// The pms driver (pms.cu:1-118) on device-resident 1-based COO.
static int run_pms(Context &c, const int *rows1, const int *cols1, int m, int nnz, int *low1,
                   int *ess, hipStream_t s, const RunCfg &cfg, openph_stats *stats,
                   float *time_track) {
    if (m < 0 || nnz < 0) return OPENPH_EINVAL;
    if (m == 0) {
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->num_iters = 1;
        }
        return OPENPH_OK;
    }
    if (cfg.lanes < 0) return OPENPH_EINVAL;
    // the initial layout gives every column a segment of >= slot0 entries
    const size_t layout = (size_t)nnz + (size_t)m * (size_t)cfg.slot0;
    if (layout > cfg.pool_max) return OPENPH_ENOMEM;
    size_t pool_want = cfg.pool_elems
                           ? std::max(cfg.pool_elems, layout)
                           : layout + std::max<size_t>((size_t)nnz * 8, (size_t)1 << 25);
    pool_want = std::min(pool_want, cfg.pool_max);
    int rc = ensure(c, m, nnz, pool_want);
    if (rc) return rc;
    Dev &d = c.d;
    d.m = m;
    d.nnz = nnz;
    d.ntiles = (m + kTile - 1) / kTile;
    d.slot0 = cfg.slot0;
    if (!c.num_cu) {
        hipDeviceProp_t prop;
        OPH_CHECK(hipGetDeviceProperties(&prop, c.device));
        c.num_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    }
    if (!c.ev0) {
        OPH_CHECK(hipEventCreate(&c.ev0));
        OPH_CHECK(hipEventCreate(&c.ev1));
    }
    // logical pool size: the requested one, even when the context holds a
    // larger buffer from an earlier call
    size_t lcap = std::min(c.cap_pool, pool_want);
    launch_ctrl_init(d, (unsigned long long)lcap, s);
    launch_setup(d, rows1, cols1, s);
    // the control block is read back into pinned memory: a copy to pageable
    // memory is staged synchronously by the runtime and serializes host
    // threads that drive other streams
    if (!c.h_ctrl) OPH_CHECK(hipHostMalloc((void **)&c.h_ctrl, sizeof(Ctrl), hipHostMallocDefault));
    if (!c.h_word) OPH_CHECK(hipHostMalloc((void **)&c.h_word, 4 * sizeof(unsigned long long),
                                           hipHostMallocDefault));
    Ctrl &h = *c.h_ctrl;
    OPH_CHECK(hipMemcpyAsync(&h, d.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s));
    OPH_CHECK(hipStreamSynchronize(s));
    if ((rc = err_from_bits(h.err))) return rc;
    launch_slots(d, s);
    // Phase II lanes per column: short input columns (simplicial complexes
    // of small dimension) take 8 lanes with 128-entry LDS buffers, longer
    // ones (cubical complexes) a full wave with 1024-entry buffers.
    int lanes = cfg.lanes, wlds = cfg.wlds;
    if (!lanes) {
        lanes = h.maxlen <= 4 ? 8 : 64;
        wlds = h.maxlen <= 4 ? 128 : 1024;
    }
    const int grid = lanes > 1 ? phase2_blocks_per_cu(lanes, wlds) * c.num_cu : 0;
    OPH_CHECK(hipEventRecord(c.ev0, s));
    launch_phase0(d, s);

    int ncompact = 0, noverflow = 0;
    unsigned long long peak = nnz;
    unsigned long long prev_add = 0;
    int prev_marks = 0;
    int stalls = 0;
    int iter = 1;
    bool converged = false;
    if (time_track && c.iter_events.size() < 2) {
        // grown lazily below
    }
    std::vector<hipEvent_t> &evs = c.iter_events;
    if (time_track) {
        if (evs.empty()) {
            hipEvent_t e;
            OPH_CHECK(hipEventCreate(&e));
            evs.push_back(e);
        }
        OPH_CHECK(hipEventRecord(evs[0], s));
    }
    while (true) {
        // Phase I and the worklist. When no low repeats (after the first
        // iteration) the matrix is reduced; the marks and clears made in
        // that pass are then facts about the reduced matrix and change
        // nothing.
        launch_reset_flags(d, s);
        launch_phase1_detect(d, s);
        launch_phase1_mark(d, s);
        launch_apply1(d, s);
        OPH_CHECK(hipMemcpyAsync(&h, d.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s));
        OPH_CHECK(hipStreamSynchronize(s));
        if (h.err) return err_from_bits(h.err) ? err_from_bits(h.err) : OPENPH_EINVAL;
        peak = std::max(peak, h.pool_top);
        if (iter > 1 && !h.any_dup) {
            converged = true;
            break;
        }
        if (iter > cfg.max_iters) return OPENPH_EITER;
        // Phase II grids sized by the work, so that a small pass does not
        // hold LDS on every CU while other streams' work waits for it.
        const int nwork = h.nwork;
        int g1 = grid, g2 = 0;
        if (lanes > 1) {
            const int per_block = kBlock / lanes;
            g1 = std::min(grid, (nwork + per_block - 1) / per_block);
            g2 = std::min(c.num_cu, nwork);
        }
        if (stats) {
            size_t k = 2 * (size_t)(iter - 1);
            while (c.p2_events.size() < k + 2) {
                hipEvent_t e;
                OPH_CHECK(hipEventCreate(&e));
                c.p2_events.push_back(e);
            }
            OPH_CHECK(hipEventRecord(c.p2_events[k], s));
            if (nwork > 0 || lanes == 1) launch_phase2(d, lanes, wlds, g1, g2, s);
            OPH_CHECK(hipEventRecord(c.p2_events[k + 1], s));
        } else if (nwork > 0 || lanes == 1) {
            launch_phase2(d, lanes, wlds, g1, g2, s);
        }
        launch_apply2(d, s);
        if (time_track && (size_t)iter <= (size_t)m) {
            while (evs.size() <= (size_t)iter) {
                hipEvent_t e;
                OPH_CHECK(hipEventCreate(&e));
                evs.push_back(e);
            }
            OPH_CHECK(hipEventRecord(evs[iter], s));
        }
        OPH_CHECK(hipMemcpyAsync(&h, d.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s));
        OPH_CHECK(hipStreamSynchronize(s));
        if (h.err) return OPENPH_EINVAL;
        peak = std::max(peak, h.pool_top);
        // pms makes progress every iteration while lows repeat (the first
        // occurrence of the largest repeated low is marked, or its repeats
        // are reduced); no progress without an overflow means a fault.
        bool progress = h.additions != prev_add || h.marks != prev_marks;
        if (!h.overflow && !progress && h.any_dup) return OPENPH_EITER;
        // a full pool with no progress twice in a row (the second time right
        // after a compaction) means the pool budget cannot hold the work
        stalls = (h.overflow && !progress) ? stalls + 1 : 0;
        if (stalls >= 2) return OPENPH_ENOMEM;
        prev_add = h.additions;
        prev_marks = h.marks;
        if (h.overflow) noverflow++;
        if (h.overflow || h.pool_top * 4 > (unsigned long long)lcap * 3) {
            rc = compact(c, cfg, s, &ncompact, lcap);
            if (rc) return rc;
        }
        iter++;
    }
    (void)converged;
    launch_final(d, low1, ess, s);
    OPH_CHECK(hipEventRecord(c.ev1, s));
    OPH_CHECK(hipMemcpyAsync(&h, d.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s));
    OPH_CHECK(hipStreamSynchronize(s));
    if (stats) {
        stats->num_iters = iter;
        stats->compactions = ncompact;
        stats->overflows = noverflow;
        stats->additions = (int64_t)h.additions;
        stats->merge_elems = (int64_t)h.merge_elems;
        stats->pool_peak = (int64_t)std::max(peak, h.pool_top);
        float ms = 0.f;
        OPH_CHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
        stats->device_ms = ms;
        stats->m = m;
        stats->lanes = lanes;
        stats->group_lds = wlds;
        stats->spills = h.spills;
        stats->handoffs = h.handoffs;
        stats->phase2_launches = iter - 1;
        float p2 = 0.f;
        for (int i = 0; i < iter - 1; i++) {
            float t = 0.f;
            OPH_CHECK(hipEventElapsedTime(&t, c.p2_events[2 * i], c.p2_events[2 * i + 1]));
            p2 += t;
        }
        stats->phase2_ms = p2;
    }
    if (time_track) {
        time_track[0] = 0.f;
        for (int i = 1; i < iter && (size_t)i < evs.size(); i++) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, evs[i - 1], evs[i]) == hipSuccess) time_track[i] = ms;
        }
    }
    return OPENPH_OK;
}

static bool known_algo(const char *a) {
    static const char *names[] = {"pms", "standard", "twist", "ph_row",
                                  "standard_parallel", "twist_parallel", "ph_row_parallel"};
    if (!a) return false;
    for (const char *n : names)
        if (std::strcmp(a, n) == 0) return true;
    return false;
}

int set_device(int device) {
    int n = 0;
    OPH_CHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return OPENPH_EINVAL;
    OPH_CHECK(hipSetDevice(device));
    return OPENPH_OK;
}

// Host-pointer run: stage through the context's device buffers.
static int run_host(const char *algstr, const int *rows1, const int *cols1, int m, int nnz,
                    const openph_options *opt, int *low1, int *ess, openph_stats *stats,
                    float *time_track) {
    if (!known_algo(algstr)) return OPENPH_EALGO;
    if (m < 0 || nnz < 0 || (nnz > 0 && (!rows1 || !cols1)) || (m > 0 && (!low1 || !ess)))
        return OPENPH_EINVAL;
    int dev = opt ? opt->device : 0;
    int rc = set_device(dev);
    if (rc) return rc;
    Context *c = context_for(dev);
    std::lock_guard<std::mutex> g(c->mu);
    if (!c->stream) OPH_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    hipStream_t s = c->stream;
    OPH_CHECK(grow(&c->h_rows, c->c_hrows, (size_t)nnz + 1));
    OPH_CHECK(grow(&c->h_cols, c->c_hcols, (size_t)nnz + 1));
    OPH_CHECK(grow(&c->h_low, c->c_hlow, (size_t)m + 1));
    OPH_CHECK(grow(&c->h_ess, c->c_hess, (size_t)m + 1));
    if (nnz) {
        OPH_CHECK(hipMemcpyAsync(c->h_rows, rows1, sizeof(int) * nnz, hipMemcpyHostToDevice, s));
        OPH_CHECK(hipMemcpyAsync(c->h_cols, cols1, sizeof(int) * nnz, hipMemcpyHostToDevice, s));
    }
    RunCfg cfg = cfg_from(opt, nnz);
    rc = run_pms(*c, c->h_rows, c->h_cols, m, nnz, c->h_low, c->h_ess, s, cfg, stats, time_track);
    if (rc) {
        (void)hipStreamSynchronize(s);
        return rc;
    }
    if (m) {
        OPH_CHECK(hipMemcpyAsync(low1, c->h_low, sizeof(int) * m, hipMemcpyDeviceToHost, s));
        OPH_CHECK(hipMemcpyAsync(ess, c->h_ess, sizeof(int) * m, hipMemcpyDeviceToHost, s));
    }
    OPH_CHECK(hipStreamSynchronize(s));
    return OPENPH_OK;
}

// Batch of `count` matrices (device COO, local 1-based indices) reduced as
// one block-diagonal matrix. low1/ess (device) receive the concatenated
// per-matrix outputs with local 1-based lows.
int run_batch(Context &c, int count, const int *ms, const int *nnzs, const int *rows1,
                     const int *cols1, int *low1, int *ess, hipStream_t s, const RunCfg &cfg,
                     openph_stats *stats) {
    if (count < 0 || (count > 0 && (!ms || !nnzs))) return OPENPH_EINVAL;
    std::vector<int> moff((size_t)count + 1), noff((size_t)count + 1);
    long long M = 0, N = 0;
    for (int k = 0; k < count; k++) {
        if (ms[k] < 0 || nnzs[k] < 0) return OPENPH_EINVAL;
        moff[k] = (int)M;
        noff[k] = (int)N;
        M += ms[k];
        N += nnzs[k];
        if (M > 0x7fffffffLL || N > 0x7fffffffLL) return OPENPH_EINVAL;
    }
    moff[count] = (int)M;
    noff[count] = (int)N;
    if (N > 0 && (!rows1 || !cols1)) return OPENPH_EINVAL;
    if (M > 0 && (!low1 || !ess)) return OPENPH_EINVAL;
    OPH_CHECK(grow(&c.b_moff, c.c_bmoff, (size_t)count + 1));
    OPH_CHECK(grow(&c.b_noff, c.c_bnoff, (size_t)count + 1));
    OPH_CHECK(grow(&c.b_rows, c.c_brows, (size_t)N + 1));
    OPH_CHECK(grow(&c.b_cols, c.c_bcols, (size_t)N + 1));
    OPH_CHECK(grow(&c.b_err, c.c_berr, 1));
    int zero = 0, err = 0;
    OPH_CHECK(hipMemcpyAsync(c.b_moff, moff.data(), sizeof(int) * (count + 1),
                             hipMemcpyHostToDevice, s));
    OPH_CHECK(hipMemcpyAsync(c.b_noff, noff.data(), sizeof(int) * (count + 1),
                             hipMemcpyHostToDevice, s));
    OPH_CHECK(hipMemcpyAsync(c.b_err, &zero, sizeof(int), hipMemcpyHostToDevice, s));
    launch_batch_offsets(rows1, cols1, c.b_noff, c.b_moff, count, (int)N, c.b_rows, c.b_cols,
                         c.b_err, s);
    OPH_CHECK(hipMemcpyAsync(&err, c.b_err, sizeof(int), hipMemcpyDeviceToHost, s));
    OPH_CHECK(hipStreamSynchronize(s));
    if (err) return OPENPH_EINVAL;
    int rc = run_pms(c, c.b_rows, c.b_cols, (int)M, (int)N, low1, ess, s, cfg, stats, nullptr);
    if (rc) return rc;
    launch_batch_local_low(low1, c.b_moff, count, (int)M, s);
    OPH_CHECK(hipStreamSynchronize(s));
    return OPENPH_OK;
}

static int batch_host(int count, const int *ms, const int *nnzs, const int *rows1,
                      const int *cols1, const openph_options *opt, int *low1, int *ess,
                      openph_stats *stats) {
    if (count < 0 || (count > 0 && (!ms || !nnzs))) return OPENPH_EINVAL;
    long long M = 0, N = 0;
    for (int k = 0; k < count; k++) {
        if (ms[k] < 0 || nnzs[k] < 0) return OPENPH_EINVAL;
        M += ms[k];
        N += nnzs[k];
    }
    if (M > 0x7fffffffLL || N > 0x7fffffffLL) return OPENPH_EINVAL;
    if ((N > 0 && (!rows1 || !cols1)) || (M > 0 && (!low1 || !ess))) return OPENPH_EINVAL;
    int dev = opt ? opt->device : 0;
    int rc = set_device(dev);
    if (rc) return rc;
    Context *c = context_for(dev);
    std::lock_guard<std::mutex> g(c->mu);
    if (!c->stream) OPH_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    hipStream_t s = c->stream;
    OPH_CHECK(grow(&c->h_rows, c->c_hrows, (size_t)N + 1));
    OPH_CHECK(grow(&c->h_cols, c->c_hcols, (size_t)N + 1));
    OPH_CHECK(grow(&c->h_low, c->c_hlow, (size_t)M + 1));
    OPH_CHECK(grow(&c->h_ess, c->c_hess, (size_t)M + 1));
    if (N) {
        OPH_CHECK(hipMemcpyAsync(c->h_rows, rows1, sizeof(int) * N, hipMemcpyHostToDevice, s));
        OPH_CHECK(hipMemcpyAsync(c->h_cols, cols1, sizeof(int) * N, hipMemcpyHostToDevice, s));
    }
    RunCfg cfg = cfg_from(opt, (int)N);
    rc = run_batch(*c, count, ms, nnzs, c->h_rows, c->h_cols, c->h_low, c->h_ess, s, cfg, stats);
    if (rc) {
        (void)hipStreamSynchronize(s);
        return rc;
    }
    if (M) {
        OPH_CHECK(hipMemcpyAsync(low1, c->h_low, sizeof(int) * M, hipMemcpyDeviceToHost, s));
        OPH_CHECK(hipMemcpyAsync(ess, c->h_ess, sizeof(int) * M, hipMemcpyDeviceToHost, s));
    }
    OPH_CHECK(hipStreamSynchronize(s));
    return OPENPH_OK;
}

}  // namespace oph

using namespace oph;

extern "C" {

const char *openph_version(void) { return "openph-mi355x 0.1.0"; }

const char *openph_strerror(int status) {
    switch (status) {
        case OPENPH_OK: return "ok";
        case OPENPH_EINVAL: return "invalid argument or malformed boundary matrix";
        case OPENPH_EUNSORTED: return "COO input is not column-major with ascending rows";
        case OPENPH_EALGO: return "unknown algorithm string";
        case OPENPH_ENOMEM: return "out of device memory or column-pool budget";
        case OPENPH_EDEVICE: return "HIP runtime error";
        case OPENPH_EITER: return "iteration cap reached";
        case OPENPH_EDIMS: return "too many distinct pseudo-dimensions";
        case OPENPH_EBUSY: return "admission refused";
        default: return "unknown status";
    }
}

int openph_v2(const char *algstr, const int *rows1, const int *cols1, int m, int nnz,
              const int *low_true1, const openph_options *opt, int *low1, int *ess,
              float *trackers, openph_stats *stats) {
    (void)low_true1;
    float *tt = (trackers && opt && opt->track) ? trackers + 4 * ((size_t)m + 1) : nullptr;
    return run_host(algstr, rows1, cols1, m, nnz, opt, low1, ess, stats, tt);
}

int openph_pms_device(const int *rows1, const int *cols1, int m, int nnz, int *low1, int *ess,
                      void *stream, const openph_options *opt, openph_stats *stats) {
    if (m < 0 || nnz < 0) return OPENPH_EINVAL;
    int dev = opt ? opt->device : 0;
    int rc = set_device(dev);
    if (rc) return rc;
    Context *c = context_for(dev);
    std::lock_guard<std::mutex> g(c->mu);
    RunCfg cfg = cfg_from(opt, nnz);
    return run_pms(*c, rows1, cols1, m, nnz, low1, ess, (hipStream_t)stream, cfg, stats, nullptr);
}

void openph(char *algstr, int *h_rows_in, int *h_cols_in, int m, int col_width, int *h_low_true,
            int nnz, int *h_low, int *h_ess, float *err_linf, float *err_lone, float *err_redu,
            float *err_ess, float *time_track, int *num_iters, int gpu_number) {
    (void)col_width;
    (void)h_low_true;
    (void)err_linf;
    (void)err_lone;
    (void)err_redu;
    (void)err_ess;
    openph_options opt;
    std::memset(&opt, 0, sizeof(opt));
    opt.device = gpu_number;
    openph_stats st;
    std::memset(&st, 0, sizeof(st));
    int rc = run_host(algstr, h_rows_in, h_cols_in, m, nnz, &opt, h_low, h_ess, &st, time_track);
    if (num_iters) num_iters[0] = rc ? rc : st.num_iters;
}

int openph_pms_batch(int count, const int *ms, const int *nnzs, const int *rows1,
                     const int *cols1, const openph_options *opt, int *low1, int *ess,
                     openph_stats *stats) {
    return batch_host(count, ms, nnzs, rows1, cols1, opt, low1, ess, stats);
}

int openph_pms_batch_device(int count, const int *ms, const int *nnzs, const int *rows1,
                            const int *cols1, int *low1, int *ess, void *stream,
                            const openph_options *opt, openph_stats *stats) {
    int dev = opt ? opt->device : 0;
    int rc = set_device(dev);
    if (rc) return rc;
    Context *c = context_for(dev);
    std::lock_guard<std::mutex> g(c->mu);
    long long N = 0;
    for (int k = 0; k < count && nnzs; k++) N += nnzs[k] > 0 ? nnzs[k] : 0;
    RunCfg cfg = cfg_from(opt, (int)std::min<long long>(N, 0x7fffffffLL));
    return run_batch(*c, count, ms, nnzs, rows1, cols1, low1, ess, (hipStream_t)stream, cfg,
                     stats);
}

}  // extern "C"
