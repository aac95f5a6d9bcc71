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

Context *context_for(int device) {
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
        r.schedule = opt->schedule;
        r.chunk = opt->chunk_iters;
        r.timing = opt->time_phase2 != 0;
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
    // the room a compaction leaves free at the least: a small pool grows to
    // it (within the budget) instead of failing
    const size_t need = live + live / 16 + 1024;
    if (want < need) want = need;
    if (want > cfg.pool_max) want = cfg.pool_max;
    if (need > want) return OPENPH_ENOMEM;
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

// Reads the control block into pinned memory and waits for it.
static int read_ctrl(Context &c, hipStream_t s) {
    OPH_CHECK(hipMemcpyAsync(c.h_ctrl, c.d.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s));
    OPH_CHECK(hipStreamSynchronize(s));
    return OPENPH_OK;
}

// A Phase II pass cut short by a full pool (kHaltSuspend): the pool is
// compacted, the same pass runs again over the same worklist (each column
// goes on from the state it committed, against the pivots of the
// iteration's start), and k_apply2 then ends the iteration, so that the
// iteration does what it would have done with a larger pool.
static int resume_suspended(Context &c, const RunCfg &cfg, int lanes, int wlds, int grid,
                            int long_grid, hipStream_t s, int *ncompact, size_t &lcap) {
    int rc = compact(c, cfg, s, ncompact, lcap);
    if (rc) return rc;
    Dev &d = c.d;
    launch_resume2(d, s);
    launch_phase2(d, lanes, wlds, grid, long_grid, s);
    launch_apply2(d, s);
    // the finished iteration's end clock and trackers
    if (d.its) launch_iter_end_stamp(d, s, 1);
    if (d.ltrue) launch_track_iter(d, s);
    return OPENPH_OK;
}

static int halt_status(const Ctrl &h) {
    switch (h.halt) {
        case kHaltStuck: return OPENPH_EITER;
        case kHaltIters: return OPENPH_EITER;
        case kHaltNoMem: return OPENPH_ENOMEM;
        case kHaltSuspend: return OPENPH_EITER;  // always resumed before this is asked
        case kHaltError: {
            int rc = err_from_bits(h.err);
            return rc ? rc : OPENPH_EINVAL;
        }
        default: return OPENPH_OK;
    }
}

static int grow_events(std::vector<hipEvent_t> &v, size_t n) {
    while (v.size() < n) {
        hipEvent_t e;
        OPH_CHECK(hipEventCreate(&e));
        v.push_back(e);
    }
    return OPENPH_OK;
}

// The Phase II launch shape of one reduction.
struct Shape {
    int lanes, wlds, grid, long_grid;
};

// One loop iteration as queued on a stream (pms.cu:44-114): Phase I and the
// worklist, Phase II on fixed persistent grids (the kernels walk the
// worklist, whatever its length), the bookkeeping, and the device-side
// checks that stop the remaining queued iterations (Ctrl::halt). Phase II
// is bracketed by the events e0/e1 or, inside graphs, by device-clock
// stamps, when either is given.
static int enqueue_iteration(const Dev &d, const Shape &sh, hipStream_t s, hipEvent_t e0,
                             hipEvent_t e1, unsigned long long *stamps) {
    launch_phase1_detect(d, s);  // k_first starts the iteration
    launch_phase1_mark(d, s);
    launch_apply1(d, s);
    if (e0) OPH_CHECK(hipEventRecord(e0, s));
    if (stamps) launch_stamp(d, stamps, 0, kStampRing, s);
    launch_phase2(d, sh.lanes, sh.wlds, sh.grid, sh.long_grid, s);
    if (stamps) launch_stamp(d, stamps, 1, kStampRing, s);
    if (e1) OPH_CHECK(hipEventRecord(e1, s));
    launch_apply2(d, s);  // ends the iteration: Ctrl::halt for the next
    if (d.its) launch_iter_end_stamp(d, s);
    if (d.ltrue) launch_track_iter(d, s);
    return OPENPH_OK;
}

// OPENPH_DEBUG=1 in the environment reports recoverable runtime refusals
// (a graph capture that falls back to queued launches) and the Phase II
// time of every iteration of the sync schedule on stderr.
static bool debug_on() {
    static const bool on = std::getenv("OPENPH_DEBUG") != nullptr;
    return on;
}

static void debug_note(const char *what, hipError_t e) {
    if (debug_on()) std::fprintf(stderr, "openph: %s: %s\n", what, hipGetErrorString(e));
}

static void drop_graph(GraphEntry &g) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g = GraphEntry();
}

// What a whole-reduction graph adds around its iterations: the set-up of
// rows1/cols1 (ctrl init, COO -> pool, validation guard, slots, Phase 0)
// and the outputs into low1/ess, gated on convergence.
struct WholeArgs {
    const int *rows1, *cols1;
    int *low1, *ess;
    size_t lcap, pool_elems;
    int max_iters;
};

// The set-up of a reduction whose Phase II shape is known without waiting
// for it: a malformed input is made inert by the guard and reported by the
// first read of the control block.
static void enqueue_setup(const Dev &d, const WholeArgs &w, hipStream_t s) {
    launch_ctrl_init(d, (unsigned long long)w.lcap, w.max_iters, s);
    launch_setup(d, w.rows1, w.cols1, s);
    if (w.pool_elems) launch_pool_cap(d, w.pool_elems, w.lcap, s);
    launch_settle(d, s);
    launch_phase0(d, s);
    if (d.fuse_first) launch_first(d, s);
}

// The cached graph of `iters` iterations for the current buffers and shape
// (with `whole`, of the whole reduction around them), captured on the
// context's own stream (the caller's may be the null stream, which cannot
// be captured) when there is none. A timed graph brackets Phase II with
// device-clock stamps: event records inside a capture are refused by this
// runtime once PyTorch has initialised it.
static int graph_for(Context &c, const Shape &sh, int iters, bool timed, const WholeArgs *whole,
                     GraphEntry **out) {
    const Dev &d = c.d;
    GraphEntry *hit = nullptr, *victim = &c.graphs[0];
    for (GraphEntry &g : c.graphs) {
        const bool same_whole =
            whole ? g.whole && g.rows1 == whole->rows1 && g.cols1 == whole->cols1 &&
                        g.low1 == whole->low1 && g.ess == whole->ess && g.lcap == whole->lcap &&
                        g.pool_elems == whole->pool_elems && g.max_iters == whole->max_iters
                  : !g.whole;
        if (g.valid && std::memcmp(&g.key, &d, sizeof(Dev)) == 0 && g.lanes == sh.lanes &&
            g.wlds == sh.wlds && g.grid == sh.grid && g.long_grid == sh.long_grid &&
            g.iters == iters && g.timed == timed && same_whole) {
            hit = &g;
            break;
        }
        if (!g.valid || g.used < victim->used) victim = &g;
    }
    if (!hit) {
        drop_graph(*victim);
        if (!c.stream) OPH_CHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        if (timed && !c.d_stamps) {
            OPH_CHECK(hipMalloc((void **)&c.d_stamps, kStampRing * sizeof(unsigned long long)));
            OPH_CHECK(hipHostMalloc((void **)&c.h_stamps, kStampRing * sizeof(unsigned long long),
                                    hipHostMallocDefault));
            OPH_CHECK(hipDeviceGetAttribute(&c.wall_khz, hipDeviceAttributeWallClockRate,
                                            c.device));
        }
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamBeginCapture(c.stream, hipStreamCaptureModeRelaxed);
        if (e != hipSuccess) {
            debug_note("hipStreamBeginCapture", e);
            last_hip_error() = e;
            return OPENPH_EDEVICE;
        }
        int rc = OPENPH_OK;
        if (whole) enqueue_setup(d, *whole, c.stream);
        for (int k = 0; k < iters && !rc; k++)
            rc = enqueue_iteration(d, sh, c.stream, nullptr, nullptr,
                                   timed ? c.d_stamps : nullptr);
        if (whole && !rc) launch_final(d, whole->low1, whole->ess, c.stream, true);
        if (rc) debug_note("capture of one iteration", last_hip_error());
        e = hipStreamEndCapture(c.stream, &graph);
        if (rc || e != hipSuccess) {
            if (e != hipSuccess) debug_note("hipStreamEndCapture", e);
            if (graph) (void)hipGraphDestroy(graph);
            if (e != hipSuccess) last_hip_error() = e;
            return rc ? rc : OPENPH_EDEVICE;
        }
        hipGraphExec_t exec = nullptr;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            debug_note("hipGraphInstantiate", e);
            (void)hipGraphDestroy(graph);
            last_hip_error() = e;
            return OPENPH_EDEVICE;
        }
        victim->valid = true;
        victim->key = d;
        victim->lanes = sh.lanes;
        victim->wlds = sh.wlds;
        victim->grid = sh.grid;
        victim->long_grid = sh.long_grid;
        victim->iters = iters;
        victim->timed = timed;
        victim->whole = whole != nullptr;
        if (whole) {
            victim->rows1 = whole->rows1;
            victim->cols1 = whole->cols1;
            victim->low1 = whole->low1;
            victim->ess = whole->ess;
            victim->lcap = whole->lcap;
            victim->pool_elems = whole->pool_elems;
            victim->max_iters = whole->max_iters;
        }
        victim->graph = graph;
        victim->exec = exec;
        hit = victim;
    }
    hit->used = ++c.graph_clock;
    *out = hit;
    return OPENPH_OK;
}

// Tracker slot `it` (tracker.cu:92-134) from the four exact integers of one
// state: sum |low - low_true|, max |low - low_true|, #unreduced, #essential.
static void fill_track(const Track *tr, int it, int m, long long ess_true, unsigned long long l1,
                       unsigned long long linf, unsigned long long unreduced,
                       unsigned long long ess) {
    if (!tr || !tr->low_true1 || it < 0 || it > m) return;
    if (tr->err_lone) tr->err_lone[it] = (float)l1;
    if (tr->err_linf) tr->err_linf[it] = (float)linf;
    if (tr->err_redu) tr->err_redu[it] = (float)unreduced / (float)m;
    if (tr->err_ess)
        tr->err_ess[it] =
            (long long)ess == ess_true ? 1.0f : (float)ess_true / (float)(long long)ess;
}

// Trackers of the current state into slot `it` (tracker.cu:92-134); reads
// the control block in any case.
static int track_state(Context &c, const Track *tr, int it, hipStream_t s, int m,
                       long long ess_true) {
    const bool slot = tr && tr->low_true1 && it <= m;
    if (slot) launch_track(c.d, c.t_low, s);
    int rc = read_ctrl(c, s);
    if (rc || !slot) return rc;
    const Ctrl &h = *c.h_ctrl;
    fill_track(tr, it, m, ess_true, h.trk_l1, (unsigned long long)h.trk_linf,
               (unsigned long long)h.trk_unreduced, (unsigned long long)h.trk_ess);
    return OPENPH_OK;
}

// The pms driver (pms.cu:1-118) on device-resident 1-based COO. With
// cfg.no_alloc it allocates nothing: the caller's context holds every buffer
// (ensure() then only checks sizes) and host mirrors, no events are created
// (stats carry no times) and the iterations are queued, never captured.
int run_pms(Context &c, const int *rows1, const int *cols1, int m, int nnz, int *low1,
                   int *ess, hipStream_t s, const RunCfg &cfg, openph_stats *stats,
                   const Track *tr) {
    if (m < 0 || nnz < 0) return OPENPH_EINVAL;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (m == 0) {
        if (stats) stats->num_iters = 1;
        return OPENPH_OK;
    }
    if (cfg.lanes < 0 || cfg.schedule < kSchedAuto || cfg.schedule > kSchedGraph || cfg.chunk < 0)
        return OPENPH_EINVAL;
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
    d.nmem = 0;  // not a running batch of the server
    d.mdup = nullptr;
    d.mstart = nullptr;
    if (!c.num_cu) {
        hipDeviceProp_t prop;
        OPH_CHECK(hipGetDeviceProperties(&prop, c.device));
        c.num_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    }
    const bool events = !cfg.no_alloc;
    if (cfg.no_alloc && (!c.h_ctrl || !c.h_word || (tr && (tr->low_true1 || tr->time_track))))
        return OPENPH_EINVAL;
    if (events && !c.ev0) {
        OPH_CHECK(hipEventCreate(&c.ev0));
        OPH_CHECK(hipEventCreate(&c.ev1));
    }
    // the control block is read back into pinned memory: a copy to pageable
    // memory is staged synchronously by the runtime and serializes host
    // threads that drive other streams
    if (!c.h_ctrl) OPH_CHECK(hipHostMalloc((void **)&c.h_ctrl, sizeof(Ctrl), hipHostMallocDefault));
    if (!c.h_word) OPH_CHECK(hipHostMalloc((void **)&c.h_word, 4 * sizeof(unsigned long long),
                                           hipHostMallocDefault));
    // trackers: low_true (1-based, 0 = none) checked, shifted down and
    // uploaded; its essential count (compute_ess_true, kernels.cu:54-62) and
    // norms (openph.cu:106-111) are taken here
    const bool tracking = tr && tr->low_true1;
    long long ess_true = 0;
    float norm1 = 0.f, norminf = 0.f;
    if (tracking) {
        std::vector<int> lt0((size_t)m);
        std::vector<char> paired((size_t)m, 0);
        long long l1 = 0;
        int linf = 0;
        for (int j = 0; j < m; j++) {
            const int v = tr->low_true1[j];
            if (v < 0 || v > m) return OPENPH_EINVAL;
            lt0[j] = v - 1;
            l1 += v;
            linf = std::max(linf, v);
            if (v > 0) {
                paired[j] = 1;
                paired[v - 1] = 1;
            }
        }
        for (int j = 0; j < m; j++) ess_true += !paired[j];
        norm1 = (float)l1;
        norminf = (float)linf;
        OPH_CHECK(grow(&c.t_low, c.c_tlow, (size_t)m + 1));
        OPH_CHECK(hipMemcpyAsync(c.t_low, lt0.data(), sizeof(int) * m, hipMemcpyHostToDevice, s));
    }
    // logical pool size: the requested one, even when the context holds a
    // larger buffer from an earlier call
    size_t lcap = std::min(c.cap_pool, pool_want);
    // the schedule: graph replay when the last reduction on this workspace
    // had the same size (its graphs are likely cached: a capture costs more
    // than it saves once), queued chunks otherwise. The trackers and the
    // iteration times (time_track) of the queued and graph schedules are
    // kept per iteration on the device and read once per chunk; the sync
    // schedule reads them after every iteration.
    const bool same_size = c.last_m == m && c.last_nnz == nnz;
    int sched = cfg.schedule;
    if (sched == kSchedAuto) sched = same_size ? kSchedGraph : kSchedQueued;
    if (cfg.no_alloc && sched == kSchedGraph) sched = kSchedQueued;
    // Without trackers the chunked schedules read the device once per
    // chunk and no more: the set-up is not waited for when the Phase II
    // shape is known without it (given, or the one this size took last
    // time: any shape gives the same result), a guard kernel making
    // malformed input inert instead; and the outputs are queued behind each
    // chunk, gated on convergence, so the chunk's read also finds them done.
    const bool one_read = !tracking && sched != kSchedSync;
    const bool ahead = one_read && (cfg.lanes != 0 || (same_size && c.last_lanes != 0));
    const bool iter_times = tr && tr->time_track;
    // no host step between two iterations (no trackers, no per-iteration or
    // per-pass times): k_apply2 also starts the next one
    d.fuse_first = one_read && !iter_times && !cfg.timing ? 1 : 0;
    // A repeated reduction replays one graph from the set-up to the gated
    // outputs: launched one by one, the set-up's kernels leave the device
    // waiting on the host between them.
    const bool whole = sched == kSchedGraph && ahead && !iter_times;
    const WholeArgs wargs{rows1, cols1, low1, ess, lcap, cfg.pool_elems, cfg.max_iters};
    if (events) OPH_CHECK(hipEventRecord(c.ev0, s));
    if (!whole) {
        launch_ctrl_init(d, (unsigned long long)lcap, cfg.max_iters, s);
        launch_setup(d, rows1, cols1, s);
        if (cfg.pool_elems) launch_pool_cap(d, cfg.pool_elems, lcap, s);
    }
    int maxlen = 0;
    if (!ahead) {
        rc = read_ctrl(c, s);  // also completes the low_true upload from pageable memory
        if (rc) return rc;
        if ((rc = err_from_bits(c.h_ctrl->err))) return rc;
        lcap = c.h_ctrl->pool_cap;
        maxlen = c.h_ctrl->maxlen;
    }
    const bool ring = sched != kSchedSync && (tracking || iter_times);
    d.ltrue = nullptr;
    d.trk_part = d.trk_ring = d.its = nullptr;
    if (ring) {
        if (!c.h_ring) {
            OPH_CHECK(hipHostMalloc((void **)&c.h_ring, 4 * kIterRing * sizeof(unsigned long long),
                                    hipHostMallocDefault));
            OPH_CHECK(hipHostMalloc((void **)&c.h_its, 2 * kIterRing * sizeof(unsigned long long),
                                    hipHostMallocDefault));
        }
        if (!c.wall_khz)
            OPH_CHECK(hipDeviceGetAttribute(&c.wall_khz, hipDeviceAttributeWallClockRate,
                                            c.device));
        if (tracking) {
            OPH_CHECK(grow(&c.t_part, c.c_tpart, 4 * (size_t)track_part_blocks(m)));
            OPH_CHECK(grow(&c.t_ring, c.c_tring, 4 * (size_t)kIterRing));
            d.ltrue = c.t_low;
            d.trk_part = c.t_part;
            d.trk_ring = c.t_ring;
        }
        if (iter_times) {
            OPH_CHECK(grow(&c.t_its, c.c_tits, 2 * (size_t)kIterRing));
            d.its = c.t_its;
        }
    }
    if (tracking) {
        if (ring) launch_track_iter(d, s);  // the state before Phase 0 (iter 0)
        else if ((rc = track_state(c, tr, 0, s, m, ess_true))) return rc;
    }
    // with `ahead` the error, and the pool's size, come with the first
    // read; k_settle makes malformed input inert meanwhile
    if (!whole) launch_settle(d, s);
    // Phase II lanes per column: short input columns (simplicial complexes
    // of small dimension) take 8 lanes with 80-entry LDS buffers (a 16-entry
    // pivot buffer: 22 KB per block, so seven blocks share a CU's 160 KB),
    // longer ones (cubical complexes) a full wave with 1024-entry buffers.
    Shape sh{cfg.lanes, cfg.wlds, 0, 0};
    if (!sh.lanes && ahead) {
        sh.lanes = c.last_lanes;
        sh.wlds = c.last_wlds;
    } else if (!sh.lanes) {
        sh.lanes = maxlen <= 4 ? 8 : 64;
        sh.wlds = maxlen <= 4 ? 80 : 1024;
        c.last_lanes = sh.lanes;
        c.last_wlds = sh.wlds;
    }
    if (sh.lanes > 1) {
        // twice the resident blocks: a block whose columns are done makes
        // room for one that has not started, which evens out the load
        sh.grid = 2 * phase2_blocks_per_cu(sh.lanes, sh.wlds) * c.num_cu;
        sh.long_grid = c.num_cu;
    }
    // long columns (the 1024-entry shapes): iterations whose worklist fits
    // the GPU's 256-lane teams at once (three per CU) use them
    d.wide_max = sh.wlds >= 1024 && cfg.wide ? 3 * c.num_cu : 0;
    const bool timed = stats != nullptr && events && cfg.timing;

    if (!whole) launch_phase0(d, s);
    if (!whole && d.fuse_first) launch_first(d, s);
    // iteration i is timed from evs[2i] to evs[2i+1] (tic/toc around the
    // loop body, pms.cu:46-103; the trackers are outside)
    std::vector<hipEvent_t> &evs = c.iter_events;
    int ncompact = 0, reads = 0, p2_launches = 0;
    float p2_ms = 0.f;
    const Ctrl &h = *c.h_ctrl;
    if (sched == kSchedSync) {
        while (true) {
            const int next = h.iter + 1;
            if (iter_times && next <= m) {
                if ((rc = grow_events(evs, 2 * (size_t)next + 2))) return rc;
                OPH_CHECK(hipEventRecord(evs[2 * next], s));
            }
            launch_phase1_detect(d, s);  // k_first starts the iteration
            launch_phase1_mark(d, s);
            launch_apply1(d, s);
            if ((rc = read_ctrl(c, s))) return rc;
            reads++;
            if (h.halt == kHaltConverged) break;
            if (h.halt) return halt_status(h);
            const int iter = h.iter;
            // Phase II grids sized by the work, so that a small pass does not
            // hold LDS on every CU while other streams' work waits for it
            Shape g = sh;
            const int nwork = h.nwork;
            if (sh.lanes > 1) {
                const int per_block = kBlock / sh.lanes;
                g.grid = std::min(sh.grid, (nwork + per_block - 1) / per_block);
                g.long_grid = std::min(c.num_cu, nwork);
            }
            const bool run2 = nwork > 0 || sh.lanes == 1;
            if (timed) {
                if ((rc = grow_events(c.p2_events, 2 * (size_t)iter))) return rc;
                OPH_CHECK(hipEventRecord(c.p2_events[2 * (iter - 1)], s));
                if (run2) launch_phase2(d, g.lanes, g.wlds, g.grid, g.long_grid, s);
                OPH_CHECK(hipEventRecord(c.p2_events[2 * (iter - 1) + 1], s));
            } else if (run2) {
                launch_phase2(d, g.lanes, g.wlds, g.grid, g.long_grid, s);
            }
            p2_launches++;
            launch_apply2(d, s);
            if (iter_times && iter <= m) OPH_CHECK(hipEventRecord(evs[2 * iter + 1], s));
            if (tracking) {
                if ((rc = track_state(c, tr, iter, s, m, ess_true))) return rc;
            } else if ((rc = read_ctrl(c, s))) {
                return rc;
            }
            reads++;
            if (h.halt == kHaltSuspend) {
                while (h.halt == kHaltSuspend) {
                    if ((rc = resume_suspended(c, cfg, g.lanes, g.wlds, g.grid, g.long_grid, s,
                                               &ncompact, lcap)))
                        return rc;
                    if (iter_times && iter <= m)
                        OPH_CHECK(hipEventRecord(evs[2 * iter + 1], s));
                    if ((rc = read_ctrl(c, s))) return rc;
                    reads++;
                }
                // the trackers of the finished iteration
                if (tracking && (rc = track_state(c, tr, iter, s, m, ess_true))) return rc;
            }
            if (h.halt == kHaltCompact) {
                if ((rc = compact(c, cfg, s, &ncompact, lcap))) return rc;
                launch_set_halt(d, kHaltNone, s);
            } else if (h.halt) {
                return halt_status(h);
            }
        }
        if (timed)
            for (int i = 0; i < p2_launches; i++) {
                float t = 0.f;
                OPH_CHECK(hipEventElapsedTime(&t, c.p2_events[2 * i], c.p2_events[2 * i + 1]));
                p2_ms += t;
                if (debug_on())
                    std::fprintf(stderr, "openph: m=%d iteration %d: Phase II %.3f ms\n", m, i + 1, t);
            }
    } else {
        // chunks of iterations; the first as long as the last reduction of
        // this size took, so that a repeated reduction reads the control
        // block once
        const int next_k = cfg.chunk > 0 ? cfg.chunk : 3;
        int k = (c.last_m == m && c.last_nnz == nnz && c.last_iters > 0) ? c.last_iters : 4;
        int iter0 = 0;
        while (true) {
            k = std::min(k, kIterRing / 2);
            if (d.its) launch_clear_stamps(d, s);
            GraphEntry *ge = nullptr;
            const bool first_whole = whole && iter0 == 0;
            if (sched == kSchedGraph &&
                graph_for(c, sh, k, timed, first_whole ? &wargs : nullptr, &ge) != OPENPH_OK) {
                // capture or instantiation refused: queue the same launches
                (void)hipGetLastError();
                sched = kSchedQueued;
            }
            if (first_whole && !ge) enqueue_setup(d, wargs, s);
            if (ge) {
                OPH_CHECK(hipGraphLaunch(ge->exec, s));
                if (timed)
                    OPH_CHECK(hipMemcpyAsync(c.h_stamps, c.d_stamps,
                                             kStampRing * sizeof(unsigned long long),
                                             hipMemcpyDeviceToHost, s));
            } else {
                if (timed && (rc = grow_events(c.chunk_events, 2 * (size_t)k))) return rc;
                for (int q = 0; q < k; q++)
                    if ((rc = enqueue_iteration(d, sh, s, timed ? c.chunk_events[2 * q] : nullptr,
                                                timed ? c.chunk_events[2 * q + 1] : nullptr,
                                                nullptr)))
                        return rc;
            }
            if (one_read) {
                if (!(first_whole && ge)) launch_final(d, low1, ess, s, true);
                if (events) OPH_CHECK(hipEventRecord(c.ev1, s));
            }
            if (d.trk_ring)
                OPH_CHECK(hipMemcpyAsync(c.h_ring, d.trk_ring,
                                         4 * kIterRing * sizeof(unsigned long long),
                                         hipMemcpyDeviceToHost, s));
            if (d.its)
                OPH_CHECK(hipMemcpyAsync(c.h_its, d.its, 2 * kIterRing * sizeof(unsigned long long),
                                         hipMemcpyDeviceToHost, s));
            if ((rc = read_ctrl(c, s))) return rc;
            reads++;
            if (ahead && iter0 == 0) lcap = h.pool_cap;  // set by the set-up
            while (h.halt == kHaltSuspend) {
                if ((rc = resume_suspended(c, cfg, sh.lanes, sh.wlds, sh.grid, sh.long_grid, s,
                                           &ncompact, lcap)))
                    return rc;
                if (d.trk_ring)
                    OPH_CHECK(hipMemcpyAsync(c.h_ring, d.trk_ring,
                                             4 * kIterRing * sizeof(unsigned long long),
                                             hipMemcpyDeviceToHost, s));
                if (d.its)
                    OPH_CHECK(hipMemcpyAsync(c.h_its, d.its,
                                             2 * kIterRing * sizeof(unsigned long long),
                                             hipMemcpyDeviceToHost, s));
                if ((rc = read_ctrl(c, s))) return rc;
                reads++;
            }
            // iterations of this chunk that reached Phase II: all that began,
            // but the one that found convergence (or the cap) at its start
            const int begun = h.iter - iter0;
            const int ran2 =
                begun - ((h.halt == kHaltConverged || h.halt == kHaltIters) ? 1 : 0);
            if (d.trk_ring && iter0 == 0) {
                const unsigned long long *q = c.h_ring;  // before Phase 0
                fill_track(tr, 0, m, ess_true, q[0], q[1], q[2], q[3]);
            }
            for (int q = 0; q < ran2 && q < k; q++) {
                const int i = iter0 + q + 1;
                if (i > m) break;
                const int e = i % kIterRing;
                if (d.trk_ring) {
                    const unsigned long long *x = c.h_ring + 4 * e;
                    fill_track(tr, i, m, ess_true, x[0], x[1], x[2], x[3]);
                }
                if (d.its) {
                    const unsigned long long a = c.h_its[2 * e], b = c.h_its[2 * e + 1];
                    tr->time_track[i] =
                        c.wall_khz > 0 && b >= a ? (float)((double)(b - a) / c.wall_khz) : 0.f;
                }
            }
            if (timed && ge) {
                for (int q = 0; q < ran2 && q < k; q++) {
                    const int i = iter0 + q + 1;  // the iteration's stamps
                    const unsigned long long a = c.h_stamps[(2 * (i - 1)) % kStampRing];
                    const unsigned long long b = c.h_stamps[(2 * (i - 1) + 1) % kStampRing];
                    if (c.wall_khz > 0 && b >= a) p2_ms += (float)((double)(b - a) / c.wall_khz);
                }
            } else if (timed) {
                for (int q = 0; q < ran2 && q < k; q++) {
                    float t = 0.f;
                    OPH_CHECK(hipEventElapsedTime(&t, c.chunk_events[2 * q],
                                                  c.chunk_events[2 * q + 1]));
                    p2_ms += t;
                }
            }
            p2_launches += std::max(ran2, 0);
            iter0 = h.iter;
            if (h.halt == kHaltConverged) break;
            if (h.halt == kHaltCompact) {
                if ((rc = compact(c, cfg, s, &ncompact, lcap))) return rc;
                launch_set_halt(d, kHaltNone, s);
                if (d.fuse_first) launch_first(d, s);  // the next iteration's start
            } else if (h.halt) {
                return halt_status(h);
            }
            k = next_k;
        }
    }
    const int num_iters = h.iter;
    c.last_m = m;
    c.last_nnz = nnz;
    c.last_iters = num_iters;
    if (tracking) {
        // the state the reference tracks after its last iteration
        // (pms.cu:94-100): ess_hat_final, update_classes_final; its
        // trackers arrive with the last read of the control block
        launch_final_ess(d, s);
        launch_classes_final(d, s);
        launch_track(d, c.t_low, s);
    }
    if (!one_read) {
        launch_final(d, low1, ess, s);
        if (events) OPH_CHECK(hipEventRecord(c.ev1, s));
        if ((rc = read_ctrl(c, s))) return rc;
    }
    if (tracking) {
        fill_track(tr, num_iters - 1, m, ess_true, h.trk_l1, (unsigned long long)h.trk_linf,
                   (unsigned long long)h.trk_unreduced, (unsigned long long)h.trk_ess);
        // scaling by the norms of low_true, over all m+1 entries
        // (openph.cu:188-193)
        for (int i = 0; i <= m; i++) {
            if (tr->err_lone) tr->err_lone[i] = tr->err_lone[i] / norm1;
            if (tr->err_linf) tr->err_linf[i] = tr->err_linf[i] / norminf;
        }
    }
    if (stats) {
        stats->num_iters = num_iters;
        stats->compactions = ncompact;
        stats->overflows = h.overflows;
        stats->additions = (int64_t)h.additions;
        stats->merge_elems = (int64_t)h.merge_elems;
        stats->pool_peak = (int64_t)std::max<unsigned long long>(h.pool_peak, (unsigned long long)nnz);
        float ms = 0.f;
        if (events) OPH_CHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
        stats->device_ms = ms;
        stats->m = m;
        stats->lanes = sh.lanes;
        stats->group_lds = sh.wlds;
        stats->spills = h.spills;
        stats->handoffs = h.handoffs;
        stats->phase2_launches = p2_launches;
        stats->phase2_ms = p2_ms;
        stats->schedule = sched;
        stats->host_reads = reads;
    }
    if (iter_times) {
        tr->time_track[0] = 0.f;
        if (sched == kSchedSync)
            for (int i = 1; i < num_iters && i <= m; i++) {
                float ms = 0.f;
                OPH_CHECK(hipEventElapsedTime(&ms, evs[2 * i], evs[2 * i + 1]));
                tr->time_track[i] = ms;
            }
    }
    return OPENPH_OK;
}

// ---------------------------------------------------------------------------
// Running batches (the server): members join between chunks of iterations and
// leave when their own lows stop repeating.
// ---------------------------------------------------------------------------

int live_start(Context &c, LiveBatch &lb, const RunCfg &cfg, int cls, int max_members,
               size_t pool_elems, hipStream_t s) {
    Dev &d = c.d;
    if (!c.num_cu) {
        hipDeviceProp_t prop;
        OPH_CHECK(hipGetDeviceProperties(&prop, c.device));
        c.num_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    }
    if (!c.h_ctrl) OPH_CHECK(hipHostMalloc((void **)&c.h_ctrl, sizeof(Ctrl), hipHostMallocDefault));
    if (!c.h_word) OPH_CHECK(hipHostMalloc((void **)&c.h_word, 4 * sizeof(unsigned long long),
                                           hipHostMallocDefault));
    OPH_CHECK(grow(&lb.d_mstart, lb.c_mstart, (size_t)max_members + 1));
    OPH_CHECK(grow(&lb.d_mdup, lb.c_mdup, (size_t)max_members + 1));
    // an append's offsets, sized for a full batch here: a device free while
    // the batch runs would wait for every stream of the device
    OPH_CHECK(grow(&c.b_moff, c.c_bmoff, (size_t)max_members + 1));
    OPH_CHECK(grow(&c.b_noff, c.c_bnoff, (size_t)max_members + 1));
    OPH_CHECK(grow(&c.b_err, c.c_berr, 1));
    if (lb.c_hmdup < (size_t)max_members + 1) {
        if (lb.h_mdup) (void)hipHostFree(lb.h_mdup);
        lb.h_mdup = nullptr;
        OPH_CHECK(hipHostMalloc((void **)&lb.h_mdup, sizeof(int) * (max_members + 1),
                                hipHostMallocDefault));
        lb.c_hmdup = (size_t)max_members + 1;
    }
    d.m = 0;
    d.nnz = 0;
    d.ntiles = 0;
    d.slot0 = cfg.slot0;
    d.ltrue = nullptr;
    d.trk_part = d.trk_ring = d.its = nullptr;
    d.mdup = lb.d_mdup;
    d.mstart = lb.d_mstart;
    d.nmem = 0;
    lb.cls = cls;
    lb.lanes = cfg.lanes > 1 ? cfg.lanes : (cls == 0 ? 8 : 64);
    lb.wlds = cfg.lanes > 1 ? cfg.wlds : (cls == 0 ? 80 : 1024);
    if (!phase2_shape_ok(lb.lanes, lb.wlds)) return OPENPH_EINVAL;
    lb.grid = 2 * phase2_blocks_per_cu(lb.lanes, lb.wlds) * c.num_cu;
    lb.long_grid = c.num_cu;
    d.wide_max = lb.wlds >= 1024 && cfg.wide ? 3 * c.num_cu : 0;
    d.fuse_first = 0;  // members join between chunks: every iteration starts with k_first
    lb.mstart.assign(1, 0);
    lb.slot_of_len.clear();
    lb.nslots = 0;
    lb.compactions = 0;
    lb.lcap = std::min(c.cap_pool, std::min(pool_elems, cfg.pool_max));
    launch_ctrl_init(d, (unsigned long long)lb.lcap, 1 << 30, s);
    c.h_ctrl->iter = 0;
    c.h_ctrl->pool_top = 0;
    c.h_ctrl->halt = kHaltNone;
    lb.running = true;
    return OPENPH_OK;
}

int live_append(Context &c, LiveBatch &lb, int count, const int *ms, const int *nnzs,
                const std::vector<int> *lens, const int *rows1, const int *cols1, int nnz,
                hipStream_t s) {
    Dev &d = c.d;
    if (count <= 0) return OPENPH_OK;
    const int a = d.m;
    std::vector<int> moff((size_t)count + 1), noff((size_t)count + 1);
    long long M = a, N = 0;
    for (int k = 0; k < count; k++) {
        moff[k] = (int)M;
        noff[k] = (int)N;
        M += ms[k];
        N += nnzs[k];
    }
    moff[count] = (int)M;
    noff[count] = (int)N;
    if (M > 0x7fffffffLL || N != nnz) return OPENPH_EINVAL;
    // pseudo-dimension slots: lengths not seen in the batch get new ones
    int maxl = (int)lb.slot_of_len.size() - 1;
    for (int k = 0; k < count; k++)
        for (int L : lens[k]) maxl = std::max(maxl, L);
    if ((size_t)maxl + 1 > lb.slot_of_len.size()) lb.slot_of_len.resize((size_t)maxl + 1, -1);
    for (int k = 0; k < count; k++)
        for (int L : lens[k])
            if (lb.slot_of_len[L] < 0) lb.slot_of_len[L] = lb.nslots++;
    if (lb.nslots > kMaxSlots) return OPENPH_EDIMS;
    OPH_CHECK(grow(&c.b_moff, c.c_bmoff, (size_t)count + 1));
    OPH_CHECK(grow(&c.b_noff, c.c_bnoff, (size_t)count + 1));
    OPH_CHECK(grow(&c.b_err, c.c_berr, 1));
    for (int k = 1; k <= count; k++) lb.mstart.push_back(moff[k]);
    // offsets, slot map and member starts go up from pinned staging, reused
    // once the previous append's uploads are done
    const size_t nm = lb.mstart.size(), ns = (size_t)maxl + 1;
    const size_t need = 2 * ((size_t)count + 1) + ns + nm;
    if (!lb.ev_stage) OPH_CHECK(hipEventCreateWithFlags(&lb.ev_stage, hipEventDisableTiming));
    else OPH_CHECK(hipEventSynchronize(lb.ev_stage));
    if (lb.c_hstage < need) {
        // pinned allocations are slow and their frees wait for the device:
        // room for a full batch, grown by doubling
        const size_t n = std::max(2 * need, 3 * (lb.c_mstart + 1) + 1024);
        if (lb.h_stage) (void)hipHostFree(lb.h_stage);
        lb.h_stage = nullptr;
        lb.c_hstage = 0;
        OPH_CHECK(hipHostMalloc((void **)&lb.h_stage, sizeof(int) * n, hipHostMallocDefault));
        lb.c_hstage = n;
    }
    int *hm = lb.h_stage, *hn = hm + count + 1, *hs = hn + count + 1, *hk = hs + ns;
    std::memcpy(hm, moff.data(), sizeof(int) * (count + 1));
    std::memcpy(hn, noff.data(), sizeof(int) * (count + 1));
    std::memcpy(hs, lb.slot_of_len.data(), sizeof(int) * ns);
    std::memcpy(hk, lb.mstart.data(), sizeof(int) * nm);
    OPH_CHECK(hipMemcpyAsync(c.b_moff, hm, sizeof(int) * (count + 1), hipMemcpyHostToDevice, s));
    OPH_CHECK(hipMemcpyAsync(c.b_noff, hn, sizeof(int) * (count + 1), hipMemcpyHostToDevice, s));
    OPH_CHECK(hipMemcpyAsync(d.slotmap, hs, sizeof(int) * ns, hipMemcpyHostToDevice, s));
    launch_batch_offsets(rows1, cols1, c.b_noff, c.b_moff, count, nnz, c.b_rows, c.b_cols,
                         c.b_err, s);
    d.m = (int)M;
    d.ntiles = (int)((M + kTile - 1) / kTile);
    launch_setup_range(d, a, 0, nnz, c.b_rows, c.b_cols, s);
    launch_slots_range(d, a, lb.nslots, s);
    launch_phase0_range(d, a, s);
    d.nmem = (int)nm - 1;
    OPH_CHECK(hipMemcpyAsync(lb.d_mstart, hk, sizeof(int) * nm, hipMemcpyHostToDevice, s));
    OPH_CHECK(hipEventRecord(lb.ev_stage, s));
    return OPENPH_OK;
}

int live_chunk_launch(Context &c, LiveBatch &lb, int iters, hipStream_t s) {
    Dev &d = c.d;
    const Shape sh{lb.lanes, lb.wlds, lb.grid, lb.long_grid};
    for (int q = 0; q < iters; q++) {
        int rc = enqueue_iteration(d, sh, s, nullptr, nullptr, nullptr);
        if (rc) return rc;
    }
    if (d.nmem)
        OPH_CHECK(hipMemcpyAsync(lb.h_mdup, lb.d_mdup, sizeof(int) * d.nmem,
                                 hipMemcpyDeviceToHost, s));
    OPH_CHECK(hipMemcpyAsync(c.h_ctrl, c.d.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s));
    return OPENPH_OK;
}

int live_chunk_finish(Context &c, LiveBatch &lb, const RunCfg &cfg, hipStream_t s) {
    Dev &d = c.d;
    OPH_CHECK(hipStreamSynchronize(s));
    const Ctrl &h = *c.h_ctrl;
    if (h.halt == kHaltCompact || h.halt == kHaltSuspend) {
        RunCfg cc = cfg;
        cc.pool_max = std::min(cfg.pool_max, c.cap_pool2 ? std::max(c.cap_pool, c.cap_pool2)
                                                         : c.cap_pool);
        int rc;
        if (h.halt == kHaltSuspend) {
            // the rest of the iteration follows on the stream; a pass that
            // runs out again is caught at the end of the next chunk
            rc = resume_suspended(c, cc, lb.lanes, lb.wlds, lb.grid, lb.long_grid, s,
                                  &lb.compactions, lb.lcap);
        } else {
            rc = compact(c, cc, s, &lb.compactions, lb.lcap);
            launch_set_halt(d, kHaltNone, s);
        }
        if (rc) return rc;
        c.h_ctrl->halt = kHaltNone;
        return OPENPH_OK;
    }
    if (h.halt && h.halt != kHaltConverged) return halt_status(h);
    return OPENPH_OK;
}

int live_retire(Context &c, LiveBatch &lb, int k, int *low1, int *ess, hipStream_t s) {
    if (k < 0 || k + 1 >= (int)lb.mstart.size()) return OPENPH_EINVAL;
    launch_retire_range(c.d, lb.mstart[k], lb.mstart[k + 1], low1, ess, s);
    return OPENPH_OK;
}

void live_trim(Context &c, LiveBatch &lb, const std::vector<char> &retired, hipStream_t s) {
    Dev &d = c.d;
    int n = (int)lb.mstart.size() - 1;
    while (n > 0 && retired[n - 1]) n--;
    if (n == (int)lb.mstart.size() - 1) return;
    lb.mstart.resize((size_t)n + 1);
    d.m = lb.mstart[n];
    d.ntiles = (d.m + kTile - 1) / kTile;
    d.nmem = n;
    (void)s;
}

// The algorithms other than pms run their own order on the device.
enum SerialAlgo { kAlgoPms = 0, kAlgoStandard, kAlgoTwist, kAlgoRow, kAlgoRowAll };
static SerialAlgo serial_algo(const char *a) {
    if (!std::strcmp(a, "standard") || !std::strcmp(a, "standard_parallel")) return kAlgoStandard;
    if (!std::strcmp(a, "twist") || !std::strcmp(a, "twist_parallel")) return kAlgoTwist;
    if (!std::strcmp(a, "ph_row")) return kAlgoRow;
    if (!std::strcmp(a, "ph_row_parallel")) return kAlgoRowAll;
    return kAlgoPms;
}

// standard (standard.cu:18-54) and twist (twist.cu:16-87) on the device.
// The column order comes from the column lengths (pseudo-dimension = length
// - 1, dimensions.cu:54-104): standard takes every column in index order;
// twist takes the pseudo-dimensions from the highest down to 1, each in index
// order. num_iters is the reference's: one iteration per column taken, plus
// one. time_track[i] is the device time of the i-th column.
// ph_row (ph_row.cu:38-121) and ph_row_parallel (ph_row_parallel.cu:20-84)
// take the rows from the last to the first: num_iters = m + 1 and
// time_track[i] is the device time of the i-th row taken. ph_row_parallel's
// ess is the reference's: 0 exactly at the lows.
// For all four the err_* trackers, which the reference evaluates over all m
// columns after every step, are left as the caller set them.
static int run_serial(Context &c, const int *rows1, const int *cols1, int m, int nnz, int *low1,
                      int *ess, hipStream_t s, const RunCfg &cfg, SerialAlgo algo,
                      openph_stats *stats, const Track *tr) {
    const bool twist = algo == kAlgoTwist, rowalg = algo == kAlgoRow || algo == kAlgoRowAll;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (m == 0) {
        if (stats) stats->num_iters = 1;
        return OPENPH_OK;
    }
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
    if (!c.ev0) {
        OPH_CHECK(hipEventCreate(&c.ev0));
        OPH_CHECK(hipEventCreate(&c.ev1));
    }
    if (!c.h_ctrl) OPH_CHECK(hipHostMalloc((void **)&c.h_ctrl, sizeof(Ctrl), hipHostMallocDefault));
    if (!c.h_word) OPH_CHECK(hipHostMalloc((void **)&c.h_word, 4 * sizeof(unsigned long long),
                                           hipHostMallocDefault));
    size_t lcap = std::min(c.cap_pool, pool_want);
    launch_ctrl_init(d, (unsigned long long)lcap, cfg.max_iters, s);
    launch_setup(d, rows1, cols1, s);
    if (cfg.pool_elems) launch_pool_cap(d, cfg.pool_elems, lcap, s);
    std::vector<int> len((size_t)m);
    OPH_CHECK(hipMemcpyAsync(len.data(), d.len, sizeof(int) * m, hipMemcpyDeviceToHost, s));
    if ((rc = read_ctrl(c, s))) return rc;
    if ((rc = err_from_bits(c.h_ctrl->err))) return rc;
    lcap = c.h_ctrl->pool_cap;
    std::vector<int> order;
    if (rowalg) {
    } else if (!twist) {
        order.reserve((size_t)m);
        for (int j = 0; j < m; j++) order.push_back(j);
    } else {
        const int maxlen = c.h_ctrl->maxlen;
        std::vector<int> count((size_t)maxlen + 2, 0), at((size_t)maxlen + 2, 0);
        for (int j = 0; j < m; j++) count[len[j]]++;
        int o = 0;
        for (int L = maxlen; L >= 2; L--) {
            at[L] = o;
            o += count[L];
        }
        order.resize((size_t)o);
        for (int j = 0; j < m; j++)
            if (len[j] >= 2) order[at[len[j]]++] = j;
    }
    // steps: columns taken, or rows for the row algorithms
    const int n = rowalg ? m : (int)order.size();
    OPH_CHECK(grow(&c.s_order, c.c_sorder, rowalg ? (size_t)5 * m + 4 : (size_t)n + 1));
    if (n && !rowalg)
        OPH_CHECK(hipMemcpyAsync(c.s_order, order.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    const size_t mm = (size_t)m;
    const RowLists rl{c.s_order,          c.s_order + mm,     c.s_order + 2 * mm,
                      c.s_order + 3 * mm, c.s_order + 4 * mm, c.s_order + 5 * mm};
    const bool times = tr && tr->time_track;
    if (times) OPH_CHECK(grow(&c.s_stamps, c.c_sstamps, (size_t)n + 1));
    OPH_CHECK(hipEventRecord(c.ev0, s));
    if (rowalg) launch_row_lists(d, rl, s);
    int k0 = 0, ncompact = 0, stalls = 0, launches = 0;
    const Ctrl &h = *c.h_ctrl;
    while (k0 < n) {
        if (rowalg)
            launch_ph_row(d, rl, algo == kAlgoRowAll, times ? c.s_stamps + 1 : nullptr,
                          times && launches == 0 ? c.s_stamps : nullptr, s);
        else
            launch_serial(d, c.s_order + k0, n - k0, twist, times ? c.s_stamps + 1 + k0 : nullptr,
                          times && k0 == 0 ? c.s_stamps : nullptr, s);
        launches++;
        if ((rc = read_ctrl(c, s))) return rc;
        if (h.err) return err_from_bits(h.err) ? err_from_bits(h.err) : OPENPH_EINVAL;
        if (!h.overflow) break;
        // a full pool: compact, then take the same column (or row) up again;
        // nwork is the position reached, or the row kernel's steps done
        stalls = h.nwork == 0 ? stalls + 1 : 0;
        if (stalls >= 2) return OPENPH_ENOMEM;
        if (!rowalg) k0 += h.nwork;
        if ((rc = compact(c, cfg, s, &ncompact, lcap))) return rc;
        launch_iter_begin(d, s);  // clears the overflow flag
    }
    // standard and twist keep the ess they set column by column, with no
    // final pass (standard.cu:18-54, twist.cu:16-87): twist never takes the
    // columns of pseudo-dimension 0 (twist.cu:30), whose lows stay as given
    if (algo == kAlgoRowAll) launch_ess_at_lows(d, s);
    if (algo == kAlgoRowAll || algo == kAlgoStandard || algo == kAlgoTwist)
        launch_output(d, low1, ess, s);
    else
        launch_final(d, low1, ess, s);
    OPH_CHECK(hipEventRecord(c.ev1, s));
    if ((rc = read_ctrl(c, s))) return rc;
    if (stats) {
        stats->num_iters = n + 1;
        stats->compactions = ncompact;
        stats->overflows = launches - 1;
        stats->additions = (int64_t)h.additions;
        stats->merge_elems = (int64_t)h.merge_elems;
        stats->pool_peak = (int64_t)std::max<unsigned long long>(h.pool_top, (unsigned long long)nnz);
        float ms = 0.f;
        OPH_CHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
        stats->device_ms = ms;
        stats->m = m;
        stats->lanes = kLongBlockLanes;
        stats->spills = h.spills;
        stats->host_reads = launches;
    }
    if (times) {
        int khz = 0;
        OPH_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c.device));
        std::vector<unsigned long long> st((size_t)n + 1);
        OPH_CHECK(hipMemcpy(st.data(), c.s_stamps, sizeof(unsigned long long) * (n + 1),
                            hipMemcpyDeviceToHost));
        tr->time_track[0] = 0.f;
        for (int k = 1; k <= n && k <= m; k++)
            tr->time_track[k] = khz > 0 && st[k] >= st[k - 1] ? (float)((double)(st[k] - st[k - 1]) / khz) : 0.f;
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
                    const Track *tr) {
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
    const SerialAlgo algo = serial_algo(algstr);
    if (algo != kAlgoPms)
        rc = run_serial(*c, c->h_rows, c->h_cols, m, nnz, c->h_low, c->h_ess, s, cfg, algo, stats,
                        tr);
    else
        rc = run_pms(*c, c->h_rows, c->h_cols, m, nnz, c->h_low, c->h_ess, s, cfg, stats, tr);
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
        case OPENPH_ESCRATCH:
            return "input larger than the static scratch of the no-allocation entry point";
        default: return "unknown status";
    }
}

int openph_v2(const char *algstr, const int *rows1, const int *cols1, int m, int nnz,
              const int *low_true1, const openph_options *opt, int *low1, int *ess,
              float *trackers, openph_stats *stats) {
    Track tr;
    if (trackers && opt && opt->track && m >= 0) {
        const size_t n = (size_t)m + 1;
        tr.time_track = trackers + 4 * n;
        if (low_true1) {
            tr.low_true1 = low_true1;
            tr.err_linf = trackers;
            tr.err_lone = trackers + n;
            tr.err_redu = trackers + 2 * n;
            tr.err_ess = trackers + 3 * n;
        }
    }
    return run_host(algstr, rows1, cols1, m, nnz, opt, low1, ess, stats, &tr);
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
    openph_options opt;
    std::memset(&opt, 0, sizeof(opt));
    opt.device = gpu_number;
    openph_stats st;
    std::memset(&st, 0, sizeof(st));
    Track tr;
    tr.low_true1 = h_low_true;
    tr.err_linf = err_linf;
    tr.err_lone = err_lone;
    tr.err_redu = err_redu;
    tr.err_ess = err_ess;
    tr.time_track = time_track;
    int rc = run_host(algstr, h_rows_in, h_cols_in, m, nnz, &opt, h_low, h_ess, &st, &tr);
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
