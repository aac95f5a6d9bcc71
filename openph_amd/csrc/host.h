// Host-side internals shared by the C ABI (openph.cpp) and the request
// server (server.cpp). Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/openph.h"
#include "engine.h"

namespace oph {

#define OPH_CHECK(expr)                                  \
    do {                                                 \
        hipError_t e_ = (expr);                          \
        if (e_ != hipSuccess) {                          \
            last_hip_error() = e_;                       \
            return OPENPH_EDEVICE;                       \
        }                                                \
    } while (0)

inline hipError_t &last_hip_error() {
    static thread_local hipError_t e = hipSuccess;
    return e;
}

// Grows a device buffer to at least `need` elements. Growth is geometric
// (x1.5) so that a workspace serving varying sizes reallocates rarely:
// hipFree waits for the whole device.
template <typename T>
inline hipError_t grow(T **p, size_t &cap, size_t need) {
    if (need <= cap && *p) return hipSuccess;
    if (*p) {
        hipError_t e = hipFree(*p);
        if (e != hipSuccess) return e;
        *p = nullptr;
    }
    size_t n = std::max<size_t>(need, 1);
    if (cap) n = std::max(n, cap + cap / 2);
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess && n > need) {
        (void)hipGetLastError();
        n = std::max<size_t>(need, 1);
        e = hipMalloc((void **)p, n * sizeof(T));
    }
    if (e != hipSuccess) {
        cap = 0;
        *p = nullptr;
        return e;
    }
    cap = n;
    return hipSuccess;
}

// An instantiated graph of `iters` loop iterations (Phase I, worklist,
// Phase II on fixed persistent grids, bookkeeping, checks), valid for the
// device buffers and shape it was captured with.
struct GraphEntry {
    bool valid = false;
    Dev key;
    int lanes = 0, wlds = 0, grid = 0, long_grid = 0, iters = 0;
    bool timed = false;
    // a whole reduction: the set-up and Phase 0 before the iterations and
    // the gated outputs after them, for these inputs, outputs and pool
    bool whole = false;
    const int *rows1 = nullptr, *cols1 = nullptr;
    int *low1 = nullptr, *ess = nullptr;
    size_t lcap = 0, pool_elems = 0;
    int max_iters = 0;
    unsigned long long used = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

constexpr int kStampRing = 128;  // device-clock stamps kept per chunk (2 per iteration)

// Device workspace of one GPU, reused by every call on that GPU.
struct Context {
    int device = 0;
    std::mutex mu;
    Dev d;
    size_t cap_m = 0, cap_tiles = 0, cap_pool = 0, cap_pool2 = 0, cap_slotmap = 0;
    size_t cap_scan = 0;
    // per-column arrays share cap_m; remember each allocation separately
    size_t c_low = 0, c_arglow = 0, c_classes = 0, c_ess = 0, c_left = 0, c_beta = 0,
           c_first = 0, c_clear = 0, c_newpiv = 0, c_slot = 0, c_len = 0, c_off = 0,
           c_off2 = 0, c_tmax = 0, c_tpre = 0, c_ctrl = 0, c_cap = 0, c_work = 0, c_pivdesc = 0, c_longw = 0;
    int num_cu = 0;
    // staging for host-pointer entry points
    int *h_rows = nullptr, *h_cols = nullptr, *h_low = nullptr, *h_ess = nullptr;
    size_t c_hrows = 0, c_hcols = 0, c_hlow = 0, c_hess = 0;
    // batches: per-matrix offsets and the block-diagonal COO
    int *b_moff = nullptr, *b_noff = nullptr, *b_rows = nullptr, *b_cols = nullptr;
    int *b_err = nullptr;
    size_t c_bmoff = 0, c_bnoff = 0, c_brows = 0, c_bcols = 0, c_berr = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> iter_events;
    std::vector<hipEvent_t> p2_events;  // start/end pairs of Phase II launches
    Ctrl *h_ctrl = nullptr;                // pinned mirror of d.ctrl
    unsigned long long *h_word = nullptr;  // pinned scratch words
    // queued / graph schedules: Phase II events of the iterations of one
    // chunk, cached graphs, and the iteration count of the last reduction
    // of each size (the first chunk's length)
    std::vector<hipEvent_t> chunk_events;
    GraphEntry graphs[4];
    unsigned long long graph_clock = 0;
    int last_m = -1, last_nnz = -1, last_iters = 0;
    int last_lanes = 0, last_wlds = 0;  // the Phase II shape the size picked
    unsigned long long *d_stamps = nullptr, *h_stamps = nullptr;  // graph timing
    int wall_khz = 0;                                            // device clock rate
    // trackers: 0-based low_true on the device; per-iteration trackers and
    // clock stamps of the queued / graph schedules (device rings and their
    // pinned host mirrors)
    int *t_low = nullptr;
    size_t c_tlow = 0;
    unsigned long long *t_part = nullptr, *t_ring = nullptr, *t_its = nullptr;
    size_t c_tpart = 0, c_tring = 0, c_tits = 0;
    unsigned long long *h_ring = nullptr, *h_its = nullptr;
    // serial algorithms: column order and per-column clock stamps
    int *s_order = nullptr;
    unsigned long long *s_stamps = nullptr;
    size_t c_sorder = 0, c_sstamps = 0;
    // filtration builders: sort keys and values (in, out), positions,
    // column offsets, pairwise distances and sort scratch
    unsigned long long *f_key = nullptr, *f_key2 = nullptr;
    unsigned *f_val = nullptr, *f_val2 = nullptr;
    int *f_pos = nullptr, *f_nnz = nullptr, *f_off = nullptr;
    double *f_dist = nullptr;
    unsigned char *f_tmp = nullptr;
    size_t c_fkey = 0, c_fkey2 = 0, c_fval = 0, c_fval2 = 0, c_fpos = 0, c_fnnz = 0, c_foff = 0,
           c_fdist = 0, c_ftmp = 0;
    bool init = false;

    Context() = default;
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    // Releases the device workspace (the per-device default contexts live
    // until the process ends; server workers free theirs).
    ~Context() {
        void *bufs[] = {d.low, d.arglow, d.classes, d.ess, d.left, d.beta, d.first, d.clear,
                        d.newpiv, d.slot, d.len, d.cap, d.work, d.longw, d.pivdesc, d.off, d.off2,
                        d.scan_tmp, d.slotmap, d.tile_max, d.tile_pre, d.pool, d.pool2, d.ctrl,
                        h_rows, h_cols, h_low, h_ess, b_moff, b_noff, b_rows, b_cols, b_err,
                        t_low, t_part, t_ring, t_its, s_order, s_stamps, f_key, f_key2, f_val,
                        f_val2, f_pos, f_nnz, f_off, f_dist, f_tmp};
        for (void *p : bufs)
            if (p) (void)hipFree(p);
        if (h_ring) (void)hipHostFree(h_ring);
        if (h_its) (void)hipHostFree(h_its);
        if (h_ctrl) (void)hipHostFree(h_ctrl);
        if (h_word) (void)hipHostFree(h_word);
        if (h_stamps) (void)hipHostFree(h_stamps);
        if (d_stamps) (void)hipFree(d_stamps);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        for (hipEvent_t e : iter_events) (void)hipEventDestroy(e);
        for (hipEvent_t e : p2_events) (void)hipEventDestroy(e);
        for (hipEvent_t e : chunk_events) (void)hipEventDestroy(e);
        for (GraphEntry &g : graphs) {
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            if (g.graph) (void)hipGraphDestroy(g.graph);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// How the host drives the iteration loop.
enum Schedule : int {
    kSchedAuto = 0,
    kSchedSync = 1,    // read the control block twice per iteration; Phase II
                       // grids sized by the worklist
    kSchedQueued = 2,  // queue `chunk` iterations between reads; fixed grids
    kSchedGraph = 3,   // the same chunks replayed from a cached hipGraph
};

// Per-iteration trackers of the drop-in (tracker.cu:92-134). Arrays of m+1
// floats, any of which may be null; low_true1 is the caller's 1-based array.
struct Track {
    const int *low_true1 = nullptr;
    float *err_linf = nullptr, *err_lone = nullptr, *err_redu = nullptr, *err_ess = nullptr;
    float *time_track = nullptr;
};

struct RunCfg {
    int schedule = kSchedAuto;
    int chunk = 0;     // iterations per chunk after the first (0 = 3)
    int slot0 = 32;
    int max_iters = 1000000;
    size_t pool_elems = 0;
    size_t pool_max = 0xFFFFFFFFull;
    int lanes = 0;
    int wlds = 0;
    bool track = false;
    bool timing = false;    // Phase II passes timed (openph_options.time_phase2)
    bool no_alloc = false;  // the context's buffers only: see run_pms
    bool wide = true;       // 256-lane teams for short worklists of long columns
};

RunCfg cfg_from(const openph_options *opt, int nnz);

// A batch of matrices that members join and leave between chunks of
// iterations (the server's workers). Members are appended at the end of the
// block-diagonal matrix; one whose lows no longer repeat is output and its
// columns made inert, and the batch ends when none is left. All members share
// one Phase II shape: class 0 for matrices whose input columns hold at most 4
// entries (simplicial complexes up to dimension 3), class 1 otherwise.
struct LiveBatch {
    bool running = false;
    int cls = 0;
    int lanes = 0, wlds = 0, grid = 0, long_grid = 0;  // Phase II shape
    std::vector<int> mstart;       // member k owns columns [mstart[k], mstart[k+1])
    std::vector<int> slot_of_len;  // input column length -> pseudo-dimension slot
    int nslots = 0;
    size_t lcap = 0;               // logical pool capacity
    int compactions = 0;
    int *d_mstart = nullptr, *d_mdup = nullptr, *h_mdup = nullptr;
    size_t c_mstart = 0, c_mdup = 0, c_hmdup = 0;
    // pinned staging of an append's offsets, slot map and member starts,
    // reused once the event recorded after their upload has completed
    int *h_stage = nullptr;
    size_t c_hstage = 0;
    hipEvent_t ev_stage = nullptr;
    LiveBatch() = default;
    LiveBatch(const LiveBatch &) = delete;
    LiveBatch &operator=(const LiveBatch &) = delete;
    ~LiveBatch() {
        if (d_mstart) (void)hipFree(d_mstart);
        if (d_mdup) (void)hipFree(d_mdup);
        if (h_mdup) (void)hipHostFree(h_mdup);
        if (h_stage) (void)hipHostFree(h_stage);
        if (ev_stage) (void)hipEventDestroy(ev_stage);
    }
};
// Starts an empty batch of class cls on the workspace, with room for
// max_members members and a pool of pool_elems entries.
int live_start(Context &c, LiveBatch &lb, const RunCfg &cfg, int cls, int max_members,
               size_t pool_elems, hipStream_t s);
// Appends `count` members: ms/nnzs their sizes, lens[k] the distinct column
// lengths of member k's input (their union with the batch's must stay within
// kMaxSlots), rows1/cols1 their local 1-based COO concatenated on the device
// (nnz entries). The workspace must have room for the columns and entries.
// Everything is enqueued on s; nothing waits for the device.
int live_append(Context &c, LiveBatch &lb, int count, const int *ms, const int *nnzs,
                const std::vector<int> *lens, const int *rows1, const int *cols1, int nnz,
                hipStream_t s);
// Queues `iters` iterations and the copies of the control block and the
// members' duplicate flags (lb.h_mdup) to the host, without waiting.
int live_chunk_launch(Context &c, LiveBatch &lb, int iters, hipStream_t s);
// Waits for the chunk, compacts the pool when asked. Returns an OPENPH_E*
// code when the reduction failed; c.h_ctrl->halt is kHaltConverged when no
// low of any member repeats.
int live_chunk_finish(Context &c, LiveBatch &lb, const RunCfg &cfg, hipStream_t s);
// Member k's 1-based local low and ess into low1/ess (device), its columns
// made inert; trailing retired members' columns are given back.
int live_retire(Context &c, LiveBatch &lb, int k, int *low1, int *ess, hipStream_t s);
// Drops retired members at the end of the batch (their columns can be
// reused); `retired` flags every member.
void live_trim(Context &c, LiveBatch &lb, const std::vector<char> &retired, hipStream_t s);
// The workspace of a device, created on first use and kept for the process.
Context *context_for(int device);
// Filtrations built on the device into rows1/cols1 (1-based COO), queued on
// `s` without waiting (filtrations.hip).
int build_vr(Context &c, const double *points, int n, int dim, int *rows1, int *cols1,
             double *values, int *dims, hipStream_t s);
int build_cubical(Context &c, const float *field, int s0, int s1, int s2, int *rows1, int *cols1,
                  float *values, int *dims, hipStream_t s);
// Sizes a workspace for batches of up to m columns, nnz entries and a pool of
// pool_elems entries, so that serving does not allocate (hipFree waits for
// the whole device).
int reserve(Context &c, size_t m, size_t nnz, size_t pool_elems);
// The same for the filtration builders: matrices of up to m columns built
// from up to n points (filtrations.hip).
int reserve_builders(Context &c, size_t m, int n);
int set_device(int device);
// The pms reduction of one matrix held on the device (1-based COO); low1/ess
// are device outputs, tr (may be null) the drop-in's trackers.
int run_pms(Context &c, const int *rows1, const int *cols1, int m, int nnz, int *low1, int *ess,
            hipStream_t s, const RunCfg &cfg, openph_stats *stats, const Track *tr);
// Reduction of a batch of `count` matrices held on the device (local 1-based
// COO); low1/ess are device outputs.
int run_batch(Context &c, int count, const int *ms, const int *nnzs, const int *rows1,
              const int *cols1, int *low1, int *ess, hipStream_t s, const RunCfg &cfg,
              openph_stats *stats);

}  // namespace oph
