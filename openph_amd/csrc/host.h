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
                        h_rows, h_cols, h_low, h_ess, b_moff, b_noff, b_rows, b_cols, b_err};
        for (void *p : bufs)
            if (p) (void)hipFree(p);
        if (h_ctrl) (void)hipHostFree(h_ctrl);
        if (h_word) (void)hipHostFree(h_word);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        for (hipEvent_t e : iter_events) (void)hipEventDestroy(e);
        for (hipEvent_t e : p2_events) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

struct RunCfg {
    int slot0 = 32;
    int max_iters = 1000000;
    size_t pool_elems = 0;
    size_t pool_max = 0xFFFFFFFFull;
    int lanes = 0;
    int wlds = 0;
    bool track = false;
    bool timing = true;
};

RunCfg cfg_from(const openph_options *opt, int nnz);
// Sizes a workspace for batches of up to m columns, nnz entries and a pool of
// pool_elems entries, so that serving does not allocate (hipFree waits for
// the whole device).
int reserve(Context &c, size_t m, size_t nnz, size_t pool_elems);
int set_device(int device);
// Reduction of a batch of `count` matrices held on the device (local 1-based
// COO); low1/ess are device outputs.
int run_batch(Context &c, int count, const int *ms, const int *nnzs, const int *rows1,
              const int *cols1, int *low1, int *ess, hipStream_t s, const RunCfg &cfg,
              openph_stats *stats);

}  // namespace oph
