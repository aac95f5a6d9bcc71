// pms on a device-resident matrix without any allocation, for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc). The reference has no counterpart:
// its openph() allocates and frees every buffer per call
// (src/cuda/openph.cu:72-230).
//
// Every buffer of the reduction (per-column state, tile scans, the control
// block and both column pools) is carved out of one file-scope device array,
// `scratch`, of fixed size. Each GPU holds its own copy from the library's
// load until the process ends. Host code never uses the array's name as an
// address (on the host it is only a stand-in): the first call on a device
// takes the device address with hipGetSymbolAddress, with that device
// current, and keeps it for later calls.
//
// Limits: m <= kNaMaxM columns, nnz <= kNaMaxNnz entries, and two pools of
// kNaPool int32 entries each: the initial layout (nnz + slot_elems per
// column) must fit one of them, and a reduction whose live columns outgrow it
// returns OPENPH_ENOMEM. Larger inputs are refused with OPENPH_ESCRATCH
// before anything runs.
//
// The scratch keeps the last call's contents; every buffer is initialised by
// the reduction's own set-up kernels. Calls on one device share the array and
// would race, so they are serialised by a per-device lock; each call returns
// once its outputs are written.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "host.h"

namespace oph {
namespace {

constexpr size_t kNaMaxM = size_t(3) << 16;      // 196,608 columns
constexpr size_t kNaMaxNnz = 3 * kNaMaxM;        // 589,824 COO entries
constexpr size_t kNaPool = size_t(8) << 20;      // int32 entries per pool
constexpr size_t kNaCols = kNaMaxM + 1;
constexpr size_t kNaTiles = (kNaMaxM + kTile - 1) / kTile + 1;
constexpr int kNaIntArrays = 16;  // the int / unsigned per-column arrays of Dev

constexpr size_t al(size_t x) { return (x + 255) & ~size_t(255); }

constexpr size_t kScratchBytes = kNaIntArrays * al(kNaCols * 4)   // low .. off2
                                 + al((kNaCols + 1) * 4)          // slotmap
                                 + al(kNaCols * sizeof(int4))     // pivdesc
                                 + 2 * al(kNaTiles * kMaxSlots * 4)  // tile_max, tile_pre
                                 + al((kNaTiles + 1) * 8)         // scan_tmp
                                 + al(sizeof(Ctrl))               // ctrl
                                 + 2 * al(kNaPool * 4);           // pool, pool2

constexpr int kNaMaxDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

struct NaDevice {
    std::mutex mu;
    bool ready = false;
    unsigned char *ptr = nullptr;  // device address of this GPU's scratch
    Context *ctx = nullptr;        // views of the scratch; never freed
    Ctrl h_ctrl;                   // host mirror of the control block
    unsigned long long h_word[4] = {0, 0, 0, 0};
};

NaDevice g_dev[kNaMaxDevices];

// Dev pointers carved from the scratch, and capacities that make every size
// check of the reduction pass without growing anything.
void carve(Context &c, unsigned char *base) {
    size_t o = 0;
    auto take = [&](size_t bytes) {
        unsigned char *p = base + o;
        o += al(bytes);
        return p;
    };
    Dev &d = c.d;
    int **ints[] = {&d.low, &d.arglow, &d.classes, &d.ess, &d.left, &d.beta, &d.first,
                    &d.clear, &d.newpiv, &d.slot, &d.len, &d.cap, &d.work, &d.longw};
    for (int **p : ints) *p = reinterpret_cast<int *>(take(kNaCols * 4));
    d.off = reinterpret_cast<unsigned *>(take(kNaCols * 4));
    d.off2 = reinterpret_cast<unsigned *>(take(kNaCols * 4));
    static_assert(sizeof(ints) / sizeof(ints[0]) + 2 == kNaIntArrays, "per-column arrays");
    d.slotmap = reinterpret_cast<int *>(take((kNaCols + 1) * 4));
    d.pivdesc = reinterpret_cast<int4 *>(take(kNaCols * sizeof(int4)));
    d.tile_max = reinterpret_cast<int *>(take(kNaTiles * kMaxSlots * 4));
    d.tile_pre = reinterpret_cast<int *>(take(kNaTiles * kMaxSlots * 4));
    d.scan_tmp = reinterpret_cast<unsigned long long *>(take((kNaTiles + 1) * 8));
    d.ctrl = reinterpret_cast<Ctrl *>(take(sizeof(Ctrl)));
    d.pool = reinterpret_cast<int *>(take(kNaPool * 4));
    d.pool2 = reinterpret_cast<int *>(take(kNaPool * 4));
    c.c_low = c.c_arglow = c.c_classes = c.c_ess = c.c_left = c.c_beta = c.c_first = c.c_clear =
        c.c_newpiv = c.c_slot = c.c_len = c.c_cap = c.c_work = c.c_longw = c.c_off = c.c_off2 =
            c.c_pivdesc = kNaCols;
    c.cap_slotmap = kNaCols + 1;
    c.c_tmax = c.c_tpre = kNaTiles * kMaxSlots;
    c.cap_scan = kNaTiles + 1;
    c.c_ctrl = 1;
    c.cap_pool = c.cap_pool2 = kNaPool;
}

}  // namespace
}  // namespace oph

using namespace oph;

extern "C" {

int openph_pms_device_noalloc(const int *rows1, const int *cols1, int m, int nnz, int *low1,
                              int *ess, void *stream, const openph_options *opt,
                              openph_stats *stats) {
    if (m < 0 || nnz < 0) return OPENPH_EINVAL;
    if ((size_t)m > kNaMaxM || (size_t)nnz > kNaMaxNnz) return OPENPH_ESCRATCH;
    const int dev = opt ? opt->device : 0;
    if (dev < 0 || dev >= kNaMaxDevices) return OPENPH_EINVAL;
    int rc = set_device(dev);
    if (rc) return rc;
    NaDevice &nd = g_dev[dev];
    std::lock_guard<std::mutex> g(nd.mu);
    if (!nd.ready) {
        // once per device, with the device current: the address of this
        // GPU's copy of the array
        void *p = nullptr;
        OPH_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        nd.ptr = static_cast<unsigned char *>(p);
        nd.ctx = new Context();
        nd.ctx->device = dev;
        carve(*nd.ctx, nd.ptr);
        nd.ctx->h_ctrl = &nd.h_ctrl;
        nd.ctx->h_word = nd.h_word;
        nd.ready = true;
    }
    RunCfg cfg = cfg_from(opt, nnz);
    cfg.no_alloc = true;
    cfg.pool_max = std::min<size_t>(cfg.pool_max, kNaPool);
    if (cfg.pool_elems > cfg.pool_max) cfg.pool_elems = cfg.pool_max;
    return run_pms(*nd.ctx, rows1, cols1, m, nnz, low1, ess, (hipStream_t)stream, cfg, stats,
                   nullptr);
}

}  // extern "C"
