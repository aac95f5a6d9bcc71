// gfx950 kernels of the pms boundary-matrix reduction.
//
// State arrays are indexed by column j (0-based) and follow the reference's
// meaning (src/cuda/algorithms/pms/pms.cu:1-118):
//   low[j]     max row of the current column j, -1 if zero or cleared
//   arglow[i]  the column whose final low is i, -1 if none known yet
//   classes[j] 0 unknown, -1 negative (final pivot), 1 positive (cleared)
//   ess[j]     1 while cell j may be essential
//   left[i]    leftmost column holding row i (INT_MAX if none), beta.cu:2-15
//   beta[j]    max{i : left[i] = j}, beta.cu:2-15
// and, new here:
//   pivdesc[i] {arglow[i], off, len, left[i]}: the pivot column of row i
//              with the left rule of row i, so each step of a Phase II
//              chain finds its operand and its rule with one 16-byte load
//   work[]     the columns Phase II has to reduce in this iteration
//   longw[]    the columns handed to the long-column pass; between passes
//              the compaction's map of published pivots
// Columns live in a pool of int32 row indices: column j occupies
// pool[off[j] .. off[j]+len[j]) in ascending order inside a segment of
// cap[j] >= len[j] entries (at least slot0), so most Phase II results are
// written back in place; a column that outgrows its segment gets a new one
// from a bump allocator, and the host compacts the pool when it fills (inside
// a Phase II pass too: the pass is suspended, and runs again after it).
#include "engine.h"

namespace oph {

// ---------------------------------------------------------------------------
// wave / block helpers (wave64)
// ---------------------------------------------------------------------------

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

template <typename T>
struct MaxOp {
    __device__ T operator()(T a, T b) const { return a > b ? a : b; }
};
template <typename T>
struct SumOp {
    __device__ T operator()(T a, T b) const { return a + b; }
};

// x of the lane that DPP control CTRL selects, identity where it selects
// none or the row is masked off
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int dpp_or(int identity, int x) {
    return __builtin_amdgcn_update_dpp(identity, x, CTRL, ROW_MASK, 0xF, false);
}

// Inclusive scan of an int over the wave in registers (all 64 lanes
// active): row shifts by 1, 2, 4 and 8 inside each row of 16 lanes, then the
// row totals carried by the row broadcasts (lane 15 into rows 1 and 3, lane
// 31 into rows 2 and 3).
template <typename Op>
__device__ __forceinline__ int wave_inclusive_scan(int x, Op op, int identity) {
    x = op(x, dpp_or<0x111, 0xF>(identity, x));  // row_shr:1
    x = op(x, dpp_or<0x112, 0xF>(identity, x));  // row_shr:2
    x = op(x, dpp_or<0x114, 0xF>(identity, x));  // row_shr:4
    x = op(x, dpp_or<0x118, 0xF>(identity, x));  // row_shr:8
    x = op(x, dpp_or<0x142, 0xA>(identity, x));  // row_bcast:15
    x = op(x, dpp_or<0x143, 0xC>(identity, x));  // row_bcast:31
    return x;
}

// Exclusive scan over the whole block (blockDim.x multiple of 64). `lds`
// needs blockDim.x/64 entries. Returns the exclusive prefix; *total gets the
// block aggregate. Contains two __syncthreads(): every thread must call it.
// Ints scan the wave in registers; wider types through lane shuffles.
template <typename T, typename Op>
__device__ T block_exclusive_scan(T v, Op op, T identity, T *lds, T *total) {
    const int lane = lane_id(), wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    T incl, excl;
    if constexpr (sizeof(T) == sizeof(int) && T(-1) < T(0)) {
        incl = wave_inclusive_scan(v, op, identity);
        excl = dpp_or<0x138, 0xF>(identity, incl);  // wave_shr:1
    } else {
        incl = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            T t = __shfl_up(incl, d, 64);
            if (lane >= d) incl = op(incl, t);
        }
        excl = __shfl_up(incl, 1, 64);
        if (lane == 0) excl = identity;
    }
    if (lane == 63) lds[wid] = incl;
    __syncthreads();
    T pre = identity, tot = identity;
    for (int w = 0; w < nw; w++) {
        T x = lds[w];
        if (w < wid) pre = op(pre, x);
        tot = op(tot, x);
    }
    __syncthreads();
    *total = tot;
    return op(pre, excl);
}

// Wave-aggregated bump allocation: every lane of the wave calls it (lanes
// with active = false ask for 0). One 64-bit atomic per wave.
__device__ __forceinline__ unsigned long long wave_alloc(unsigned long long *top, unsigned need,
                                                         bool active) {
    const int lane = lane_id();
    unsigned long long v = active ? (unsigned long long)need : 0ull;
    unsigned long long incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        unsigned long long t = __shfl_up(incl, d, 64);
        if (lane >= d) incl += t;
    }
    unsigned long long total = __shfl(incl, 63, 64);
    unsigned long long base = 0;
    if (lane == 63 && total) base = atomicAdd(top, total);
    base = __shfl(base, 63, 64);
    return base + (incl - v);
}

// Block-aggregated append to a list, in thread order inside the block: one
// global atomic per block. Every thread of the block calls it.
__device__ __forceinline__ void block_append(int *list, int *count, int value, bool take) {
    __shared__ int wcnt[kBlock / 64];
    __shared__ int wbase[kBlock / 64];
    __shared__ int bbase;
    const unsigned long long b = __ballot(take);
    const int lane = lane_id(), wid = threadIdx.x >> 6;
    if (lane == 0) wcnt[wid] = __popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
            wbase[w] = tot;
            tot += wcnt[w];
        }
        bbase = tot ? atomicAdd(count, tot) : 0;
    }
    __syncthreads();
    if (take) list[bbase + wbase[wid] + __popcll(b & ((1ull << lane) - 1))] = value;
}

// The per-column passes of an iteration take kItems consecutive columns per
// thread. Their loads are issued together and waited for once: a pass over
// columns costs the latency of its chain of dependent loads (column state,
// then what it indexes), so no load waits behind another one it does not
// depend on.
//
// v[k] = a[base + k] for base + k < n, else fill. A full run is one 16-byte
// load: base is a multiple of kItems and every array starts 256-byte aligned.
__device__ __forceinline__ void load_run(const int *a, int base, int n, int fill,
                                         int (&v)[kItems]) {
    static_assert(kItems == 4, "one int4 per run");
    if (base + kItems <= n) {
        const int4 x = *reinterpret_cast<const int4 *>(a + base);
        v[0] = x.x;
        v[1] = x.y;
        v[2] = x.z;
        v[3] = x.w;
    } else {  // the last run: clamped loads, no branch per column
        int raw[kItems];
#pragma unroll
        for (int k = 0; k < kItems; k++) raw[k] = a[base + k < n ? base + k : n - 1];
#pragma unroll
        for (int k = 0; k < kItems; k++) v[k] = base + k < n ? raw[k] : fill;
    }
}

// v[k] = a[i[k]] for i[k] >= 0, else fill; the loads of a negative index
// read a[0] (every state array has m + 1 >= 1 entries) and are dropped.
__device__ __forceinline__ void gather_run(const int *a, const int (&i)[kItems], int fill,
                                           int (&v)[kItems]) {
    int raw[kItems];
#pragma unroll
    for (int k = 0; k < kItems; k++) raw[k] = a[i[k] > 0 ? i[k] : 0];
#pragma unroll
    for (int k = 0; k < kItems; k++) v[k] = i[k] >= 0 ? raw[k] : fill;
}

// Keeps the loads of v ahead of a branch on x that follows (the halt test),
// so that they are in flight together with the load of x.
__device__ __forceinline__ int issued(const int (&v)[kItems], int x) {
    asm volatile("" : "+v"(x) : "v"(v[0]), "v"(v[1]), "v"(v[2]), "v"(v[3]));
    return __builtin_amdgcn_readfirstlane(x);
}

__device__ __forceinline__ unsigned slot_size(int len, int slot0) {
    return (unsigned)(len > slot0 ? len : slot0);
}

// ---------------------------------------------------------------------------
// Tiled exclusive scan of per-column sizes (pool layout).
//   mode 0: setup layout, size = max(len, slot0)
//   mode 1: compaction, size = len for published pivots (read, never written
//           again), 0 for other cleared columns (never read again),
//           max(len, slot0) else. Whether a column is a published pivot
//           comes from arglow (longw[j] = 1 + the row it is the pivot of,
//           written by k_pivot_rows), not from classes: an algorithm may
//           mark a pivot's column positive, and its data must survive.
// Offsets go to `out`; the total to scan_tmp[ntiles] and ctrl->pool_top.
// ---------------------------------------------------------------------------

__device__ __forceinline__ unsigned col_size(const Dev &d, int j, int mode) {
    if (j >= d.m) return 0u;
    int n = d.len[j];
    if (mode == 0) return slot_size(n, d.slot0);
    if (d.longw[j]) return (unsigned)n;
    if (d.classes[j] == 1) return 0u;
    return slot_size(n, d.slot0);
}

// The scan runs over columns [a, d.m) in tiles of kTile.
__global__ void __launch_bounds__(kBlock) k_scan_tiles(Dev d, int mode, int a) {
    __shared__ unsigned long long lds[kBlock / 64];
    const int base = a + blockIdx.x * kTile + threadIdx.x * kItems;
    unsigned long long sum = 0;
#pragma unroll
    for (int k = 0; k < kItems; k++) sum += col_size(d, base + k, mode);
    unsigned long long tot;
    block_exclusive_scan(sum, SumOp<unsigned long long>(), 0ull, lds, &tot);
    if (threadIdx.x == 0) d.scan_tmp[blockIdx.x] = tot;
}

// nt tiles; with append the offsets start at the current pool top, which
// then moves past them (columns added to a running batch).
__global__ void __launch_bounds__(1024) k_scan_partials(Dev d, int nt, int append) {
    __shared__ unsigned long long lds[16];
    __shared__ unsigned long long s_base;
    if (threadIdx.x == 0) s_base = append ? d.ctrl->pool_top : 0ull;
    __syncthreads();
    unsigned long long running = s_base;
    for (int c0 = 0; c0 < nt; c0 += blockDim.x) {
        int t = c0 + threadIdx.x;
        unsigned long long v = t < nt ? d.scan_tmp[t] : 0ull;
        unsigned long long tot;
        unsigned long long ex = block_exclusive_scan(v, SumOp<unsigned long long>(), 0ull, lds, &tot);
        if (t < nt) d.scan_tmp[t] = ex + running;
        running += tot;
    }
    if (threadIdx.x == 0) {
        d.scan_tmp[nt] = running;
        d.ctrl->pool_top = running;
    }
}

__global__ void __launch_bounds__(kBlock) k_scan_apply(Dev d, int mode, unsigned *out, int a) {
    __shared__ unsigned long long lds[kBlock / 64];
    const int base = a + blockIdx.x * kTile + threadIdx.x * kItems;
    unsigned v[kItems];
    unsigned long long sum = 0;
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        v[k] = col_size(d, base + k, mode);
        sum += v[k];
    }
    unsigned long long tot;
    unsigned long long ex = block_exclusive_scan(sum, SumOp<unsigned long long>(), 0ull, lds, &tot);
    ex += d.scan_tmp[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        int j = base + k;
        if (j < d.m) {
            out[j] = (unsigned)ex;
            if (mode == 0) d.cap[j] = (int)v[k];
        }
        ex += v[k];
    }
}

static void launch_scan(const Dev &d, int mode, unsigned *out, hipStream_t s, int a = 0,
                        bool append = false) {
    const int nt = (d.m - a + kTile - 1) / kTile;
    if (nt <= 0) return;
    k_scan_tiles<<<nt, kBlock, 0, s>>>(d, mode, a);
    k_scan_partials<<<1, 1024, 0, s>>>(d, nt, append ? 1 : 0);
    k_scan_apply<<<nt, kBlock, 0, s>>>(d, mode, out, a);
}

// ---------------------------------------------------------------------------
// Setup: COO (1-based) -> pool + per-column state
// ---------------------------------------------------------------------------

// Also clears `zero` (n entries): the presence map of column lengths that
// k_init_cols fills.
__global__ void k_ctrl_init(Ctrl *c, unsigned long long cap, int max_iters, int *zero, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) zero[i] = 0;
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        c->pool_top = 0;
        c->pool_cap = cap;
        c->additions = 0;
        c->merge_elems = 0;
        c->err = 0;
        c->any_dup = 0;
        c->overflow = 0;
        c->nslots = 0;
        c->maxlen = 0;
        c->marks = 0;
        c->nwork = 0;
        c->nlong = 0;
        c->spills = 0;
        c->handoffs = 0;
        c->halt = kHaltNone;
        c->iter = 0;
        c->max_iters = max_iters;
        c->stalls = 0;
        c->prev_marks = 0;
        c->overflows = 0;
        c->suspended = 0;
        c->prev_add = 0;
        c->pool_peak = 0;
        for (int k = 0; k < kStripes; k++) {
            c->st_add[k] = 0;
            c->st_el[k] = 0;
            c->st_spill[k] = 0;
        }
    }
}

// The Phase II counters of a block into its stripe.
__device__ __forceinline__ void stripe_add(Ctrl *c, unsigned long long nadd, unsigned long long nel,
                                           int spills) {
    const int k = blockIdx.x % kStripes;
    if (nadd) {
        atomicAdd(&c->st_add[k], nadd);
        atomicAdd(&c->st_el[k], nel);
    }
    if (spills) atomicAdd(&c->st_spill[k], spills);
}

__device__ __forceinline__ void fold_stripes(Ctrl *c) {
    for (int k = 0; k < kStripes; k++) {
        c->additions += c->st_add[k];
        c->merge_elems += c->st_el[k];
        c->spills += c->st_spill[k];
        c->st_add[k] = 0;
        c->st_el[k] = 0;
        c->st_spill[k] = 0;
    }
}

void launch_ctrl_init(const Dev &d, unsigned long long cap, int max_iters, hipStream_t s) {
    const int n = d.m > 0 ? d.m + 2 : 0;  // slotmap has m + 2 entries
    k_ctrl_init<<<n > 0 ? (n + kBlock - 1) / kBlock : 1, kBlock, 0, s>>>(d.ctrl, cap, max_iters,
                                                                      d.slotmap, n);
}

// A caller-given pool size counts from the exact initial layout, which the
// setup scan leaves in pool_top: capacity = max(requested, layout), at most
// the allocation.
__global__ void k_pool_cap(Ctrl *c, unsigned long long requested, unsigned long long alloc) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        unsigned long long cap = requested > c->pool_top ? requested : c->pool_top;
        c->pool_cap = cap < alloc ? cap : alloc;
    }
}

void launch_pool_cap(const Dev &d, unsigned long long requested, unsigned long long alloc,
                     hipStream_t s) {
    k_pool_cap<<<1, 64, 0, s>>>(d.ctrl, requested, alloc);
}

// The first positions of sorted cols1[lo, hi) whose column (1-based) is >=
// ka and >= kb, found together by one wave: lanes 0-31 search for ka and
// lanes 32-63 for kb, each half probing 32 evenly spaced positions per step
// (a 32-ary search: 4 dependent loads for half a million entries where a
// binary search takes 19). In unsorted input the results are some positions
// in [lo, hi]. Every lane of the wave calls it.
__device__ __forceinline__ void wave_lower_bounds(const int *__restrict__ cols1, int lo, int hi,
                                                  int ka, int kb, int &ra, int &rb) {
    const int lane = lane_id(), q = lane & 31;
    const bool upper = lane >= 32;
    const int key = upper ? kb : ka;
    long long l = lo, r = hi;  // the answer lies in [l, r]
    while (true) {
        const bool more = r - l > 32;
        if (!__ballot(more)) break;
        const long long step = more ? (r - l + 31) / 32 : 0;
        const long long pos = l + q * step;
        const bool less = more && pos < r && cols1[pos] < key;
        const unsigned long long bal = __ballot(less);
        const int nl = __popcll(upper ? bal >> 32 : bal & 0xffffffffull);
        if (more) {
            const long long cand = l + nl * step;
            l = nl ? l + (nl - 1) * step + 1 : l;
            r = cand < r ? cand : r;
        }
    }
    const long long pos = l + q;
    const unsigned long long bal = __ballot(pos < r && cols1[pos] < key);
    const int res = (int)l + __popcll(upper ? bal >> 32 : bal & 0xffffffffull);
    ra = __shfl(res, 0, 64);
    rb = __shfl(res, 32, 64);
}

// Column extents and initial state (openph.cu:86-90 fills, kernels.cu:40-52
// compute_low), one wave per 64 columns: the wave finds where its columns'
// COO entries begin and end (wave_lower_bounds), then reads those entries
// 64 at a time and each column boundary among them gives the starts of the
// columns up to it. The COO start of a column goes to off2. The longest
// column reaches Ctrl::maxlen with one atomic per block: one per lane on
// that single address took 23 us of a Torus-100 setup.
// Columns [a, d.m) and their COO entries [e0, e1) (global 1-based indices).
// whole (a = 0, e0 = 0, nothing in the pool): column j's segment holds
// len + slot0 entries, so its offset is its COO start + j * slot0 and the
// layout needs no scan; its length is marked present in slotmap.
template <int CPW>
__global__ void __launch_bounds__(kBlock) k_init_cols(Dev d, const int *__restrict__ rows1,
                                                      const int *__restrict__ cols1, int a,
                                                      int e0, int e1, int whole) {
    constexpr int NW = kBlock / 64;
    constexpr int U = 4;        // windows of 64 entries read at once
    constexpr int Q = CPW / 64;  // columns per lane
    __shared__ int s_start[NW][CPW + 1];
    __shared__ int s_max;
    const int lane = lane_id(), wid = threadIdx.x >> 6;
    const long long jw = a + (((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6) * CPW;
    const int key0 = jw < d.m ? (int)jw + 1 : d.m + 1;             // the wave's first column, 1-based
    const int keyE = jw + CPW < d.m ? (int)jw + CPW + 1 : d.m + 1;  // one past its last
    if (threadIdx.x == 0) s_max = 0;
    int lo0, loE;
    wave_lower_bounds(cols1, e0, e1, key0, keyE, lo0, loE);
    if (loE < lo0) loE = lo0;  // unsorted input; flagged by k_entries
    for (int i = lane; i <= CPW; i += 64) s_start[wid][i] = loE;
    // column c1 > c0 at entry e: the columns (c0, c1] of the wave start at e
    const int nwin = (loE - lo0 + 63) >> 6;
    for (int w0 = 0; w0 < nwin; w0 += U) {
        int c[U], pc[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int e = lo0 + ((w0 + u) << 6) + lane;
            const int ec = e < loE ? e : loE - 1;
            c[u] = cols1[ec];
            pc[u] = cols1[ec > lo0 ? ec - 1 : lo0];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int e = lo0 + ((w0 + u) << 6) + lane;
            if (e < loE) {
                const int from = e > lo0 && pc[u] + 1 > key0 ? pc[u] + 1 : key0;
                const int to = c[u] < keyE - 1 ? c[u] : keyE - 1;
                for (int k = from; k <= to; k++) s_start[wid][k - key0] = e;
            }
        }
    }
    __syncthreads();
    int lo[Q], hi[Q], n[Q], l[Q];
    int wmax = 0;
    unsigned long long present = 0;  // lengths below 64 of the lane's columns
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const long long jl = jw + q * 64 + lane;
        lo[q] = s_start[wid][q * 64 + lane];
        hi[q] = s_start[wid][q * 64 + lane + 1];
        if (hi[q] < lo[q]) hi[q] = lo[q];  // unsorted input; flagged by k_entries
        n[q] = jl < d.m ? hi[q] - lo[q] : 0;
        wmax = n[q] > wmax ? n[q] : wmax;
        if (jl < d.m && n[q] < 64) present |= 1ull << n[q];
        // the last row: the initial low (compute_low, kernels.cu:40-52)
        l[q] = n[q] ? rows1[hi[q] - 1] - 1 : -1;
        if (l[q] < -1 || l[q] >= d.m) l[q] = -1;  // out of range; flagged by k_entries
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int x = __shfl_xor(wmax, o, 64);
        wmax = x > wmax ? x : wmax;
        present |= __shfl_xor(present, o, 64);
    }
    if (lane == 0 && wmax) atomicMax(&s_max, wmax);
    // the lengths present, one store per length and wave instead of one per
    // column (all columns of a complex share a few lengths)
    if (whole && lane < 64 && (present >> lane & 1) && lane <= d.m) d.slotmap[lane] = 1;
    __syncthreads();
    if (threadIdx.x == 0 && s_max) atomicMax(&d.ctrl->maxlen, s_max);
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const long long jl = jw + q * 64 + lane;
        if (jl >= d.m) break;
        const int j = (int)jl;
        d.off2[j] = (unsigned)lo[q];
        d.len[j] = n[q];
        if (whole) {
            d.off[j] = (unsigned)((unsigned long long)lo[q] + (unsigned long long)j * d.slot0);
            d.cap[j] = n[q] + d.slot0;
            if (n[q] >= 64 && n[q] <= d.m) d.slotmap[n[q]] = 1;  // (beyond m only in malformed input)
            if (j == 0)
                d.ctrl->pool_top = (unsigned long long)e1 + (unsigned long long)d.m * d.slot0;
        }
        d.low[j] = l[q];
        d.arglow[j] = -1;
        d.classes[j] = 0;
        d.ess[j] = 1;
        d.left[j] = kIntMax;
        d.beta[j] = -1;
        d.first[j] = kIntMax;
        d.clear[j] = 0;
        d.newpiv[j] = -1;
        d.pivdesc[j] = make_int4(-1, 0, 0, 0);
    }
}

// Columns per wave: 256 for large matrices (one search per 256 columns; 4x
// fewer waves), 64 below a million columns (more waves in flight).
constexpr int kInitColsLarge = 1 << 20;

static int init_cols_blocks(int m, int cpw) {
    const long long waves = ((long long)m + cpw - 1) / cpw;
    return (int)((waves * 64 + kBlock - 1) / kBlock);
}

static void launch_init_cols(const Dev &d, const int *rows1, const int *cols1, int a, int e0,
                             int e1, int whole, hipStream_t s) {
    const int n = d.m - a;
    if (n >= kInitColsLarge)
        k_init_cols<256><<<init_cols_blocks(n, 256), kBlock, 0, s>>>(d, rows1, cols1, a, e0, e1,
                                                                      whole);
    else
        k_init_cols<64><<<init_cols_blocks(n, 64), kBlock, 0, s>>>(d, rows1, cols1, a, e0, e1,
                                                                    whole);
}

// Validation, pool copy (0-based) and left[] by atomicMin. A row occurs in
// many columns (an edge of a Vietoris-Rips complex in every triangle on it,
// a few hundred columns apart), and device-scope atomics run at some 20 G/s
// for the whole GPU, slower still on one address. So a block takes
// 1,024 x PER consecutive entries, the minimum column per row first in an
// LDS hash table, and issues one atomic per distinct row: 219 K atomics
// instead of 302 K (1,024-entry tiles) on Torus-100. Larger tiles cut the
// atomics further (4.6 M instead of 22.8 M on a batch of 512 VR 2-skeleta
// with 16,384-entry tiles), but on a small matrix their fewer blocks take
// longer than the atomics they save (Torus-100: 31 blocks, 38.5 against
// 31 us), so the tile grows with the input (PER entries per thread). A row
// that finds no slot within kProbe probes (cubical grids: few repeats, many
// distinct rows) takes its atomic at once.
constexpr int kEntriesBlock = 1024;
constexpr int kEntriesHash = 8192;  // slots (key and value: 64 KB of LDS)
constexpr int kProbe = 8;
constexpr long long kEntriesLarge = 4 << 20;  // entries from which PER = 16

// Entries [e0, e1), of columns [a, d.m): a row below a (another matrix of
// a batch) is out of range.
template <int PER>
__global__ void __launch_bounds__(kEntriesBlock) k_entries(Dev d, const int *__restrict__ rows1,
                                                          const int *__restrict__ cols1, int a,
                                                          int e0, int e1) {
    constexpr int H = kEntriesHash;
    constexpr int R = 4;  // entries whose loads are issued together
    __shared__ int hkey[H];
    __shared__ int hval[H];
    for (int i = threadIdx.x; i < H; i += kEntriesBlock) {
        hkey[i] = -1;
        hval[i] = kIntMax;
    }
    __syncthreads();
    int bad = 0;
    const long long tile0 = (long long)e0 + (long long)blockIdx.x * (kEntriesBlock * PER);
    for (int k0 = 0; k0 < PER; k0 += R) {
        int c[R], r[R], b[R];
#pragma unroll
        for (int u = 0; u < R; u++) {
            const long long el = tile0 + (long long)(k0 + u) * kEntriesBlock + threadIdx.x;
            b[u] = el < e1 ? 0 : -1;  // -1: no entry
            const int e = el < e1 ? (int)el : e1 - 1;  // coalesced
            c[u] = cols1[e] - 1;
            r[u] = rows1[e] - 1;
            const int pc = cols1[e > e0 ? e - 1 : e] - 1, pr = rows1[e > e0 ? e - 1 : e] - 1;
            if (b[u] == 0) {
                if (c[u] < a || c[u] >= d.m || r[u] < a || r[u] >= c[u]) b[u] |= kErrRange;
                if (e > e0) {
                    if (c[u] < pc) b[u] |= kErrOrder;
                    else if (c[u] == pc && pr >= r[u]) b[u] |= kErrRows;
                }
                bad |= b[u];
            }
        }
        int o2[R], ln[R], of[R];
#pragma unroll
        for (int u = 0; u < R; u++) {
            const int cc = b[u] & kErrRange ? a : c[u];  // a valid column for the loads
            o2[u] = (int)d.off2[cc];
            ln[u] = d.len[cc];
            of[u] = (int)d.off[cc];
        }
#pragma unroll
        for (int u = 0; u < R; u++) {
            if (b[u] & kErrRange) continue;  // (and no entry: b = -1)
            const int e = (int)(tile0 + (long long)(k0 + u) * kEntriesBlock + threadIdx.x);
            const int q = e - o2[u];
            if (q >= 0 && q < ln[u]) d.pool[(unsigned)of[u] + q] = r[u];
            unsigned h = ((unsigned)r[u] * 2654435761u) >> 19;  // 13 bits = H
            bool in = false;
            for (int p = 0; p < kProbe; p++) {
                const int old = atomicCAS(&hkey[h], -1, r[u]);
                if (old == -1 || old == r[u]) {
                    atomicMin(&hval[h], c[u]);
                    in = true;
                    break;
                }
                h = (h + 1) & (H - 1);
            }
            if (!in) atomicMin(&d.left[r[u]], c[u]);
        }
    }
    if (bad) atomicOr(&d.ctrl->err, bad);
    __syncthreads();
    for (int i = threadIdx.x; i < H; i += kEntriesBlock) {
        const int key = hkey[i];
        if (key > -1) atomicMin(&d.left[key], hval[i]);
    }
}
static_assert(kEntriesHash == 1 << 13, "k_entries hashes to 13 bits");

static void launch_entries(const Dev &d, const int *rows1, const int *cols1, int a, int e0,
                           int e1, hipStream_t s) {
    const long long n = (long long)e1 - e0;
    if (n <= 0) return;
    if (n >= kEntriesLarge) {
        constexpr long long T = kEntriesBlock * 16;
        k_entries<16><<<(int)((n + T - 1) / T), kEntriesBlock, 0, s>>>(d, rows1, cols1, a, e0, e1);
    } else {
        constexpr long long T = kEntriesBlock * 4;
        k_entries<4><<<(int)((n + T - 1) / T), kEntriesBlock, 0, s>>>(d, rows1, cols1, a, e0, e1);
    }
}

__global__ void k_beta(Dev d, int a) {
    int i = a + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.m) return;
    int l = d.left[i];
    d.pivdesc[i].w = l;
    if (l != kIntMax) atomicMax(&d.beta[l], i);
}

void launch_setup(const Dev &d, const int *rows1, const int *cols1, hipStream_t s) {
    if (d.m > 0) launch_init_cols(d, rows1, cols1, 0, 0, d.nnz, 1, s);
    launch_entries(d, rows1, cols1, 0, 0, d.nnz, s);
}

// After the set-up of the whole matrix, one pass:
// * when the set-up found the input malformed, every column is made empty
//   (length 0, no low) and the loop halted, so that the kernels queued
//   behind it, which index by column lengths and lows, touch nothing out of
//   range (lows are in range already: k_init_cols clamps them); the host
//   reports the error at its next read of the control block;
// * else beta (beta.cu:2-15) from left, the left rule into pivdesc, and in
//   block 0 the present lengths numbered in increasing order: slot of a
//   pseudo-dimension (dimensions.cu:61-63, pseudo-dim = len - 1). (Finding
//   beta[j] among column j's own entries, which holds every i with
//   left[i] = j, instead of an atomic per column took longer: the gathers of
//   left[] cost more than the atomics, 10 against 3.4 us on Torus-100.)
__global__ void __launch_bounds__(kBlock) k_settle(Dev d) {
    __shared__ int lds[kBlock / 64];
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (d.ctrl->err) {
        if (j == 0) {
            d.ctrl->maxlen = 0;
            d.ctrl->halt = kHaltError;
        }
        if (j < d.m) {
            d.len[j] = 0;
            d.low[j] = -1;
        }
        return;
    }
    if (blockIdx.x == 0) {
        const int mx = d.ctrl->maxlen;
        int run = 0;
        for (int i0 = 0; i0 <= mx; i0 += kBlock) {
            const int i = i0 + threadIdx.x;
            const int present = i <= mx && d.slotmap[i] ? 1 : 0;
            int tot;
            const int ex = block_exclusive_scan(present, SumOp<int>(), 0, lds, &tot);
            if (i <= mx) d.slotmap[i] = present ? run + ex : -1;
            run += tot;
        }
        if (threadIdx.x == 0) {
            d.ctrl->nslots = run;
            if (run > kMaxSlots) atomicOr(&d.ctrl->err, kErrSlots);
        }
    }
    if (j < d.m) {
        const int l = d.left[j];
        d.pivdesc[j].w = l;
        if (l != kIntMax) atomicMax(&d.beta[l], j);
    }
}

void launch_settle(const Dev &d, hipStream_t s) {
    k_settle<<<d.m > 0 ? (d.m + kBlock - 1) / kBlock : 1, kBlock, 0, s>>>(d);
}

void launch_setup_range(const Dev &d, int a, int e0, int e1, const int *rows1, const int *cols1,
                        hipStream_t s) {
    if (d.m > a)
        launch_init_cols(d, rows1, cols1, a, e0, e1, 0, s);
    launch_scan(d, 0, d.off, s, a, a > 0);
    launch_entries(d, rows1, cols1, a, e0, e1, s);
}

__global__ void k_assign_slot(Dev d, int a) {
    int j = a + blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d.m) return;
    int s = d.slotmap[d.len[j]];
    d.slot[j] = s < 0 ? 0 : (s >= kMaxSlots ? kMaxSlots - 1 : s);
}


// Columns [a, d.m) added to a running batch: left/beta and the slots from a
// slot map the host keeps (Ctrl::nslots set with it).
__global__ void k_set_nslots(Ctrl *c, int n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) c->nslots = n;
}

void launch_slots_range(const Dev &d, int a, int nslots, hipStream_t s) {
    const int gm = (d.m - a + kBlock - 1) / kBlock;
    if (gm <= 0) return;
    k_beta<<<gm, kBlock, 0, s>>>(d, a);
    k_assign_slot<<<gm, kBlock, 0, s>>>(d, a);
    k_set_nslots<<<1, 64, 0, s>>>(d.ctrl, nslots);
}

// Publishes column j as the pivot of row l.
__device__ __forceinline__ void publish_pivot(const Dev &d, int l, int j) {
    d.arglow[l] = j;
    d.pivdesc[l] = make_int4(j, (int)d.off[j], d.len[j], d.left[l]);
}

// ---------------------------------------------------------------------------
// Phase 0: mark_pivots_and_clear (pms_functions.cu:33-53)
// ---------------------------------------------------------------------------

// assign: also the column's slot, from the numbered lengths (k_settle)
__global__ void k_phase0(Dev d, int a, int assign) {
    int j = a + blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d.m) return;
    if (assign) {
        const int sl = d.slotmap[d.len[j]];
        d.slot[j] = sl < 0 ? 0 : (sl >= kMaxSlots ? kMaxSlots - 1 : sl);
    }
    int lj = d.low[j];
    if (lj > -1 && lj == d.beta[j]) {
        d.classes[j] = -1;
        publish_pivot(d, lj, j);
        d.clear[lj] = 1;
        d.ess[lj] = 0;
        d.ess[j] = 0;
    }
}

// clear_positives (pms_functions.cu:131-140), reset of the first-occurrence
// table and construction of the Phase II worklist: columns still unknown
// whose low has a pivot to their left.
// One tile of kTile columns per block, kItems consecutive columns per
// thread; the worklist gets the block's columns in column order with one
// atomic per block (the appends of all blocks go to one counter, and
// same-address atomics serialize).
__global__ void __launch_bounds__(kBlock) k_apply1(Dev d) {
    __shared__ int lds[kBlock / 64];
    __shared__ int sbase;
    const int base = blockIdx.x * kTile + threadIdx.x * kItems;
    int cl[kItems], lo[kItems], cr[kItems];
    const int halt = d.ctrl->halt;
    load_run(d.classes, base, d.m, 1, cl);
    load_run(d.low, base, d.m, -1, lo);
    load_run(d.clear, base, d.m, 0, cr);
    if (issued(lo, halt)) return;  // uniform over the grid: no thread reaches the barrier
    int lt[kItems];
#pragma unroll
    for (int k = 0; k < kItems; k++) lt[k] = cr[k] || cl[k] != 0 ? -1 : lo[k];
    int p[kItems];
    gather_run(d.arglow, lt, -1, p);
    if (base + kItems <= d.m) {
        *reinterpret_cast<int4 *>(d.first + base) = make_int4(kIntMax, kIntMax, kIntMax, kIntMax);
    } else {
#pragma unroll
        for (int k = 0; k < kItems; k++)
            if (base + k < d.m) d.first[base + k] = kIntMax;
    }
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        const int t = base + k;
        if (t < d.m && cr[k]) {
            d.low[t] = -1;
            d.classes[t] = 1;
            d.clear[t] = 0;
        }
    }
    bool take[kItems];
    int cnt = 0;
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        take[k] = p[k] > -1 && p[k] < base + k;
        cnt += take[k] ? 1 : 0;
    }
    int tot;
    const int ex = block_exclusive_scan(cnt, SumOp<int>(), 0, lds, &tot);
    if (threadIdx.x == 0) sbase = tot ? atomicAdd(&d.ctrl->nwork, tot) : 0;
    __syncthreads();
    int o = sbase + ex;
#pragma unroll
    for (int k = 0; k < kItems; k++)
        if (take[k]) d.work[o++] = base + k;
}

// Phase 0; the first iteration's k_first applies its clears (first[] is
// reset by k_init_cols).
void launch_phase0(const Dev &d, hipStream_t s) {
    int gm = (d.m + kBlock - 1) / kBlock;
    if (!d.m) return;
    k_phase0<<<gm, kBlock, 0, s>>>(d, 0, 1);
}

// Phase 0 of columns [a, d.m) added to a running batch, its clears and the
// reset of first[] (the rest of k_apply1: the next iteration builds the
// worklist).
__global__ void k_clear_range(Dev d, int a) {
    const int t = a + blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.m) return;
    if (d.clear[t]) {
        d.low[t] = -1;
        d.classes[t] = 1;
        d.clear[t] = 0;
    }
    d.first[t] = kIntMax;
}

void launch_phase0_range(const Dev &d, int a, hipStream_t s) {
    const int gm = (d.m - a + kBlock - 1) / kBlock;
    if (gm <= 0) return;
    k_phase0<<<gm, kBlock, 0, s>>>(d, a, 0);
    k_clear_range<<<gm, kBlock, 0, s>>>(d, a);
}

void launch_apply1(const Dev &d, hipStream_t s) {
    if (d.ntiles) k_apply1<<<d.ntiles, kBlock, 0, s>>>(d);
}

// ---------------------------------------------------------------------------
// Iteration control (pms.cu:42-114). The loop's decisions are taken on the
// device, so that the host can queue several iterations, or replay them from
// a graph, and read the control block once at the end. Every kernel of an
// iteration returns at once while Ctrl::halt is set.
// ---------------------------------------------------------------------------

__device__ __forceinline__ void iter_begin(Ctrl *c) {
    c->iter += 1;
    c->any_dup = 0;
    c->overflow = 0;
    c->suspended = 0;
    c->nwork = 0;
    c->nlong = 0;
}

__global__ void k_iter_begin(Ctrl *c) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || c->halt) return;
    iter_begin(c);
}

// The end of an iteration, after Phase II: fold the counters and set
// Ctrl::halt. While lows repeat, pms marks a column or adds one in every
// iteration (the first occurrence of the largest repeated low is marked, or
// its repeats are reduced), so an iteration without either is a fault unless
// the pool ran out; two overflowing passes in a row without progress (the
// second right after a compaction) mean the pool budget cannot hold the work.
// Run by one thread of k_apply2: it reads only what the earlier kernels of
// the iteration wrote, and k_apply2 writes none of it.
__device__ void iter_end(Ctrl *c) {
    fold_stripes(c);  // also after a halt: the counters of a pass that ran
    if (c->halt) return;
    if (c->err) {
        c->halt = kHaltError;
        return;
    }
    const bool progress = c->additions != c->prev_add || c->marks != c->prev_marks;
    if (c->pool_top > c->pool_peak) c->pool_peak = c->pool_top;
    if (!c->overflow && !progress && c->any_dup && !c->suspended) {
        c->halt = kHaltStuck;
        return;
    }
    c->stalls = (c->overflow && !progress) ? c->stalls + 1 : 0;
    if (c->stalls >= 2) {
        c->halt = kHaltNoMem;
        return;
    }
    c->prev_add = c->additions;
    c->prev_marks = c->marks;
    if (c->overflow) c->overflows += 1;
    if (c->overflow || c->pool_top * 4 > c->pool_cap * 3) c->halt = kHaltCompact;
}

// A Phase II pass that ran out of pool is suspended instead of ended: the
// host compacts the pool and runs the same pass again over the same worklist
// (each column goes on from the state it committed; the pivots it reads are
// still those of the iteration's start, since k_apply2 published nothing),
// then k_apply2 ends the iteration. The iteration then does what it would
// have done with a larger pool: the same additions, the same iteration count.
// Two suspensions in a row without progress mean the pool budget cannot hold
// the work. Run by one thread of k_apply2, which otherwise does nothing while
// Ctrl::overflow is set.
__device__ void iter_suspend(Ctrl *c) {
    fold_stripes(c);
    if (c->halt) return;
    if (c->err) {
        c->halt = kHaltError;
        return;
    }
    const bool progress = c->additions != c->prev_add || c->marks != c->prev_marks;
    c->stalls = progress ? 0 : c->stalls + 1;
    if (c->stalls >= 2) {
        c->halt = kHaltNoMem;
        return;
    }
    c->prev_add = c->additions;
    c->prev_marks = c->marks;
    c->overflows += 1;
    c->suspended = 1;
    if (c->pool_top > c->pool_peak) c->pool_peak = c->pool_top;
    c->halt = kHaltSuspend;
}

__global__ void k_resume2(Ctrl *c) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    c->halt = kHaltNone;
    c->overflow = 0;
    c->nlong = 0;
}

void launch_resume2(const Dev &d, hipStream_t s) { k_resume2<<<1, 64, 0, s>>>(d.ctrl); }

__global__ void k_set_halt(Ctrl *c, int value) {
    if (threadIdx.x == 0 && blockIdx.x == 0) c->halt = value;
}

// Device clock at a point of the current iteration (which = 0 before
// Phase II, 1 after it) into a ring of `cap` entries; used to time Phase II
// inside graphs, which this runtime does not let record events once another
// library in the process has initialised it.
__global__ void k_stamp(Ctrl *c, unsigned long long *stamps, int which, int cap) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || c->halt) return;
    stamps[(2 * (c->iter - 1) + which) % cap] = (unsigned long long)wall_clock64();
}

void launch_stamp(const Dev &d, unsigned long long *stamps, int which, int cap, hipStream_t s) {
    k_stamp<<<1, 64, 0, s>>>(d.ctrl, stamps, which, cap);
}

void launch_iter_begin(const Dev &d, hipStream_t s) { k_iter_begin<<<1, 64, 0, s>>>(d.ctrl); }
void launch_set_halt(const Dev &d, int value, hipStream_t s) {
    k_set_halt<<<1, 64, 0, s>>>(d.ctrl, value);
}

// ---------------------------------------------------------------------------
// Phase I: lowstar detection (pms_functions.cu:55-85) as a segmented scan.
//
// For the columns of one pseudo-dimension in index order, column j with
// L = low[j] >= 0 is a duplicate when an earlier column has the same low
// (first[L] < j); ceil(j) = max of the duplicate lows before j in the same
// pseudo-dimension (-1 if none). j is marked when it is the first occurrence
// of L, classes[j] == 0 and L > ceil(j). Rows of different pseudo-dimensions
// are disjoint, so first[] can be global.
//
// first[] needs atomics only for unknown columns whose low has no pivot yet:
// a final pivot p of row L is the first column with low L (no column left of
// p can ever reach low L: p is the first column whose span holds a vector
// with low L), and every other column with low L is therefore a duplicate.
// This keeps the many columns that share a pivot's low off the atomic.
// ---------------------------------------------------------------------------

// Columns that share a low tend to be neighbours in filtration order (VR
// simplices born at the same diameter, cubical cells of one vertex), and
// device-scope atomics on one address serialize. Each block therefore first
// takes the minimum per distinct low of its tile in an LDS hash table and
// then issues one global atomicMin per distinct low.
constexpr int kHash = 2 * kTile;  // open addressing, load factor <= 1/2

// The first occurrences of lows for Phase I: columns base..base+kItems-1
// with lo[k] >= 0 and no pivot for it (p[k] = arglow[lo[k]] < 0) take the
// minimum column per low in the block's LDS table, then one atomicMin per
// distinct low. Every thread of the block calls both, with a block barrier
// between the init and the insert.
__device__ __forceinline__ void first_hash_init(int *hkey, int *hval) {
    for (int i = threadIdx.x; i < kHash; i += kBlock) {
        hkey[i] = -1;
        hval[i] = kIntMax;
    }
}

__device__ __forceinline__ void first_hash(const Dev &d, int base, const int (&lo)[kItems],
                                           const int (&p)[kItems], int *hkey, int *hval) {
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        const int lj = lo[k];
        if (lj < 0 || p[k] > -1) continue;
        unsigned h = ((unsigned)lj * 2654435761u) >> 21;  // 11 bits = kHash
        while (true) {
            const int old = atomicCAS(&hkey[h], -1, lj);
            if (old == -1 || old == lj) break;
            h = (h + 1) & (kHash - 1);
        }
        atomicMin(&hval[h], base + k);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kHash; i += kBlock) {
        const int key = hkey[i];
        if (key > -1) atomicMin(&d.first[key], hval[i]);
    }
}

__global__ void __launch_bounds__(kBlock) k_first(Dev d) {
    __shared__ int hkey[kHash];
    __shared__ int hval[kHash];
    const int base = blockIdx.x * kTile + threadIdx.x * kItems;
    int lo[kItems], cl[kItems];
    int cr[kItems];
    const int halt = d.ctrl->halt;
    load_run(d.low, base, d.m, -1, lo);
    load_run(d.classes, base, d.m, 1, cl);
    load_run(d.clear, base, d.m, 0, cr);
    if (issued(lo, halt)) return;
    // Phase 0's clears (clear_positives, pms_functions.cu:131-140): pending
    // only in the first iteration, every later clear is applied in the
    // iteration that sets it
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        if (cr[k]) {
            const int j = base + k;
            d.low[j] = -1;
            d.classes[j] = 1;
            d.clear[j] = 0;
        }
    }
    // only unknown columns whose low has no pivot yet take part
#pragma unroll
    for (int k = 0; k < kItems; k++)
        if (cl[k] != 0 || cr[k]) lo[k] = -1;
    int p[kItems];
    gather_run(d.arglow, lo, 0, p);
    // the iteration starts here: none of this kernel's blocks reads the
    // flags reset below
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        iter_begin(d.ctrl);
        if (d.its) d.its[2 * (d.ctrl->iter % kIterRing)] = (unsigned long long)wall_clock64();
    }
    if (blockIdx.x == 0)  // k_tile marks the members whose lows repeat
        for (int k = threadIdx.x; k < d.nmem; k += kBlock) d.mdup[k] = 0;
    first_hash_init(hkey, hval);
    __syncthreads();
    first_hash(d, base, lo, p, hkey, hval);
}
static_assert(kHash == 2048, "k_first hashes to 11 bits");

// dup = L when column j repeats an earlier column's low L, else -1;
// fst = 1 when j is the first column with its low. p = arglow[L] and
// f = first[L], loaded together (f matters only when L has no pivot).
__device__ __forceinline__ void dup_state(int j, int lj, int p, int f, int &dup, int &fst) {
    dup = -1;
    fst = 0;
    if (lj < 0) return;
    if (p > -1) {
        if (p == j) fst = 1;
        else dup = lj;
        return;
    }
    if (f == j) fst = 1;
    else if (f < j) dup = lj;
}

__device__ __forceinline__ int segment_of(const int *off, int count, int x) {
    // largest k with off[k] <= x (off has count+1 ascending entries)
    int lo = 0, hi = count;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (off[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Per tile of kTile columns: max duplicate low per slot, and the global
// "some low occurs twice" flag (is_reduced, pms_functions.cu:7-27).
__global__ void __launch_bounds__(kBlock) k_tile(Dev d) {
    __shared__ int smax[kMaxSlots];
    __shared__ int sdup;
    const int base = blockIdx.x * kTile + threadIdx.x * kItems;
    int lo[kItems], sl[kItems];
    const int halt = d.ctrl->halt;
    load_run(d.low, base, d.m, -1, lo);
    load_run(d.slot, base, d.m, 0, sl);
    if (issued(lo, halt)) return;
    int p[kItems], f[kItems];
    gather_run(d.arglow, lo, -1, p);
    gather_run(d.first, lo, kIntMax, f);
    if (threadIdx.x < kMaxSlots) smax[threadIdx.x] = -1;
    if (threadIdx.x == 0) sdup = 0;
    __syncthreads();
    int dupseen = 0;
    int mk = -1;  // member of the last repeating column (running batches)
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        const int j = base + k;
        int dup, fst;
        dup_state(j, lo[k], p[k], f[k], dup, fst);
        if (dup > -1) {
            atomicMax(&smax[sl[k]], dup);
            dupseen = 1;
            if (d.nmem) {
                if (mk < 0 || j >= d.mstart[mk + 1]) mk = segment_of(d.mstart, d.nmem, j);
                d.mdup[mk] = 1;
            }
        }
    }
    if (dupseen) sdup = 1;
    __syncthreads();
    if (threadIdx.x < kMaxSlots) d.tile_max[blockIdx.x * kMaxSlots + threadIdx.x] = smax[threadIdx.x];
    if (threadIdx.x == 0 && sdup) atomicOr(&d.ctrl->any_dup, 1);
}

// Convergence test (is_reduced, pms_functions.cu:7-27: no non-negative low
// occurs twice, from k_tile's flag; the first iteration always runs, as in
// pms.cu:39-44), the iteration cap, and the exclusive max-scan of the tile
// maxima over tiles, per slot. Block s scans slot s: each of its 1024
// threads takes a contiguous run of tiles (independent loads), the block
// scans the runs' maxima and each thread writes its run's prefixes. Every
// block takes the loop's exit decision from the same inputs; block 0
// writes it.
__global__ void __launch_bounds__(1024) k_tilescan(Dev d) {
    __shared__ int s_halt;
    __shared__ int lds[1024 / 64];
    if (threadIdx.x == 0) {
        const Ctrl *c = d.ctrl;
        int h = c->halt;
        if (!h) {
            if (c->err) h = kHaltError;
            else if (c->iter > 1 && !c->any_dup) h = kHaltConverged;
            else if (c->iter > c->max_iters) h = kHaltIters;
            if (h && blockIdx.x == 0) d.ctrl->halt = h;
        }
        s_halt = h;
    }
    __syncthreads();
    if (s_halt) return;
    const int nslots = d.ctrl->nslots;
    const int s = blockIdx.x;
    if (s >= nslots || s >= kMaxSlots) return;
    const int per = (d.ntiles + 1023) / 1024;
    const int t0 = threadIdx.x * per < d.ntiles ? threadIdx.x * per : d.ntiles;
    const int t1 = t0 + per < d.ntiles ? t0 + per : d.ntiles;
    const int *tm = d.tile_max + s;
    int run = -1;
#pragma unroll 4
    for (int t = t0; t < t1; t++) {
        const int v = tm[t * kMaxSlots];
        run = v > run ? v : run;
    }
    int tot;
    int pre = block_exclusive_scan(run, MaxOp<int>(), -1, lds, &tot);
#pragma unroll 4
    for (int t = t0; t < t1; t++) {
        d.tile_pre[t * kMaxSlots + s] = pre;
        const int v = tm[t * kMaxSlots];
        pre = v > pre ? v : pre;
    }
}

// Up to this many tiles each k_mark block reads the earlier tiles' maxima
// itself (at most 2 tiles per thread) instead of a k_tilescan pass.
constexpr int kFuseTiles = 512;
static_assert(kFuseTiles <= 2 * kBlock, "two earlier tiles per thread");
static_assert(kMaxSlots % 4 == 0, "a tile's slot maxima as int4");

// The 16 slot maxima of tile t (64 contiguous bytes), or -1 when !valid.
__device__ __forceinline__ void load_tile_slots(const int *tm, int t, bool valid,
                                                int (&v)[kMaxSlots]) {
    const int4 *q = reinterpret_cast<const int4 *>(tm + (valid ? t : 0) * kMaxSlots);
#pragma unroll
    for (int i = 0; i < kMaxSlots / 4; i++) {
        const int4 x = q[i];
        v[4 * i] = valid ? x.x : -1;
        v[4 * i + 1] = valid ? x.y : -1;
        v[4 * i + 2] = valid ? x.z : -1;
        v[4 * i + 3] = valid ? x.w : -1;
    }
}

// FUSED (few tiles): k_tilescan's work done here, by every block from the
// same inputs: the loop's exit (block 0 writes it) and the maximum over the
// earlier tiles, read by each block itself. One launch less per iteration.
// All slots are scanned together: one wave scan per slot in registers and a
// single block barrier for the lot.
template <bool FUSED>
__global__ void __launch_bounds__(kBlock) k_mark(Dev d) {
    constexpr int NW = kBlock / 64;
    __shared__ int exl[kMaxSlots][kBlock];
    __shared__ int s_agg[kMaxSlots][NW];
    __shared__ int s_tm[kMaxSlots][NW];
    __shared__ int s_cnt[NW];
    const int base = blockIdx.x * kTile + threadIdx.x * kItems;
    const int lane = lane_id(), wid = threadIdx.x >> 6;
    int lo[kItems], sl[kItems], cl[kItems];
    const Ctrl *c = d.ctrl;
    const int halt0 = c->halt, err = c->err, iter = c->iter, any_dup = c->any_dup;
    const int max_iters = c->max_iters, nslots = c->nslots;
    load_run(d.low, base, d.m, -1, lo);
    load_run(d.slot, base, d.m, 0, sl);
    load_run(d.classes, base, d.m, 1, cl);
    int h = issued(lo, halt0);
    if (FUSED) {
        if (!h) {
            if (err) h = kHaltError;
            else if (iter > 1 && !any_dup) h = kHaltConverged;
            else if (iter > max_iters) h = kHaltIters;
        }
        if (h) {
            if (blockIdx.x == 0 && threadIdx.x == 0) d.ctrl->halt = h;
            return;
        }
    } else if (h) {
        return;
    }
    int p[kItems], f[kItems], lf[kItems], of[kItems], ln[kItems];
    gather_run(d.arglow, lo, -1, p);
    gather_run(d.first, lo, kIntMax, f);
    // FUSED (short passes, bound by their chain of loads): what a mark
    // publishes (pivdesc: the column's place and the row's left rule),
    // loaded with the rest instead of after the scans; in the long passes
    // the extra loads would cost more bandwidth than the latency they hide
    if (FUSED) {
        gather_run(d.left, lo, 0, lf);
        load_run(reinterpret_cast<const int *>(d.off), base, d.m, 0, of);
        load_run(d.len, base, d.m, 0, ln);
    }
    // per slot, the largest duplicate low of the earlier tiles: FUSED, the
    // maximum over this thread's share of them (reduced over the block
    // below); else k_tilescan's prefix for this tile
    int tm[kMaxSlots];
    if (FUSED) {
        const int t0 = threadIdx.x, t1 = threadIdx.x + kBlock;
        int v1[kMaxSlots];
        load_tile_slots(d.tile_max, t0, t0 < (int)blockIdx.x, tm);
        load_tile_slots(d.tile_max, t1, t1 < (int)blockIdx.x, v1);
#pragma unroll
        for (int s = 0; s < kMaxSlots; s++) tm[s] = v1[s] > tm[s] ? v1[s] : tm[s];
    } else {
        load_tile_slots(d.tile_pre, blockIdx.x, true, tm);
    }
    int dup[kItems], fst[kItems];
#pragma unroll
    for (int k = 0; k < kItems; k++) dup_state(base + k, lo[k], p[k], f[k], dup[k], fst[k]);
    // exclusive max scan of the duplicate lows per slot over the block
#pragma unroll
    for (int s = 0; s < kMaxSlots; s++) {
        if (s < nslots) {
            int agg = -1;
#pragma unroll
            for (int k = 0; k < kItems; k++)
                if (sl[k] == s && dup[k] > agg) agg = dup[k];
            const int incl = wave_inclusive_scan(agg, MaxOp<int>(), -1);
            exl[s][threadIdx.x] = dpp_or<0x138, 0xF>(-1, incl);  // wave_shr:1
            const int wtot = __builtin_amdgcn_readlane(incl, 63);
            int wtm = -1;
            if (FUSED)
                wtm = __builtin_amdgcn_readlane(wave_inclusive_scan(tm[s], MaxOp<int>(), -1), 63);
            if (lane == 0) {
                s_agg[s][wid] = wtot;
                s_tm[s][wid] = wtm;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kMaxSlots; s++) {
        if (s < nslots) {
            int pre = exl[s][threadIdx.x];
            int tp = FUSED ? -1 : tm[s];
#pragma unroll
            for (int w = 0; w < NW; w++) {
                const int x = s_agg[s][w];
                if (w < wid && x > pre) pre = x;
                if (FUSED && s_tm[s][w] > tp) tp = s_tm[s][w];
            }
            exl[s][threadIdx.x] = pre > tp ? pre : tp;
        }
    }
    int marked = 0;
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        const int j = base + k;
        if (lo[k] < 0 || !fst[k]) continue;
        int c = exl[sl[k]][threadIdx.x];
#pragma unroll
        for (int q = 0; q < k; q++)
            if (sl[q] == sl[k] && dup[q] > c) c = dup[q];
        if (cl[k] == 0 && lo[k] > c) {
            const int L = lo[k];
            marked++;
            if (FUSED) {
                d.arglow[L] = j;
                d.pivdesc[L] = make_int4(j, of[k], ln[k], lf[k]);
            } else {
                publish_pivot(d, L, j);
            }
            d.classes[j] = -1;
            d.clear[L] = 1;
            d.ess[L] = 0;
            d.ess[j] = 0;
        }
    }
    // one atomic per block on the shared counter
    const int wm = __builtin_amdgcn_readlane(wave_inclusive_scan(marked, SumOp<int>(), 0), 63);
    if (lane == 0) s_cnt[wid] = wm;
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) tot += s_cnt[w];
        if (tot) atomicAdd(&d.ctrl->marks, tot);
    }
}

void launch_phase1_detect(const Dev &d, hipStream_t s) {
    if (!d.fuse_first) k_first<<<d.ntiles, kBlock, 0, s>>>(d);
    k_tile<<<d.ntiles, kBlock, 0, s>>>(d);
}

void launch_phase1_mark(const Dev &d, hipStream_t s) {
    if (d.ntiles <= kFuseTiles) {
        k_mark<true><<<d.ntiles, kBlock, 0, s>>>(d);
    } else {
        k_tilescan<<<kMaxSlots, 1024, 0, s>>>(d);
        k_mark<false><<<d.ntiles, kBlock, 0, s>>>(d);
    }
}

// ---------------------------------------------------------------------------
// Phase II (pms_functions.cu:87-112): reduce every unfinished column by the
// pivots known at the start of the phase.
//
// Pivots found here by the left rule (left[low_j] == j, the column is final)
// go to newpiv[] and are published after the phase, so every pivdesc[] entry
// read here names a column that is final and not written during the phase.
// clear[] / ess[] writes only ever store 1 / 0 and are facts about the matrix
// (a non-negative low of a partially reduced column is a positive, paired
// index), so their order does not matter. A column's own state (off, len,
// cap, low, classes, newpiv) changes only once its data is in the pool.
// ---------------------------------------------------------------------------

// c = a xor b for ascending lists; returns |c|.
__device__ __forceinline__ int merge_xor(const int *__restrict__ a, int na,
                                         const int *__restrict__ b, int nb,
                                         int *__restrict__ c) {
    int i = 0, k = 0, n = 0;
    while (i < na && k < nb) {
        int x = a[i], y = b[k];
        if (x < y) { c[n++] = x; i++; }
        else if (y < x) { c[n++] = y; k++; }
        else { i++; k++; }
    }
    while (i < na) c[n++] = a[i++];
    while (k < nb) c[n++] = b[k++];
    return n;
}

// One lane per column over all m columns, every addition into a fresh pool
// segment. Kept as the lanes = 1 configuration of the autotuner.
__global__ void __launch_bounds__(kBlock) k_phase2_lane(Dev d) {
    if (d.ctrl->halt) return;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    bool active = false;
    int lj = -1, p = -1;
    unsigned coff = 0;
    int clen = 0;
    if (j < d.m && d.classes[j] == 0) {
        lj = d.low[j];
        if (lj > -1) {
            p = d.arglow[lj];
            active = (p > -1 && p < j);
        }
    }
    if (active) {
        coff = d.off[j];
        clen = d.len[j];
    }
    bool dirty = false, ovf = false, becomes_pivot = false;
    unsigned long long nadd = 0, nel = 0;
    const unsigned long long cap = d.ctrl->pool_cap;
    // wave-uniform loop so that wave_alloc sees every lane
    while (__any(active)) {
        unsigned need = 0, aoff = 0;
        int alen = 0;
        if (active) {
            d.ess[lj] = 0;
            aoff = d.off[p];
            alen = d.len[p];
            need = (unsigned)(alen + clen);
        }
        unsigned long long base = wave_alloc(&d.ctrl->pool_top, need, active);
        if (active) {
            if (base + need > cap) {
                ovf = true;
                active = false;
            } else {
                int *dst = d.pool + base;
                int n = merge_xor(d.pool + aoff, alen, d.pool + coff, clen, dst);
                nadd += 1;
                nel += (unsigned long long)(alen + clen + n);
                coff = (unsigned)base;
                clen = n;
                dirty = true;
                int prev = lj;
                lj = n ? dst[n - 1] : -1;
                if (lj >= prev) {
                    atomicOr(&d.ctrl->err, kErrState);
                    active = false;
                } else if (lj < 0) {
                    active = false;
                } else {
                    d.clear[lj] = 1;
                    if (d.left[lj] == j) {
                        becomes_pivot = true;
                        active = false;
                    } else {
                        p = d.arglow[lj];
                        active = (p > -1 && p < j);
                    }
                }
            }
        }
    }
    if (dirty) {
        d.off[j] = coff;
        d.len[j] = clen;
        d.cap[j] = clen;
        d.low[j] = lj;
        if (becomes_pivot) {
            d.newpiv[lj] = j;
            d.classes[j] = -1;
        }
    }
    if (ovf) d.ctrl->overflow = 1;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        nadd += __shfl_down(nadd, o, 64);
        nel += __shfl_down(nel, o, 64);
    }
    if (lane_id() == 0 && nadd) {
        atomicAdd(&d.ctrl->additions, nadd);
        atomicAdd(&d.ctrl->merge_elems, nel);
    }
}

// Lane-group merge. G lanes of one wave own one column; c = a xor b for
// ascending lists by merge path. Lane g takes the diagonals [g*K, (g+1)*K) of
// the merge of a and b (K = ceil(n/G), elements of a before equal elements of
// b), finds its start by one binary search along its first diagonal and then
// walks K steps reading one element per step. An element of a is dropped when
// the next element of b equals it; an element of b is dropped when the
// previous element of a equals it: both copies of a common element vanish
// wherever the diagonals split them. A first walk counts the survivors, an
// exclusive scan over the group gives each lane its output offset and a second
// walk writes them. The pointers may address LDS or global memory.

// number of a-elements among the first d elements of the merge
__device__ __forceinline__ int mp_split(const int *a, int na, const int *b, int nb, int d) {
    int lo = d - nb > 0 ? d - nb : 0, hi = d < na ? d : na;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] <= b[d - mid - 1]) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

template <bool WRITE>
__device__ __forceinline__ int mp_walk(const int *a, int na, const int *b, int nb, int d0, int d1,
                                       int *c) {
    int i = mp_split(a, na, b, nb, d0), k = d0 - i;
    int av = i < na ? a[i] : kIntMax;
    int bv = k < nb ? b[k] : kIntMax;
    int prev_a = i > 0 ? a[i - 1] : -1;
    int cnt = 0;
    for (int d = d0; d < d1; d++) {
        int v;
        bool keep;
        if (av <= bv) {
            v = av;
            keep = bv != av;
            prev_a = av;
            i++;
            av = i < na ? a[i] : kIntMax;
        } else {
            v = bv;
            keep = prev_a != bv;
            k++;
            bv = k < nb ? b[k] : kIntMax;
        }
        if (keep) {
            if (WRITE) c[cnt] = v;
            cnt++;
        }
    }
    return cnt;
}

// mp_walk for operands in LDS, without branches: one LDS read per step from
// whichever list the step consumed (no flat addressing: both are LDS). With
// SENT, a[na] and b[nb] hold kIntMax, so no read is guarded, and every step
// writes its element to c (a survivor's slot is the next one): the writes
// stay inside the lane's own stretch of diagonals.
template <bool SENT>
__device__ __forceinline__ int mp_walk_lds(const int *a, int na, const int *b, int nb, int d0,
                                           int d1, int *c) {
    int i = mp_split(a, na, b, nb, d0), k = d0 - i;
    int av, bv, prev_a;
    if (SENT) {
        av = a[i];
        bv = b[k];
        prev_a = a[i > 0 ? i - 1 : 0];
        prev_a = i > 0 ? prev_a : -1;
    } else {
        av = i < na ? a[i] : kIntMax;
        bv = k < nb ? b[k] : kIntMax;
        prev_a = i > 0 ? a[i - 1] : -1;
    }
    int cnt = 0;
    for (int d = d0; d < d1; d++) {
        const bool ta = av <= bv;
        const int v = ta ? av : bv;
        const bool keep = ta ? bv != av : prev_a != bv;
        prev_a = ta ? av : prev_a;
        i += ta ? 1 : 0;
        k += ta ? 0 : 1;
        int nv;
        if (SENT) {
            nv = (ta ? a + i : b + k)[0];
        } else {
            const int idx = ta ? i : k;
            nv = idx < (ta ? na : nb) ? (ta ? a : b)[idx] : kIntMax;
        }
        av = ta ? nv : av;
        bv = ta ? bv : nv;
        if (SENT) {
            c[cnt] = v;
        } else if (keep) {
            c[cnt] = v;
        }
        cnt += keep ? 1 : 0;
    }
    return cnt;
}

// mp_walk_lds<true> with one element of lookahead on each list, for long
// walks: each step issues the read the step after next may need, so the
// reads' latency overlaps the steps instead of adding to each. a[na] and
// b[nb] hold kIntMax; indices are clamped there.
__device__ __forceinline__ int mp_walk_lds_ahead(const int *a, int na, const int *b, int nb,
                                                 int d0, int d1, int *c) {
    int i = mp_split(a, na, b, nb, d0), k = d0 - i;
    int av = a[i], bv = b[k];
    int an = a[i + 1 < na ? i + 1 : na], bn = b[k + 1 < nb ? k + 1 : nb];
    int prev_a = a[i > 0 ? i - 1 : 0];
    prev_a = i > 0 ? prev_a : -1;
    int cnt = 0;
    for (int d = d0; d < d1; d++) {
        const bool ta = av <= bv;
        const int v = ta ? av : bv;
        const bool keep = ta ? bv != av : prev_a != bv;
        prev_a = ta ? av : prev_a;
        i += ta ? 1 : 0;
        k += ta ? 0 : 1;
        // the element after the new head of the list just consumed
        const int nx = ta ? a[i + 1 < na ? i + 1 : na] : b[k + 1 < nb ? k + 1 : nb];
        av = ta ? an : av;
        an = ta ? nx : an;
        bv = ta ? bv : bn;
        bn = ta ? bn : nx;
        c[cnt] = v;
        cnt += keep ? 1 : 0;
    }
    return cnt;
}

// A team merges one column: G lanes of one wave (WaveTeam) or a whole block
// (BlockTeam). It provides its lane index and size, a barrier, an exclusive
// scan with total, a broadcast from its leader and pool allocation.
template <int G>
struct WaveTeam {
    static constexpr int size = G;
    static constexpr bool wave = true;
    int gl, gw;
    __device__ int lane() const { return gl; }
    __device__ void sync() const { __syncwarp(); }
    // stores of the team to global memory visible to the team's later loads
    __device__ void fence() const { __threadfence_block(); }
    __device__ int scan(int v, int &total) const {
        int incl = v;
        if constexpr (G <= 16) {
            // a group lies inside one 16-lane row: row shifts (DPP) instead
            // of LDS permutes; a lane adds only what comes from its group
            incl += dpp_up<1>(incl, gl);
            if constexpr (G > 2) incl += dpp_up<2>(incl, gl);
            if constexpr (G > 4) incl += dpp_up<4>(incl, gl);
            if constexpr (G > 8) incl += dpp_up<8>(incl, gl);
        } else if constexpr (G == 64) {
            // the group is the wave: the register scan, and the total from
            // lane 63 into a scalar register
            incl = wave_inclusive_scan(v, [](int p, int q) { return p + q; }, 0);
            total = __builtin_amdgcn_readlane(incl, 63);
            return incl - v;
        } else if constexpr (G == 32) {
            // the register scan without the broadcast across the halves
            incl += dpp_or<0x111, 0xF>(0, incl);  // row_shr:1
            incl += dpp_or<0x112, 0xF>(0, incl);  // row_shr:2
            incl += dpp_or<0x114, 0xF>(0, incl);  // row_shr:4
            incl += dpp_or<0x118, 0xF>(0, incl);  // row_shr:8
            incl += dpp_or<0x142, 0xA>(0, incl);  // row_bcast:15 into rows 1 and 3
        } else {
#pragma unroll
            for (int o = 1; o < G; o <<= 1) {
                const int t = __shfl_up(incl, o, G);
                if (gl >= o) incl += t;
            }
        }
        total = __shfl(incl, G - 1, G);
        return incl - v;
    }
    __device__ unsigned long long bcast(unsigned long long x) const {
        return __shfl(x, gw * G, 64);
    }

  private:
    // x of the lane O below, 0 for the group's first O lanes (row_shr:O)
    template <int O>
    __device__ static int dpp_up(int x, int gl) {
        const int t = __builtin_amdgcn_update_dpp(0, x, 0x110 + O, 0xF, 0xF, false);
        return gl >= O ? t : 0;
    }
};

// The scan and the broadcast take one block barrier each: they alternate
// between two LDS buffers, and a buffer is written again only after the
// next call's barrier, which every wave reaches once done reading it.
template <int BT>
struct BlockTeam {
    static constexpr int size = BT;
    static constexpr bool wave = false;
    int *s_scan;                 // 2 * BT / 64 ints of LDS
    unsigned long long *s_word;  // two words of LDS
    mutable int ps = 0, pb = 0;  // the buffers' turns (the same on every thread)
    __device__ int lane() const { return threadIdx.x; }
    __device__ void sync() const { __syncthreads(); }
    __device__ void fence() const {
        __threadfence_block();
        __syncthreads();
    }
    __device__ int scan(int v, int &total) const {
        int *lds = s_scan + ps * (BT / 64);
        ps ^= 1;
        const int lane = lane_id(), wid = threadIdx.x >> 6;
        const int incl = wave_inclusive_scan(v, SumOp<int>(), 0);
        const int excl = dpp_or<0x138, 0xF>(0, incl);  // wave_shr:1
        if (lane == 63) lds[wid] = incl;
        __syncthreads();
        int pre = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < BT / 64; w++) {
            const int x = lds[w];
            pre += w < wid ? x : 0;
            tot += x;
        }
        total = tot;
        return pre + excl;
    }
    __device__ unsigned long long bcast(unsigned long long x) const {
        unsigned long long *w = s_word + pb;
        pb ^= 1;
        if (threadIdx.x == 0) *w = x;
        __syncthreads();
        return *w;
    }
};

template <class Team>
__device__ int team_xor(const Team &t, const int *a, int na, const int *b, int nb, int *c) {
    const int n = na + nb;
    const int K = (n + Team::size - 1) / Team::size;
    const int l = t.lane();
    const int d0 = l * K < n ? l * K : n;
    const int d1 = d0 + K < n ? d0 + K : n;
    const int cnt = mp_walk<false>(a, na, b, nb, d0, d1, nullptr);
    int total;
    const int ex = t.scan(cnt, total);
    mp_walk<true>(a, na, b, nb, d0, d1, c + ex);
    return total;
}

// The same sum with one walk, for operands in LDS: each lane writes its
// survivors to c at its first diagonal (a provisional slot, never beyond its
// own stretch of diagonals), and after the scan copies them, now without a
// dependency chain, to their final places in a's buffer, which the walks no
// longer need. The result ends in `a`.
template <class Team>
__device__ int team_xor_lds(const Team &t, int *a, int na, const int *b, int nb, int *c) {
    const int n = na + nb;
    const int K = (n + Team::size - 1) / Team::size;
    const int l = t.lane();
    const int d0 = l * K < n ? l * K : n;
    const int d1 = d0 + K < n ? d0 + K : n;
    const int cnt = mp_walk_lds<false>(a, na, b, nb, d0, d1, c + d0);
    int total;
    // every lane's walk is over after the scan (a block barrier for a
    // block team, the lanes' own wave for a wave team)
    const int ex = t.scan(cnt, total);
    // four reads, then four writes: one LDS round trip per four survivors
    for (int q = 0; q < cnt; q += 4) {
        const int x0 = c[d0 + q];
        const int x1 = q + 1 < cnt ? c[d0 + q + 1] : 0;
        const int x2 = q + 2 < cnt ? c[d0 + q + 2] : 0;
        const int x3 = q + 3 < cnt ? c[d0 + q + 3] : 0;
        a[ex + q] = x0;
        if (q + 1 < cnt) a[ex + q + 1] = x1;
        if (q + 2 < cnt) a[ex + q + 2] = x2;
        if (q + 3 < cnt) a[ex + q + 3] = x3;
    }
    return total;
}

// The same sum with the result in b's buffer, in two parts so that a caller
// can issue global loads between them. xor_walk: survivors go to provisional
// slots of c at their first diagonal, the team scan gives their places;
// returns the sum's length, with the lane's survivor count, first diagonal
// and output place in w. xor_copy (after xor_walk's scan every walk is over,
// so b is free to write): the survivors to their places in b, followed by a
// new sentinel. a may be in
// LDS (LDS_A, with its sentinel) or in the pool; b[nb] holds the sentinel
// kIntMax, so n + 1 entries of b's buffer are used.
struct XorWalk {
    int cnt, d0, ex, total;
};

template <bool LDS_A, class Team>
__device__ __forceinline__ XorWalk xor_walk(const Team &t, const int *a, int na, const int *b,
                                            int nb, int *c) {
    const int n = na + nb;
    const int K = (int)(((unsigned)n + Team::size - 1) / Team::size);
    const int l = t.lane();
    XorWalk w;
    w.d0 = l * K < n ? l * K : n;
    const int d1 = w.d0 + K < n ? w.d0 + K : n;
    if (!LDS_A) w.cnt = mp_walk<true>(a, na, b, nb, w.d0, d1, c + w.d0);
    else if constexpr (Team::size >= 32) w.cnt = mp_walk_lds_ahead(a, na, b, nb, w.d0, d1, c + w.d0);
    else w.cnt = mp_walk_lds<true>(a, na, b, nb, w.d0, d1, c + w.d0);
    w.ex = t.scan(w.cnt, w.total);
    return w;
}

// four reads, then four writes, without branches: reads past the lane's
// survivors are clamped to c[spare] and their writes go there (spare >= n is
// a slot of c that no walk uses)
__device__ __forceinline__ void xor_copy(const XorWalk &w, int *b, int *c, int spare) {
    for (int q = 0; q < w.cnt; q += 4) {
        const int s0 = w.d0 + q;
        const int x0 = c[s0];
        const int x1 = c[q + 1 < w.cnt ? s0 + 1 : spare];
        const int x2 = c[q + 2 < w.cnt ? s0 + 2 : spare];
        const int x3 = c[q + 3 < w.cnt ? s0 + 3 : spare];
        b[w.ex + q] = x0;
        (q + 1 < w.cnt ? b + w.ex + q + 1 : c + spare)[0] = x1;
        (q + 2 < w.cnt ? b + w.ex + q + 2 : c + spare)[0] = x2;
        (q + 3 < w.cnt ? b + w.ex + q + 3 : c + spare)[0] = x3;
    }
    b[w.total] = kIntMax;  // every lane: the same value
}

template <bool LDS_A, class Team>
__device__ int team_xor_into_b(const Team &t, const int *a, int na, int *b, int nb, int *c,
                               int spare) {
    // every walk is over after xor_walk's scan: b can be written
    const XorWalk w = xor_walk<LDS_A>(t, a, na, b, nb, c);
    xor_copy(w, b, c, spare);
    return w.total;
}

// The largest element of a xor b for ascending lists whose largest elements
// are equal (a column and the pivot of its low): the tops cancel, and so do
// the elements below them while both lists agree. -1 when the sum is empty.
// Every lane of a team reads the same entries and gets the same value.
__device__ __forceinline__ int xor_top_from(const int *a, int na, const int *b, int nb, int q) {
    int i = na - 2 - q, k = nb - 2 - q;
    while (true) {
        const int x = i >= 0 ? a[i] : -1;
        const int y = k >= 0 ? b[k] : -1;
        if (x != y) return x > y ? x : y;
        if (x < 0) return -1;
        i--;
        k--;
    }
}

// The same for a team of one whole wave: lane q compares the q-th pair below
// the tops, one LDS round trip for the first 64 pairs.
template <class Team>
__device__ __forceinline__ int xor_top(const Team &t, const int *a, int na, const int *b, int nb) {
    (void)t;
    if constexpr (Team::size >= 64) {
        // a whole wave, or every wave of a block team (each computes the
        // same value from the same entries)
        const int q = lane_id();
        const int i = na - 2 - q, k = nb - 2 - q;
        const int x = i >= 0 ? a[i] : -1;
        const int y = k >= 0 ? b[k] : -1;
        const unsigned long long ne = __ballot(x != y || x < 0);
        if (!ne) return xor_top_from(a, na, b, nb, 64);
        const int qs = __builtin_amdgcn_readfirstlane(__builtin_ctzll(ne));
        const int xs = __builtin_amdgcn_readlane(x, qs), ys = __builtin_amdgcn_readlane(y, qs);
        return xs > ys ? xs : ys;  // both -1 when the sum is empty
    } else {
        return xor_top_from(a, na, b, nb, 0);
    }
}

// Allocate `n` pool entries by the team leader; ~0ull when the pool is full.
template <class Team>
__device__ __forceinline__ unsigned long long team_alloc(const Team &t, const Dev &d, unsigned n) {
    unsigned long long base = 0;
    if (t.lane() == 0) {
        base = atomicAdd(&d.ctrl->pool_top, (unsigned long long)n);
        if (base + n > d.ctrl->pool_cap) base = ~0ull;
    }
    return t.bcast(base);
}

struct Phase2Counters {
    unsigned long long nadd = 0, nel = 0;
    int spills = 0;
    int ovf = 0;
};

// The Phase II chain of column j (pms_functions.cu:87-112) by one team. The
// column's contents stay in LDS buffer sB (W entries) while they fit; a
// pivot column of at most WA entries is staged into sA (a longer one is read
// from the pool), the survivors of the sum go to sC and then back to sB.
// With HANDOFF, a column whose next sum may not fit W is committed as it
// stands and queued on the long list for the long-column pass, which
// continues the same chain; without it, the team merges straight in the pool.
template <class Team, int W, int WA, bool HANDOFF>
__device__ void phase2_column(const Team &t, const Dev &d, int j, int *buf, Phase2Counters &ctr) {
    const int gl = t.lane();
    constexpr int G = Team::size;
    int lj = d.low[j];
    // a pass that runs again after a suspension meets columns whose chain
    // already ended, with no low
    int4 pd = d.pivdesc[lj > -1 ? lj : 0];
    if (lj < 0) pd.x = -1;
    unsigned coff = d.off[j];
    int clen = d.len[j], ccap = d.cap[j];
    int *sA = buf, *sB = buf + WA, *sC = buf + WA + W;
    // The column's state as bits of one int: a divergent bool carried through
    // the loop would be a lane mask, merged at every join of the loop body.
    // kGlob: the contents are in the pool instead of LDS (in LDS they end
    // with the sentinel kIntMax, so a column there holds fewer than W
    // entries); kHand: handed off; kDirty: changed; kPivot: becomes the pivot
    // of its low; kLost: progress held only in LDS could not be written back;
    // kStop: the chain ends here.
    constexpr int kGlob = 1, kHand = 2, kDirty = 4, kPivot = 8, kLost = 16, kStop = 32;
    int fl = clen >= W ? (HANDOFF ? kGlob | kHand : kGlob) : 0;
    // this column's additions, merged elements and spills, counted once its
    // state is committed: work thrown away by a full pool is not progress
    unsigned long long padd = 0, pel = 0;
    int pspill = 0;
    if (!(fl & kGlob)) {
        for (int i = gl; i < clen; i += G) sB[i] = d.pool[coff + i];
        sB[clen] = kIntMax;
    }
    t.sync();
    // Teams of 32 lanes and more take the long columns, whose pivots are
    // mostly short. The next low of a sum is known from the tops of the two
    // lists before the merge (xor_top), so such a team loads the next
    // descriptor while it walks the merge and the next pivot's entries (at
    // most PFR per lane, into registers) while it copies the survivors back:
    // a step then waits for one global load, after the merge, instead of two
    // in a row before it. Pivots read in a Phase II pass are not written
    // during it, so the early loads see what the late ones would.
    constexpr bool PF = G >= 32;
    constexpr int PFR = 4;
    int pa[PFR];
    bool have = false;  // pa holds the entries of pd's pivot
    // one exit, at the top
    while (!(fl & (kHand | kStop)) && pd.x > -1 && pd.x < j) {
        const int na = pd.z;
        const unsigned aoff = (unsigned)pd.y;
        const bool pre = PF && have;  // this step's pivot is in pa
        have = false;
        bool nxt = false;  // pdn = pivdesc[nl], loaded during this step
        // written only by its load (and read only when nxt is set): a write
        // here would make the compiler wait for every load in flight
        int4 pdn;
        int nc = 0, nl = 0;
        if (!(fl & kGlob) && na + clen < W) {
            // WA >= W: every pivot of a sum that fits in LDS fits sA
            const bool ain = WA >= W || na < WA;
            if (ain) {
                if (pre) {
#pragma unroll
                    for (int r = 0; r < PFR; r++) {
                        const int i = gl + G * r;
                        if (i < na) sA[i] = pa[r];
                    }
                } else {
                    for (int i = gl; i < na; i += G) sA[i] = d.pool[aoff + i];
                }
                sA[na] = kIntMax;
                t.sync();
            }
            if (PF && ain) {
                // The early loads are unconditional (a valid row when there
                // is no next low) and their registers are not written on any
                // other path: a write to a register with a load in flight
                // makes the compiler wait for that load at once.
                const int nlp = xor_top(t, sA, na, sB, clen);
                nxt = nlp > -1 && nlp < lj;
                pdn = d.pivdesc[nxt ? nlp : lj];
                const XorWalk w = xor_walk<true>(t, sA, na, sB, clen, sC);
                nc = w.total;
                // the next pivot's entries, when the chain goes on with it;
                // lanes past its end read its last entry
                if (nxt && pdn.x > -1 && pdn.x < j && pdn.w != j && pdn.z <= PFR * G) {
                    const int nr = (pdn.z + G - 1) / G;
#pragma unroll
                    for (int r = 0; r < PFR; r++) {
                        const int i = gl + G * r;
                        if (r < nr) pa[r] = d.pool[(unsigned)pdn.y + (i < pdn.z ? i : pdn.z - 1)];
                    }
                    have = true;
                }
                // every walk is over after xor_walk's scan
                xor_copy(w, sB, sC, W - 1);
                // Stores go after the loads they would otherwise delay: a
                // wait for a load also waits for every older store. Every
                // lane of the team stores (the same value to the same
                // address), so that no branch hides the store from the
                // compiler's wait counting.
                d.ess[lj] = 0;
                t.sync();
                nl = nc ? sB[nc - 1] : -1;
                if (nl != nlp) {  // the scan and the merge must agree
                    if (gl == 0) atomicOr(&d.ctrl->err, kErrState);
                    fl |= kStop;
                }
            } else {
                d.ess[lj] = 0;
                if (ain)
                    nc = team_xor_into_b<true>(t, sA, na, sB, clen, sC, W - 1);
                else
                    nc = team_xor_into_b<false>(t, d.pool + aoff, na, sB, clen, sC, W - 1);
                t.sync();
                nl = nc ? sB[nc - 1] : -1;
            }
        } else if (HANDOFF) {
            fl |= kHand;
        } else {
            if (gl == 0) d.ess[lj] = 0;
            // the result may not fit in LDS: merge straight into a fresh pool
            // segment, reading the current contents from LDS the first time
            // and from the pool afterwards
            const unsigned long long base = team_alloc(t, d, (unsigned)(na + clen));
            if (base == ~0ull) {
                ctr.ovf = 1;
                fl |= kStop | ((fl & (kGlob | kDirty)) == kDirty ? kLost : 0);
            } else {
                const int *cur = (fl & kGlob) ? d.pool + coff : sB;
                nc = team_xor(t, d.pool + aoff, na, cur, clen, d.pool + base);
                t.fence();
                if (!(fl & kGlob)) pspill++;
                fl |= kGlob;
                coff = (unsigned)base;
                ccap = na + clen;
                nl = nc ? d.pool[base + nc - 1] : -1;
            }
        }
        if (!(fl & (kHand | kStop))) {
            padd++;
            pel += (unsigned long long)(na + clen + nc);
            clen = nc;
            fl |= kDirty;
            if (nl >= lj) {
                if (gl == 0) atomicOr(&d.ctrl->err, kErrState);
                fl |= kStop;
            } else {
                lj = nl;
                if (lj < 0) {
                    fl |= kStop;
                } else if (nxt) {
                    pd = pdn;  // loaded during the merge
                    if (pd.w == j) fl |= kPivot | kStop;
                    d.clear[lj] = 1;
                } else {
                    // .w = left[lj]: one load for the rule and the pivot, used
                    // here so that its wait stays on this path
                    pd = d.pivdesc[lj];
                    if (pd.w == j) fl |= kPivot | kStop;
                    d.clear[lj] = 1;
                }
            }
        }
    }
    const bool glob = fl & kGlob, dirty = fl & kDirty, lost = fl & kLost;
    const bool handoff = fl & kHand, becomes_pivot = fl & kPivot;
    bool committed = false;
    if (!lost && dirty) {
        bool ok = true;
        if (!glob) {
            // write back from LDS: in place when the segment is big enough
            if (clen > ccap) {
                unsigned want = (unsigned)clen;
                if (want < (unsigned)d.slot0) want = (unsigned)d.slot0;
                want = (want + 7u) & ~7u;
                const unsigned long long base = team_alloc(t, d, want);
                if (base == ~0ull) {
                    ctr.ovf = 1;
                    ok = false;  // nothing of this column's progress is kept
                } else {
                    coff = (unsigned)base;
                    ccap = (int)want;
                }
            }
            if (ok)
                for (int i = gl; i < clen; i += G) d.pool[coff + i] = sB[i];
        }
        if (ok && gl == 0) {
            d.off[j] = coff;
            d.len[j] = clen;
            d.cap[j] = ccap;
            d.low[j] = lj;
            if (becomes_pivot) {
                d.newpiv[lj] = j;
                d.classes[j] = -1;
            }
        }
        committed = ok;
    }
    if (committed) {
        ctr.nadd += padd;
        ctr.nel += pel;
        ctr.spills += pspill;
    }
    // a handed-off column continues in the long-column pass from the state
    // just committed (or from its unchanged state)
    if (HANDOFF && handoff && (committed || !dirty) && gl == 0) {
        const int k = atomicAdd(&d.ctrl->nlong, 1);
        d.longw[k] = j;
        atomicAdd(&d.ctrl->handoffs, 1);
    }
    t.sync();  // the buffers are reused by the team's next column
}

// Lane-group shapes: G lanes per column, W LDS entries for the column and
// for the sum, WA for the pivot, kBlock / G groups per block. WA = W but for
// the simplicial shapes of 8 lanes. A 16-entry pivot buffer (99.4 % of the
// pivots of a VR 2-skeleton reduction fit, the others are read from the
// pool) brings an (8, 80) block to 22 KB: seven blocks share a CU's 160 KB
// instead of five. (8, 72) with 12 entries and (8, 64) with 16 stay under
// 20 KB: eight blocks, the most the kernel's registers allow.
constexpr int pivot_lds(int G, int W) {
    return G != 8 ? W : W == 80 ? 16 : W == 72 ? 12 : W == 64 ? 16 : W;
}

template <int G, int W>
struct GroupShape {
    static constexpr int NG = kBlock / G;
    static constexpr int WA = pivot_lds(G, W);
    static constexpr int PER = WA + 2 * W;  // LDS entries per group
    static constexpr int LDS_BYTES = NG * PER * 4;
    static_assert(LDS_BYTES <= 64 * 1024, "static LDS limit");
};

template <int G, int W>
__global__ void __launch_bounds__(kBlock) k_phase2_group(Dev d) {
    using S = GroupShape<G, W>;
    constexpr int NG = S::NG;
    __shared__ __align__(16) int lds[NG * S::PER];
    const WaveTeam<G> t{(int)(threadIdx.x % G), (int)((threadIdx.x & 63) / G)};
    const int gid = threadIdx.x / G;
    int *const buf = lds + gid * S::PER;
    if (d.ctrl->halt) return;
    const int nwork = d.ctrl->nwork;
    if (nwork <= d.wide_max) return;  // the wide pass takes short worklists
    Phase2Counters ctr;
    for (int w = blockIdx.x * NG + gid; w < nwork; w += gridDim.x * NG)
        phase2_column<WaveTeam<G>, W, S::WA, true>(t, d, d.work[w], buf, ctr);
    if (ctr.ovf && t.gl == 0) d.ctrl->overflow = 1;
    // per-group counters are identical on the group's lanes: count leaders only
    if (t.gl != 0) ctr = Phase2Counters();
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        ctr.nadd += __shfl_down(ctr.nadd, o, 64);
        ctr.nel += __shfl_down(ctr.nel, o, 64);
        ctr.spills += __shfl_down(ctr.spills, o, 64);
    }
    // then per block: the counters live on one address each, where every
    // atomic waits for the one before
    __shared__ unsigned long long s_add[kBlock / 64], s_el[kBlock / 64];
    __shared__ int s_sp[kBlock / 64];
    const int wid = threadIdx.x >> 6;
    if (lane_id() == 0) {
        s_add[wid] = ctr.nadd;
        s_el[wid] = ctr.nel;
        s_sp[wid] = ctr.spills;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long a = 0, e = 0;
        int sp = 0;
        for (int k = 0; k < kBlock / 64; k++) {
            a += s_add[k];
            e += s_el[k];
            sp += s_sp[k];
        }
        stripe_add(d.ctrl, a, e, sp);
    }
}

// Columns handed off by the lane groups of 1024-entry buffers: one block of
// kLongTeam lanes per column, 4096-entry LDS buffers (48 KB: three blocks
// per CU), merging straight in the pool beyond that.
constexpr int kLongW = 4096;
constexpr int kLongBlock = kLongBlockLanes;
constexpr int kLongTeam = 256;

__global__ void __launch_bounds__(kLongTeam) k_phase2_long(Dev d) {
    __shared__ int lds[3 * kLongW];
    __shared__ int s_scan[2 * kLongTeam / 64];
    __shared__ unsigned long long s_word[2];
    const BlockTeam<kLongTeam> t{s_scan, s_word};
    if (d.ctrl->halt) return;
    const int nlong = d.ctrl->nlong;
    Phase2Counters ctr;
    for (int w = blockIdx.x; w < nlong; w += gridDim.x)
        phase2_column<BlockTeam<kLongTeam>, kLongW, kLongW, false>(t, d, d.longw[w], lds, ctr);
    if (threadIdx.x == 0) {
        if (ctr.ovf) d.ctrl->overflow = 1;
        stripe_add(d.ctrl, ctr.nadd, ctr.nel, ctr.spills);
    }
}

// Short worklists (at most Dev::wide_max columns, the late iterations of
// reductions whose columns grow long): the GPU holds every column at once,
// so each gets a team of kLongTeam lanes with 4096-entry buffers instead of
// one wave: a merge of n entries walks n / 256 diagonals per lane instead of
// n / 64, and the chains of dependent additions, which set the pass's time,
// get shorter steps.
__global__ void __launch_bounds__(kLongTeam) k_phase2_wide(Dev d) {
    __shared__ int lds[3 * kLongW];
    __shared__ int s_scan[2 * kLongTeam / 64];
    __shared__ unsigned long long s_word[2];
    const BlockTeam<kLongTeam> t{s_scan, s_word};
    if (d.ctrl->halt) return;
    const int nwork = d.ctrl->nwork;
    if (nwork > d.wide_max) return;
    Phase2Counters ctr;
    for (int w = blockIdx.x; w < nwork; w += gridDim.x)
        phase2_column<BlockTeam<kLongTeam>, kLongW, kLongW, false>(t, d, d.work[w], lds, ctr);
    if (threadIdx.x == 0) {
        if (ctr.ovf) d.ctrl->overflow = 1;
        stripe_add(d.ctrl, ctr.nadd, ctr.nel, ctr.spills);
    }
}

// Columns handed off by lane groups whose buffers hold fewer than kMidW
// entries: one wave per column with kMidW-entry LDS buffers (4 waves, 48 KB
// per block), merging in the pool beyond that. No block barriers, and many
// more columns in flight than one block per column gives.
constexpr int kMidW = 1024;

__global__ void __launch_bounds__(kBlock) k_phase2_mid(Dev d) {
    constexpr int NW = kBlock / 64;
    __shared__ __align__(16) int lds[NW * 3 * kMidW];
    __shared__ unsigned long long s_add[NW], s_el[NW];
    __shared__ int s_sp[NW], s_ovf[NW];
    const WaveTeam<64> t{lane_id(), 0};
    const int wid = threadIdx.x >> 6;
    if (d.ctrl->halt) return;
    const int nlong = d.ctrl->nlong;
    Phase2Counters ctr;
    for (int w = blockIdx.x * NW + wid; w < nlong; w += gridDim.x * NW)
        phase2_column<WaveTeam<64>, kMidW, kMidW, false>(t, d, d.longw[w],
                                                         lds + wid * 3 * kMidW, ctr);
    // the counters are the same on every lane of a wave (one column team)
    if (lane_id() == 0) {
        s_add[wid] = ctr.nadd;
        s_el[wid] = ctr.nel;
        s_sp[wid] = ctr.spills;
        s_ovf[wid] = ctr.ovf;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long a = 0, e = 0;
        int sp = 0, ovf = 0;
        for (int k = 0; k < NW; k++) {
            a += s_add[k];
            e += s_el[k];
            sp += s_sp[k];
            ovf |= s_ovf[k];
        }
        if (ovf) d.ctrl->overflow = 1;
        if (a || sp) stripe_add(d.ctrl, a, e, sp);
    }
}

// The lane-group shapes that are compiled: {lanes, LDS entries per buffer}.
using Phase2Kernel = void (*)(Dev);

static Phase2Kernel phase2_kernel(int lanes, int wlds) {
#define OPH_SHAPE(G, W) \
    if (lanes == G && wlds == W) return k_phase2_group<G, W>;
    OPH_SHAPE(4, 32)
    OPH_SHAPE(4, 64)
    OPH_SHAPE(8, 32)
    OPH_SHAPE(8, 64)
    OPH_SHAPE(8, 72)
    OPH_SHAPE(8, 80)
    OPH_SHAPE(8, 96)
    OPH_SHAPE(8, 128)
    OPH_SHAPE(16, 64)
    OPH_SHAPE(16, 128)
    OPH_SHAPE(16, 256)
    OPH_SHAPE(32, 256)
    OPH_SHAPE(32, 512)
    OPH_SHAPE(64, 512)
    OPH_SHAPE(64, 1024)
#undef OPH_SHAPE
    return nullptr;
}

bool phase2_shape_ok(int lanes, int wlds) {
    return lanes == 1 || phase2_kernel(lanes, wlds) != nullptr;
}

int phase2_default_lds(int lanes) {
    switch (lanes) {
        case 4: return 64;
        case 8: return 80;
        case 16: return 128;
        case 32: return 512;
        case 64: return 1024;
        default: return 0;
    }
}

int phase2_blocks_per_cu(int lanes, int wlds) {
    Phase2Kernel k = phase2_kernel(lanes, wlds);
    if (!k) return 0;
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void *>(k),
                                                     kBlock, 0) != hipSuccess || n < 1)
        n = 1;
    return n;
}

void launch_phase2(const Dev &d, int lanes, int wlds, int grid, int long_grid, hipStream_t s) {
    Phase2Kernel k = phase2_kernel(lanes, wlds);
    if (!k) {
        int gm = (d.m + kBlock - 1) / kBlock;
        k_phase2_lane<<<gm, kBlock, 0, s>>>(d);
        return;
    }
    k<<<grid, kBlock, 0, s>>>(d);
    // handed-off columns: wave teams when the groups' buffers are smaller
    // than theirs, else 256-lane teams (three 48 KB blocks per CU each)
    if (wlds < kMidW) {
        k_phase2_mid<<<3 * long_grid, kBlock, 0, s>>>(d);
    } else {
        k_phase2_long<<<3 * long_grid, kLongTeam, 0, s>>>(d);
        if (d.wide_max) k_phase2_wide<<<3 * long_grid, kLongTeam, 0, s>>>(d);
    }
}

// ---------------------------------------------------------------------------
// The serial algorithms on the device: standard (standard.cu:18-54) and
// twist (twist.cu:16-87), whose GPU forms in the reference launch one
// single-thread kernel per column (standard_parallel.cu, twist_parallel.cu).
// Here one workgroup of kLongBlock lanes takes the columns of `order` one
// after another. A column is reduced by the pivots of the columns before it
// (reduce_col_host, kernels_host.cu:62-75): while its low has a pivot, the
// pivot column is added (merge path over the block, in LDS while the sum
// fits, in the pool beyond). It then becomes the pivot of its low; cell and
// low are marked non-essential and their classes set (twist.cu:44-58); with
// twist the low's own column is cleared (twist_step_host, twist.cu:1-7).
// A full pool stops the kernel at the column it was on (Ctrl::nwork = its
// position, Ctrl::overflow = 1); the host compacts and starts it again there.
// stamps (may be null) receives the device clock after each column, start
// (may be null) the clock before the first.
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(kLongBlock) k_serial(Dev d, const int *__restrict__ order, int n,
                                                       int twist, unsigned long long *stamps,
                                                       unsigned long long *start) {
    constexpr int W = kLongW;
    __shared__ int lds[3 * W];
    __shared__ int s_scan[2 * kLongBlock / 64];
    __shared__ unsigned long long s_word[2];
    const BlockTeam<kLongBlock> t{s_scan, s_word};
    const int gl = threadIdx.x;
    Phase2Counters ctr;
    if (start && gl == 0) *start = (unsigned long long)wall_clock64();
    for (int k = 0; k < n; k++) {
        const int j = order[k];
        int lj = d.low[j];
        int4 pd = lj > -1 ? d.pivdesc[lj] : make_int4(-1, 0, 0, 0);
        bool stop = false;
        if (pd.x > -1) {
            unsigned coff = d.off[j];
            int clen = d.len[j], ccap = d.cap[j];
            int *sA = lds, *sB = lds + W, *sC = lds + 2 * W;
            bool glob = clen > W, dirty = false, lost = false;
            // counted once the column's state is committed (see phase2_column)
            unsigned long long padd = 0, pel = 0;
            int pspill = 0;
            if (!glob)
                for (int i = gl; i < clen; i += kLongBlock) sB[i] = d.pool[coff + i];
            t.sync();
            while (pd.x > -1) {
                const int na = pd.z;
                const unsigned aoff = (unsigned)pd.y;
                if (gl == 0) d.ess[lj] = 0;
                int nc, nl;
                if (!glob && na + clen <= W) {
                    for (int i = gl; i < na; i += kLongBlock) sA[i] = d.pool[aoff + i];
                    t.sync();
                    nc = team_xor_lds(t, sA, na, sB, clen, sC);
                    t.sync();
                    int *tmp = sB;
                    sB = sA;
                    sA = tmp;
                    nl = nc ? sB[nc - 1] : -1;
                } else {
                    const unsigned long long base = team_alloc(t, d, (unsigned)(na + clen));
                    if (base == ~0ull) {
                        ctr.ovf = true;
                        lost = !glob && dirty;
                        break;
                    }
                    const int *cur = glob ? d.pool + coff : sB;
                    nc = team_xor(t, d.pool + aoff, na, cur, clen, d.pool + base);
                    t.fence();
                    if (!glob) pspill++;
                    glob = true;
                    coff = (unsigned)base;
                    ccap = na + clen;
                    nl = nc ? d.pool[base + nc - 1] : -1;
                }
                padd++;
                pel += (unsigned long long)(na + clen + nc);
                clen = nc;
                dirty = true;
                if (nl >= lj) {
                    if (gl == 0) atomicOr(&d.ctrl->err, kErrState);
                    break;
                }
                lj = nl;
                if (lj < 0) break;
                pd = d.pivdesc[lj];
            }
            if (ctr.ovf) {
                // keep what is in the pool; the column is taken up again
                // after compaction from the state committed here
                if (!lost && dirty && glob) {
                    if (gl == 0) {
                        d.off[j] = coff;
                        d.len[j] = clen;
                        d.cap[j] = ccap;
                        d.low[j] = lj;
                    }
                    ctr.nadd += padd;
                    ctr.nel += pel;
                    ctr.spills += pspill;
                }
                stop = true;
            } else if (dirty) {
                bool ok = true;
                if (!glob) {
                    if (clen > ccap) {
                        unsigned want = (unsigned)clen;
                        if (want < (unsigned)d.slot0) want = (unsigned)d.slot0;
                        want = (want + 7u) & ~7u;
                        const unsigned long long base = team_alloc(t, d, want);
                        if (base == ~0ull) {
                            ctr.ovf = true;
                            ok = false;
                        } else {
                            coff = (unsigned)base;
                            ccap = (int)want;
                        }
                    }
                    if (ok)
                        for (int i = gl; i < clen; i += kLongBlock) d.pool[coff + i] = sB[i];
                }
                if (ok && gl == 0) {
                    d.off[j] = coff;
                    d.len[j] = clen;
                    d.cap[j] = ccap;
                    d.low[j] = lj;
                }
                if (ok) {
                    ctr.nadd += padd;
                    ctr.nel += pel;
                    ctr.spills += pspill;
                }
                stop = !ok;
            }
        }
        if (stop) {
            if (gl == 0) {
                d.ctrl->overflow = 1;
                d.ctrl->nwork = k;
            }
            break;
        }
        if (gl == 0) {
            if (lj > -1) {
                publish_pivot(d, lj, j);
                d.classes[j] = -1;
                // the low's column is positive (standard.cu:31), unless it
                // is itself a pivot, which only a matrix that is not a
                // boundary matrix allows: its class stays negative
                if (d.classes[lj] != -1) d.classes[lj] = 1;
                d.ess[j] = 0;
                d.ess[lj] = 0;
                if (twist) d.low[lj] = -1;
            } else {
                d.classes[j] = 1;
            }
            if (stamps) stamps[k] = (unsigned long long)wall_clock64();
        }
        __syncthreads();  // the next column reads this one's pivot
    }
    if (gl == 0) {
        if (ctr.nadd) {
            atomicAdd(&d.ctrl->additions, ctr.nadd);
            atomicAdd(&d.ctrl->merge_elems, ctr.nel);
        }
        if (ctr.spills) atomicAdd(&d.ctrl->spills, ctr.spills);
    }
}

void launch_serial(const Dev &d, const int *order, int n, bool twist, unsigned long long *stamps,
                   unsigned long long *start, hipStream_t s) {
    k_serial<<<1, kLongBlock, 0, s>>>(d, order, n, twist ? 1 : 0, stamps, start);
}

// ---------------------------------------------------------------------------
// The row algorithm (ph_row.cu:38-121, ph_row_parallel.cu:20-84) on the
// device. Rows are taken from the last to the first; the leftmost column
// whose low is the row (its pivot) is added to every other column with that
// low, which moves to a smaller low. ph_row takes only the columns of the
// pivot's pseudo-dimension (get_neighbours_position_host walks the pivot's
// dimension list, ph_row.cu:25-36); ph_row_parallel takes every column with
// that low (mark_neighbours, ph_row_parallel.cu:2-10).
//
// One workgroup of kLongBlock lanes takes the rows in turn. The columns of
// each low are kept in linked lists (head[row], next[column]) so that a row
// finds them without a pass over all columns; a list may still hold a column
// whose low has moved on, which is skipped. For each row:
//   1. lane 0 walks the row's list and gathers its neighbours into cand[];
//   2. the block adds the pivot (staged in LDS once) to each neighbour;
//   3. the lanes push every moved neighbour onto the list of its new low and
//      mark ess.
// classes are not kept: they feed only the trackers, which the row
// algorithms leave as the caller set them, and the pool compaction drops the
// columns marked positive, which here may still be reduced.
// Within a row the additions are independent of each other, so their order
// does not matter but for one thing: the host loop marks a neighbour when
// it is, at that moment, the leftmost column of its new low
// (left_to_right_neighbours_host, ph_row.cu:14-21), and in its ascending
// order that is exactly the neighbour that is leftmost once the whole row is
// done. Step 3 marks that one, after the row.
// pivot[row] is the leftmost column that has had that low; a column's low
// only decreases, so it enters each list at most once and cand[] never holds
// a column twice.
// A full pool stops the kernel inside step 2; state = {row, next candidate,
// candidates} lets the host compact and start it again where it stopped.
// ---------------------------------------------------------------------------

// Lists, leftmost columns and pseudo-dimensions (initial lengths) of the
// initial lows; the resume state points at the last row.
__global__ void k_row_lists(Dev d, RowLists r) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j == 0) {
        r.state[0] = d.m - 1;
        r.state[1] = 0;
        r.state[2] = -1;
    }
    if (j >= d.m) return;
    r.plen[j] = d.len[j];
    const int l = d.low[j];
    if (l < 0) return;
    r.next[j] = atomicExch(&r.head[l], j);
    atomicMin(&r.pivot[l], j);
}

__global__ void k_fill_i(int *p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void __launch_bounds__(kLongBlock) k_ph_row(Dev d, RowLists r, int all_dims,
                                                      unsigned long long *stamps,
                                                      unsigned long long *start) {
    constexpr int W = kLongW;
    __shared__ int lds[3 * W];
    __shared__ int s_scan[2 * kLongBlock / 64];
    __shared__ unsigned long long s_word[2];
    __shared__ int s_n;
    const BlockTeam<kLongBlock> t{s_scan, s_word};
    const int gl = threadIdx.x;
    int *sP = lds, *sB = lds + W, *sC = lds + 2 * W;
    Phase2Counters ctr;
    if (start && gl == 0) *start = (unsigned long long)wall_clock64();
    int i = r.state[0], q = r.state[1], n = r.state[2];
    int progress = 0;
    bool stop = false;
    for (; i >= 0; i--, q = 0, n = -1) {
        const int pv = r.pivot[i];
        if (pv != kIntMax) {
            if (n < 0) {  // step 1
                if (gl == 0) {
                    // a list holds each column at most once: a longer walk
                    // is a broken list, reported instead of followed
                    const int pl = r.plen[pv];
                    int c = 0, steps = 0;
                    for (int j = __hip_atomic_load(&r.head[i], __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
                         j >= 0; j = r.next[j]) {
                        if (++steps > d.m || j >= d.m) {
                            atomicOr(&d.ctrl->err, kErrState);
                            break;
                        }
                        if (j != pv && d.low[j] == i && (all_dims || r.plen[j] == pl))
                            r.cand[c++] = j;
                    }
                    s_n = c;
                }
                t.fence();
                n = s_n;
                q = 0;
            }
            const unsigned poff = d.off[pv];
            const int np = d.len[pv];
            const bool pin = np <= W;
            if (n > q && pin)
                for (int k = gl; k < np; k += kLongBlock) sP[k] = d.pool[poff + k];
            t.sync();
            for (; q < n; q++) {  // step 2
                const int j = r.cand[q];
                const unsigned coff = d.off[j];
                const int clen = d.len[j], ccap = d.cap[j];
                const bool cin = clen <= W;
                if (cin)
                    for (int k = gl; k < clen; k += kLongBlock) sB[k] = d.pool[coff + k];
                t.sync();
                int nc, nl;
                unsigned roff = coff;
                int rcap = ccap;
                if (pin && cin && np + clen <= W) {
                    nc = team_xor(t, sP, np, sB, clen, sC);
                    t.sync();
                    if (nc > rcap) {
                        unsigned want = (unsigned)nc;
                        if (want < (unsigned)d.slot0) want = (unsigned)d.slot0;
                        want = (want + 7u) & ~7u;
                        const unsigned long long base = team_alloc(t, d, want);
                        if (base == ~0ull) {
                            stop = true;
                            break;
                        }
                        roff = (unsigned)base;
                        rcap = (int)want;
                    }
                    for (int k = gl; k < nc; k += kLongBlock) d.pool[roff + k] = sC[k];
                    nl = nc ? sC[nc - 1] : -1;
                } else {
                    const unsigned long long base = team_alloc(t, d, (unsigned)(np + clen));
                    if (base == ~0ull) {
                        stop = true;
                        break;
                    }
                    nc = team_xor(t, pin ? sP : d.pool + poff, np, cin ? sB : d.pool + coff,
                                  clen, d.pool + base);
                    t.fence();
                    roff = (unsigned)base;
                    rcap = np + clen;
                    ctr.spills += cin ? 1 : 0;
                    nl = nc ? d.pool[base + nc - 1] : -1;
                }
                ctr.nadd++;
                ctr.nel += (unsigned long long)(np + clen + nc);
                progress++;
                if (gl == 0) {
                    if (nl >= i) atomicOr(&d.ctrl->err, kErrState);
                    d.off[j] = roff;
                    d.len[j] = nc;
                    d.cap[j] = rcap;
                    d.low[j] = nl;
                    if (nl > -1 && j < r.pivot[nl]) r.pivot[nl] = j;
                }
                t.fence();  // sB and sC are reused by the next neighbour
            }
            if (stop) break;
            for (int k = gl; k < n; k += kLongBlock) {  // step 3
                const int j = r.cand[k];
                const int nl = d.low[j];
                if (nl < 0) continue;
                r.next[j] = atomicExch(&r.head[nl], j);
                if (!all_dims && r.pivot[nl] == j) {
                    d.ess[j] = 0;
                    d.ess[nl] = 0;
                }
            }
            if (gl == 0 && n > 0 && !all_dims) {  // ph_row.cu:6-9
                d.ess[i] = 0;
                d.ess[pv] = 0;
            }
            t.fence();
        }
        progress++;
        if (stamps && gl == 0) stamps[d.m - 1 - i] = (unsigned long long)wall_clock64();
    }
    if (gl == 0) {
        if (stop) {
            r.state[0] = i;
            r.state[1] = q;
            r.state[2] = n;
            d.ctrl->overflow = 1;
        } else {
            r.state[0] = -1;
            r.state[1] = 0;
            r.state[2] = -1;
        }
        d.ctrl->nwork = progress;
        if (ctr.nadd) {
            atomicAdd(&d.ctrl->additions, ctr.nadd);
            atomicAdd(&d.ctrl->merge_elems, ctr.nel);
        }
        if (ctr.spills) atomicAdd(&d.ctrl->spills, ctr.spills);
    }
}

// ph_row_parallel's estimate (ess_hat, kernels.cu:113-124, with arglow never
// set by the row algorithm): 0 exactly at the rows that are some column's
// low after some row, i.e. at the final lows, and 1 everywhere else.
__global__ void k_ess_lows_mark(Dev d) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.m) return;
    const int lt = d.low[t];
    if (lt > -1) d.ess[lt] = 0;
}

void launch_row_lists(const Dev &d, const RowLists &r, hipStream_t s) {
    const int gm = (d.m + kBlock - 1) / kBlock;
    if (!d.m) return;
    k_fill_i<<<gm, kBlock, 0, s>>>(r.head, d.m, -1);
    k_fill_i<<<gm, kBlock, 0, s>>>(r.pivot, d.m, kIntMax);
    k_row_lists<<<gm, kBlock, 0, s>>>(d, r);
}

void launch_ph_row(const Dev &d, const RowLists &r, bool all_dims, unsigned long long *stamps,
                   unsigned long long *start, hipStream_t s) {
    k_ph_row<<<1, kLongBlock, 0, s>>>(d, r, all_dims ? 1 : 0, stamps, start);
}

void launch_ess_at_lows(const Dev &d, hipStream_t s) {
    const int gm = (d.m + kBlock - 1) / kBlock;
    if (!d.m) return;
    k_fill_i<<<gm, kBlock, 0, s>>>(d.ess, d.m, 1);
    k_ess_lows_mark<<<gm, kBlock, 0, s>>>(d);
}

// After Phase II: publish new pivots, clear_positives, ess_hat
// (kernels.cu:113-124) and update_classes (kernels.cu:93-101). Every thread
// touches its own column first, so one pass suffices. The kernel does not
// stop on Ctrl::halt, which its first thread sets for the next iteration:
// after a halted iteration nothing is pending (newpiv and clear were reset
// by the last pass) and the pass changes nothing.
__global__ void __launch_bounds__(kBlock) k_apply2(Dev d) {
    // a Phase II pass cut short by a full pool publishes nothing until it
    // has run again after the compaction (iter_suspend)
    if (d.ctrl->overflow) {
        if (blockIdx.x == 0 && threadIdx.x == 0) iter_suspend(d.ctrl);
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) iter_end(d.ctrl);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < d.m) {
        // column t's own entries, written by no other thread of the pass:
        // loaded together before any store
        const int np = d.newpiv[t], cr = d.clear[t], lo = d.low[t], al = d.arglow[t];
        if (np > -1) {
            publish_pivot(d, t, np);
            d.newpiv[t] = -1;
        }
        if (cr) {
            d.low[t] = -1;
            d.classes[t] = 1;
            d.clear[t] = 0;
        }
        const int lt = cr ? -1 : lo;
        if (lt > -1) d.ess[lt] = 0;
        const int a = np > -1 ? np : al;
        if (a > -1) {
            d.ess[a] = 0;
            d.ess[t] = 0;
            d.classes[a] = -1;
            d.classes[t] = 1;
        }
    }
}

// k_apply2 followed by the next iteration's k_first (Dev::fuse_first: the
// queued and graph schedules without trackers or iteration times, where no
// host step comes between two iterations), one launch less per iteration.
// The first thread ends the iteration and, unless it halts, starts the next;
// the blocks then take the next iteration's first occurrences from their
// columns as just updated. The arglow read for them races with this pass's
// publications, which is harmless: a row whose pivot is published here has
// arglow set for Phase I, which then does not look at first[] for that row.
// After any halt the blocks skip that part: a resumed loop (after a pool
// compaction) starts with k_first.
__global__ void __launch_bounds__(kBlock) k_apply2_first(Dev d) {
    __shared__ int hkey[kHash];
    __shared__ int hval[kHash];
    __shared__ int s_halt;
    const int base = blockIdx.x * kTile + threadIdx.x * kItems;
    int np[kItems], cr[kItems], lo[kItems], al[kItems], cl[kItems];
    // set by Phase II only, so every block reads the same value: block 0
    // resets it only when it was clear
    const int ovf = d.ctrl->overflow;
    // column t's own entries, written by no other thread of the pass
    load_run(d.newpiv, base, d.m, -1, np);
    load_run(d.clear, base, d.m, 0, cr);
    load_run(d.low, base, d.m, -1, lo);
    load_run(d.arglow, base, d.m, -1, al);
    load_run(d.classes, base, d.m, 1, cl);
    // a Phase II pass cut short by a full pool publishes nothing until it
    // has run again after the compaction (iter_suspend)
    if (ovf) {
        if (blockIdx.x == 0 && threadIdx.x == 0) iter_suspend(d.ctrl);
        return;
    }
    if (threadIdx.x == 0) {
        if (blockIdx.x == 0) {
            iter_end(d.ctrl);
            if (!d.ctrl->halt) iter_begin(d.ctrl);
        }
        s_halt = d.ctrl->halt;
    }
    int nl[kItems];
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        const int t = base + k;
        nl[k] = -1;
        if (t >= d.m) continue;
        if (np[k] > -1) {
            publish_pivot(d, t, np[k]);
            d.newpiv[t] = -1;
        }
        if (cr[k]) {
            d.low[t] = -1;
            d.classes[t] = 1;
            d.clear[t] = 0;
        }
        const int lt = cr[k] ? -1 : lo[k];
        if (lt > -1) d.ess[lt] = 0;
        const int a = np[k] > -1 ? np[k] : al[k];
        if (a > -1) {
            d.ess[a] = 0;
            d.ess[t] = 0;
            d.classes[a] = -1;
            d.classes[t] = 1;
        }
        // the next iteration's Phase I: unknown columns with a low
        const int c = cr[k] || a > -1 ? 1 : cl[k];
        nl[k] = c == 0 ? lt : -1;
    }
    int p[kItems];
    gather_run(d.arglow, nl, 0, p);
    first_hash_init(hkey, hval);
    __syncthreads();  // s_halt and the table
    if (s_halt) return;  // uniform over the block
    first_hash(d, base, nl, p, hkey, hval);
}

void launch_apply2(const Dev &d, hipStream_t s) {
    if (d.fuse_first) {
        if (d.ntiles) k_apply2_first<<<d.ntiles, kBlock, 0, s>>>(d);
        return;
    }
    int gm = (d.m + kBlock - 1) / kBlock;
    k_apply2<<<gm, kBlock, 0, s>>>(d);
}

void launch_first(const Dev &d, hipStream_t s) {
    if (d.ntiles) k_first<<<d.ntiles, kBlock, 0, s>>>(d);
}

// ---------------------------------------------------------------------------
// Convergence: ess_hat_final (kernels.cu:103-111) and the 1-based output
// (openph.cu:195-199).
// ---------------------------------------------------------------------------

// gate: only when the loop has converged (queued behind a chunk of
// iterations before the host knows whether the chunk converged)
__global__ void k_final_ess(Dev d, int gate) {
    if (gate && d.ctrl->halt != kHaltConverged) return;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.m) return;
    int lt = d.low[t];
    if (lt > -1) {
        d.ess[t] = 0;
        d.ess[lt] = 0;
    }
}

__global__ void k_output(Dev d, int *__restrict__ low1, int *__restrict__ ess, int gate) {
    if (gate && d.ctrl->halt != kHaltConverged) return;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.m) return;
    low1[t] = d.low[t] + 1;
    ess[t] = d.ess[t];
}

void launch_final(const Dev &d, int *low1_out, int *ess_out, hipStream_t s, bool gated) {
    int gm = (d.m + kBlock - 1) / kBlock;
    if (!d.m) return;
    k_final_ess<<<gm, kBlock, 0, s>>>(d, gated ? 1 : 0);
    k_output<<<gm, kBlock, 0, s>>>(d, low1_out, ess_out, gated ? 1 : 0);
}

void launch_output(const Dev &d, int *low1_out, int *ess_out, hipStream_t s) {
    if (d.m) k_output<<<(d.m + kBlock - 1) / kBlock, kBlock, 0, s>>>(d, low1_out, ess_out, 0);
}

void launch_final_ess(const Dev &d, hipStream_t s) {
    if (d.m) k_final_ess<<<(d.m + kBlock - 1) / kBlock, kBlock, 0, s>>>(d, 0);
}

// update_classes_final (kernels.cu:79-91). The stores cannot conflict: a
// column with a low is never itself a low (it is a negative, and lows are
// positives), and ess_hat_final has zeroed ess of every low.
__global__ void k_classes_final(Dev d) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.m) return;
    int lt = d.low[t];
    if (lt > -1) {
        d.classes[t] = -1;
        d.classes[lt] = 1;
    } else if (d.ess[t] == 1) {
        d.classes[t] = 2;
    }
}

// A member [a, b) of a running batch whose lows no longer repeat: its final
// ess (k_final_ess on the range) and output, then its columns inert.
__global__ void k_final_ess_range(Dev d, int a, int b) {
    const int t = a + blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b) return;
    const int lt = d.low[t];
    if (lt > -1) {
        d.ess[t] = 0;
        d.ess[lt] = 0;
    }
}

__global__ void k_output_retire(Dev d, int a, int b, int *__restrict__ low1,
                                int *__restrict__ ess) {
    const int t = a + blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b) return;
    const int lt = d.low[t];
    low1[t - a] = lt > -1 ? lt - a + 1 : 0;
    ess[t - a] = d.ess[t];
    d.low[t] = -1;
    d.classes[t] = 1;
    d.arglow[t] = -1;
}

void launch_retire_range(const Dev &d, int a, int b, int *low1_out, int *ess_out, hipStream_t s) {
    const int g = (b - a + kBlock - 1) / kBlock;
    if (g <= 0) return;
    k_final_ess_range<<<g, kBlock, 0, s>>>(d, a, b);
    k_output_retire<<<g, kBlock, 0, s>>>(d, a, b, low1_out, ess_out);
}

void launch_classes_final(const Dev &d, hipStream_t s) {
    if (d.m) k_classes_final<<<(d.m + kBlock - 1) / kBlock, kBlock, 0, s>>>(d);
}

// ---------------------------------------------------------------------------
// Trackers (tracker.cu:92-134) as exact integer reductions: the l1 and linf
// distance of low to low_true (both 0-based, as the reference compares them
// after indexShiftDown, openph.cu:113), the unreduced count #{classes == 0}
// and the essential estimate #{ess == 1}. One 64-bit atomic per block and
// quantity; the host turns them into the reference's floats.
// ---------------------------------------------------------------------------

__global__ void k_track_zero(Ctrl *c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        c->trk_l1 = 0;
        c->trk_linf = 0;
        c->trk_unreduced = 0;
        c->trk_ess = 0;
    }
}

__global__ void __launch_bounds__(kBlock) k_track(Dev d, const int *__restrict__ low_true0) {
    __shared__ unsigned long long s_l1[kBlock / 64];
    __shared__ int s_inf[kBlock / 64], s_un[kBlock / 64], s_es[kBlock / 64];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long l1 = 0;
    int inf = 0, un = 0, es = 0;
    if (t < d.m) {
        const long long diff = (long long)d.low[t] - (long long)low_true0[t];
        const int a = (int)(diff < 0 ? -diff : diff);
        l1 = (unsigned long long)a;
        inf = a;
        un = d.classes[t] == 0;
        es = d.ess[t] == 1;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        l1 += __shfl_down(l1, o, 64);
        const int x = __shfl_down(inf, o, 64);
        inf = x > inf ? x : inf;
        un += __shfl_down(un, o, 64);
        es += __shfl_down(es, o, 64);
    }
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        s_l1[w] = l1;
        s_inf[w] = inf;
        s_un[w] = un;
        s_es[w] = es;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBlock / 64; k++) {
            l1 += s_l1[k];
            inf = s_inf[k] > inf ? s_inf[k] : inf;
            un += s_un[k];
            es += s_es[k];
        }
        if (l1) atomicAdd(&d.ctrl->trk_l1, l1);
        if (inf) atomicMax(&d.ctrl->trk_linf, inf);
        if (un) atomicAdd(&d.ctrl->trk_unreduced, un);
        if (es) atomicAdd((unsigned long long *)&d.ctrl->trk_ess, (unsigned long long)es);
    }
}

void launch_track(const Dev &d, const int *low_true0, hipStream_t s) {
    k_track_zero<<<1, 64, 0, s>>>(d.ctrl);
    if (d.m) k_track<<<(d.m + kBlock - 1) / kBlock, kBlock, 0, s>>>(d, low_true0);
}

// The same four quantities per block, overwritten (not accumulated), so that
// an instance queued past a halt, which sees the same state, writes the same
// values; k_track_fold sums them into the iteration's ring entry.
constexpr int kTrackItems = 4;  // columns per thread
constexpr int kTrackTile = kBlock * kTrackItems;

int track_part_blocks(int m) { return (m + kTrackTile - 1) / kTrackTile; }

__global__ void __launch_bounds__(kBlock) k_track_part(Dev d) {
    __shared__ unsigned long long s_l1[kBlock / 64];
    __shared__ int s_inf[kBlock / 64], s_un[kBlock / 64], s_es[kBlock / 64];
    unsigned long long l1 = 0;
    int inf = 0, un = 0, es = 0;
#pragma unroll
    for (int k = 0; k < kTrackItems; k++) {
        const int t = blockIdx.x * kTrackTile + k * kBlock + threadIdx.x;
        if (t < d.m) {
            const long long diff = (long long)d.low[t] - (long long)d.ltrue[t];
            const int a = (int)(diff < 0 ? -diff : diff);
            l1 += (unsigned long long)a;
            inf = a > inf ? a : inf;
            un += d.classes[t] == 0;
            es += d.ess[t] == 1;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        l1 += __shfl_down(l1, o, 64);
        const int x = __shfl_down(inf, o, 64);
        inf = x > inf ? x : inf;
        un += __shfl_down(un, o, 64);
        es += __shfl_down(es, o, 64);
    }
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        s_l1[w] = l1;
        s_inf[w] = inf;
        s_un[w] = un;
        s_es[w] = es;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBlock / 64; k++) {
            l1 += s_l1[k];
            inf = s_inf[k] > inf ? s_inf[k] : inf;
            un += s_un[k];
            es += s_es[k];
        }
        unsigned long long *p = d.trk_part + 4 * (size_t)blockIdx.x;
        p[0] = l1;
        p[1] = (unsigned long long)inf;
        p[2] = (unsigned long long)un;
        p[3] = (unsigned long long)es;
    }
}

__global__ void __launch_bounds__(1024) k_track_fold(Dev d, int nblocks) {
    __shared__ unsigned long long s[4][1024 / 64];
    unsigned long long l1 = 0, inf = 0, un = 0, es = 0;
    for (int b = threadIdx.x; b < nblocks; b += 1024) {
        const unsigned long long *p = d.trk_part + 4 * (size_t)b;
        l1 += p[0];
        inf = p[1] > inf ? p[1] : inf;
        un += p[2];
        es += p[3];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        l1 += __shfl_down(l1, o, 64);
        const unsigned long long x = __shfl_down(inf, o, 64);
        inf = x > inf ? x : inf;
        un += __shfl_down(un, o, 64);
        es += __shfl_down(es, o, 64);
    }
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        s[0][w] = l1;
        s[1][w] = inf;
        s[2][w] = un;
        s[3][w] = es;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 1024 / 64; k++) {
            l1 += s[0][k];
            inf = s[1][k] > inf ? s[1][k] : inf;
            un += s[2][k];
            es += s[3][k];
        }
        unsigned long long *q = d.trk_ring + 4 * (size_t)(d.ctrl->iter % kIterRing);
        q[0] = l1;
        q[1] = inf;
        q[2] = un;
        q[3] = es;
    }
}

__global__ void k_iter_end_stamp(Dev d, int again) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    unsigned long long *e = d.its + 2 * (d.ctrl->iter % kIterRing) + 1;
    if (*e == 0 || again) *e = (unsigned long long)wall_clock64();
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

void launch_iter_end_stamp(const Dev &d, hipStream_t s, int again) {
    k_iter_end_stamp<<<1, 64, 0, s>>>(d, again);
}

void launch_track_iter(const Dev &d, hipStream_t s) {
    if (!d.m) return;
    const int nb = track_part_blocks(d.m);
    k_track_part<<<nb, kBlock, 0, s>>>(d);
    k_track_fold<<<1, 1024, 0, s>>>(d, nb);
}

void launch_clear_stamps(const Dev &d, hipStream_t s) {
    k_zero_u64<<<(2 * kIterRing + kBlock - 1) / kBlock, kBlock, 0, s>>>(d.its, 2 * kIterRing);
}

// ---------------------------------------------------------------------------
// Pool compaction: live columns are copied, in column order, to the other
// pool; cleared columns are dropped (never reduced again, never a pivot).
// ---------------------------------------------------------------------------

// longw[] (free outside Phase II: the pass rebuilds the long-column list,
// also when it runs again after a suspension, which needs work[] intact)
// marks the published pivots for the compaction: longw[p] = 1 + i where
// arglow[i] = p.
__global__ void k_pivot_rows_clear(Dev d) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < d.m) d.longw[j] = 0;
}

__global__ void k_pivot_rows(Dev d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.m) return;
    const int p = d.arglow[i];
    if (p > -1 && p < d.m) d.longw[p] = i + 1;
}

__global__ void k_copy_cols(Dev d) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d.m) return;
    unsigned sz = col_size(d, j, 1);
    int n = d.len[j];
    if (sz == 0) {
        d.len[j] = 0;
        d.cap[j] = 0;
        return;
    }
    const int *src = d.pool + d.off[j];
    const unsigned no = d.off2[j];
    int *dst = d.pool2 + no;
    for (int i = 0; i < n; i++) dst[i] = src[i];
    d.cap[j] = (int)sz;
    const int r = d.longw[j] - 1;  // the row whose pivot j is, or -1
    if (r > -1) d.pivdesc[r] = make_int4(j, (int)no, n, d.left[r]);
}

void launch_compact_scan(const Dev &d, hipStream_t s) {
    const int gm = (d.m + kBlock - 1) / kBlock;
    if (d.m) {
        k_pivot_rows_clear<<<gm, kBlock, 0, s>>>(d);
        k_pivot_rows<<<gm, kBlock, 0, s>>>(d);
    }
    launch_scan(d, 1, d.off2, s);
}

void launch_compact_copy(Dev &d, hipStream_t s) {
    int gm = (d.m + kBlock - 1) / kBlock;
    k_copy_cols<<<gm, kBlock, 0, s>>>(d);
    unsigned *to = d.off;
    d.off = d.off2;
    d.off2 = to;
    int *tp = d.pool;
    d.pool = d.pool2;
    d.pool2 = tp;
}

// ---------------------------------------------------------------------------
// Batches: `count` matrices reduced as one block-diagonal matrix. Matrix k
// owns columns and rows [moff[k], moff[k+1]) and COO entries
// [noff[k], noff[k+1]); its input indices are local and 1-based.
// ---------------------------------------------------------------------------


// kBatchRun entries (or columns) per thread, kBlock apart (coalesced): one
// binary search over the members for the first, then a forward step where a
// member ends before the next (members hold thousands of entries), so that a
// pass over the batch waits for one search per thread instead of one per
// entry.
constexpr int kBatchRun = 16;

__global__ void k_batch_offsets(const int *__restrict__ rows1, const int *__restrict__ cols1,
                                const int *__restrict__ noff, const int *__restrict__ moff,
                                int count, int nnz, int *__restrict__ rows_g,
                                int *__restrict__ cols_g, int *err) {
    const long long e0 = (long long)blockIdx.x * kBlock * kBatchRun + threadIdx.x;
    if (e0 >= nnz) return;
    // the entries' loads first, all in flight with the search
    int r[kBatchRun], c[kBatchRun];
#pragma unroll
    for (int i = 0; i < kBatchRun; i++) {
        const long long el = e0 + (long long)i * kBlock;
        const int e = el < nnz ? (int)el : nnz - 1;
        r[i] = rows1[e];
        c[i] = cols1[e];
    }
    int k = segment_of(noff, count, (int)e0);
    int end = noff[k + 1], base = moff[k], mk = moff[k + 1] - base;
    int bad = 0;
#pragma unroll
    for (int i = 0; i < kBatchRun; i++) {
        const long long el = e0 + (long long)i * kBlock;
        if (el >= nnz) break;
        const int e = (int)el;
        while (e >= end && k + 1 < count) {  // the next member (possibly empty ones first)
            k++;
            end = noff[k + 1];
            base = moff[k];
            mk = moff[k + 1] - base;
        }
        bad |= r[i] < 1 || c[i] < 1 || r[i] > mk || c[i] > mk;
        rows_g[e] = r[i] + base;
        cols_g[e] = c[i] + base;
    }
    if (bad) atomicOr(err, 1);
}

__global__ void k_batch_local_low(int *__restrict__ low1, const int *__restrict__ moff, int count,
                                  int m) {
    const long long j0 = (long long)blockIdx.x * kBlock * kBatchRun + threadIdx.x;
    if (j0 >= m) return;
    int l[kBatchRun];
#pragma unroll
    for (int i = 0; i < kBatchRun; i++) {
        const long long jl = j0 + (long long)i * kBlock;
        l[i] = low1[jl < m ? (int)jl : m - 1];
    }
    int k = segment_of(moff, count, (int)j0);
    int end = moff[k + 1], base = moff[k];
#pragma unroll
    for (int i = 0; i < kBatchRun; i++) {
        const long long jl = j0 + (long long)i * kBlock;
        if (jl >= m) break;
        const int j = (int)jl;
        while (j >= end && k + 1 < count) {
            k++;
            end = moff[k + 1];
            base = moff[k];
        }
        if (l[i] > 0) low1[j] = l[i] - base;
    }
}

static int batch_run_blocks(long long n) {
    return (int)((n + (long long)kBlock * kBatchRun - 1) / ((long long)kBlock * kBatchRun));
}

void launch_batch_offsets(const int *rows1, const int *cols1, const int *noff, const int *moff,
                          int count, int nnz, int *rows_g, int *cols_g, int *err, hipStream_t s) {
    if (nnz > 0)
        k_batch_offsets<<<batch_run_blocks(nnz), kBlock, 0, s>>>(rows1, cols1, noff, moff, count,
                                                                 nnz, rows_g, cols_g, err);
}

void launch_batch_local_low(int *low1, const int *moff, int count, int m, hipStream_t s) {
    if (m > 0)
        k_batch_local_low<<<batch_run_blocks(m), kBlock, 0, s>>>(low1, moff, count, m);
}

}  // namespace oph
