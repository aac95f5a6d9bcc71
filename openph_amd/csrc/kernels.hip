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
// Columns live in a pool of int32 row indices: column j occupies
// pool[off[j] .. off[j]+len[j]) in ascending order. A column addition
// writes its result to a fresh pool segment (bump allocation), so the
// previous contents stay valid until the host compacts the pool.
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

// Exclusive scan over the whole block (blockDim.x multiple of 64). `lds`
// needs blockDim.x/64 entries. Returns the exclusive prefix; *total gets the
// block aggregate. Contains two __syncthreads(): every thread must call it.
template <typename T, typename Op>
__device__ T block_exclusive_scan(T v, Op op, T identity, T *lds, T *total) {
    const int lane = lane_id(), wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    T incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T t = __shfl_up(incl, d, 64);
        if (lane >= d) incl = op(incl, t);
    }
    T excl = __shfl_up(incl, 1, 64);
    if (lane == 0) excl = identity;
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

// ---------------------------------------------------------------------------
// Setup: COO (1-based) -> pool + per-column state
// ---------------------------------------------------------------------------

__global__ void k_ctrl_init(Ctrl *c, unsigned long long cap) {
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
    }
}

__device__ __forceinline__ int lower_bound_i(const int *a, int n, int key) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// One thread per column: column extent by binary search in the sorted cols,
// initial state (openph.cu:86-90 fills, kernels.cu:40-52 compute_low).
__global__ void k_init_cols(Dev d, const int *__restrict__ rows1, const int *__restrict__ cols1) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d.m) return;
    int lo = lower_bound_i(cols1, d.nnz, j + 1);
    int hi = lower_bound_i(cols1, d.nnz, j + 2);
    if (hi < lo) hi = lo;  // unsorted input; flagged by k_entries
    int n = hi - lo;
    d.off[j] = (unsigned)lo;
    d.len[j] = n;
    int l = n ? rows1[hi - 1] - 1 : -1;
    if (l < -1 || l >= d.m) l = -1;  // out of range; flagged by k_entries
    d.low[j] = l;
    d.arglow[j] = -1;
    d.classes[j] = 0;
    d.ess[j] = 1;
    d.left[j] = kIntMax;
    d.beta[j] = -1;
    d.first[j] = kIntMax;
    d.clear[j] = 0;
    d.newpiv[j] = -1;
    if (n) atomicMax(&d.ctrl->maxlen, n);
}

// One thread per entry: validation, pool copy (0-based), left[] by atomicMin.
__global__ void k_entries(Dev d, const int *__restrict__ rows1, const int *__restrict__ cols1) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= d.nnz) return;
    int c = cols1[e] - 1, r = rows1[e] - 1;
    int bad = 0;
    if (c < 0 || c >= d.m || r < 0 || r >= c) bad |= kErrRange;
    if (e > 0) {
        int pc = cols1[e - 1] - 1;
        if (c < pc) bad |= kErrOrder;
        else if (c == pc && rows1[e - 1] - 1 >= r) bad |= kErrRows;
    }
    if (bad) atomicOr(&d.ctrl->err, bad);
    d.pool[e] = r;
    if (!(bad & kErrRange)) atomicMin(&d.left[r], c);
}

__global__ void k_beta(Dev d) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.m) return;
    int l = d.left[i];
    if (l != kIntMax) atomicMax(&d.beta[l], i);
}

void launch_setup(const Dev &d, const int *rows1, const int *cols1, hipStream_t s) {
    int gm = (d.m + kBlock - 1) / kBlock, ge = (d.nnz + kBlock - 1) / kBlock;
    if (d.m) k_init_cols<<<gm, kBlock, 0, s>>>(d, rows1, cols1);
    if (d.nnz) k_entries<<<ge, kBlock, 0, s>>>(d, rows1, cols1);
}

// Pseudo-dimension slots: pseudo-dim = len - 1 (dimensions.cu:61-63). The
// distinct values present are numbered in increasing order.
__global__ void k_presence(Dev d) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d.m) return;
    d.slotmap[d.len[j]] = 1;
}

__global__ void k_slotmap(Dev d) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int n = 0, mx = d.ctrl->maxlen;
    for (int i = 0; i <= mx; i++) {
        if (d.slotmap[i]) d.slotmap[i] = n++;
        else d.slotmap[i] = -1;
    }
    d.ctrl->nslots = n;
    if (n > kMaxSlots) atomicOr(&d.ctrl->err, kErrSlots);
}

__global__ void k_assign_slot(Dev d) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d.m) return;
    int s = d.slotmap[d.len[j]];
    d.slot[j] = s < 0 ? 0 : (s >= kMaxSlots ? kMaxSlots - 1 : s);
}

__global__ void k_zero_i(int *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

void launch_slots(const Dev &d, hipStream_t s) {
    int gm = (d.m + kBlock - 1) / kBlock;
    int g1 = (d.m + 1 + kBlock - 1) / kBlock;
    k_zero_i<<<g1, kBlock, 0, s>>>(d.slotmap, d.m + 1);
    k_beta<<<gm, kBlock, 0, s>>>(d);
    k_presence<<<gm, kBlock, 0, s>>>(d);
    k_slotmap<<<1, 64, 0, s>>>(d);
    k_assign_slot<<<gm, kBlock, 0, s>>>(d);
}

// ---------------------------------------------------------------------------
// Phase 0: mark_pivots_and_clear (pms_functions.cu:33-53)
// ---------------------------------------------------------------------------

__global__ void k_phase0(Dev d) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d.m) return;
    int lj = d.low[j];
    if (lj > -1 && lj == d.beta[j]) {
        d.classes[j] = -1;
        d.arglow[lj] = j;
        d.clear[lj] = 1;
        d.ess[lj] = 0;
        d.ess[j] = 0;
    }
}

// clear_positives (pms_functions.cu:131-140) and reset of the first-occurrence
// table for the next Phase I.
__global__ void k_apply1(Dev d) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.m) return;
    if (d.clear[t]) {
        d.low[t] = -1;
        d.classes[t] = 1;
        d.clear[t] = 0;
    }
    d.first[t] = kIntMax;
}

void launch_phase0(const Dev &d, hipStream_t s) {
    int gm = (d.m + kBlock - 1) / kBlock;
    k_phase0<<<gm, kBlock, 0, s>>>(d);
    k_apply1<<<gm, kBlock, 0, s>>>(d);
}

void launch_apply1(const Dev &d, hipStream_t s) {
    int gm = (d.m + kBlock - 1) / kBlock;
    k_apply1<<<gm, kBlock, 0, s>>>(d);
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
// ---------------------------------------------------------------------------

__global__ void k_first(Dev d) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d.m) return;
    int lj = d.low[j];
    if (lj > -1) atomicMin(&d.first[lj], j);
}

// Per tile of kTile columns: max duplicate low per slot, and the global
// "some low occurs twice" flag (is_reduced, pms_functions.cu:7-27).
__global__ void __launch_bounds__(kBlock) k_tile(Dev d) {
    __shared__ int smax[kMaxSlots];
    __shared__ int sdup;
    if (threadIdx.x < kMaxSlots) smax[threadIdx.x] = -1;
    if (threadIdx.x == 0) sdup = 0;
    __syncthreads();
    const int base = blockIdx.x * kTile + threadIdx.x * kItems;
    int dupseen = 0;
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        int j = base + k;
        if (j < d.m) {
            int lj = d.low[j];
            if (lj > -1 && d.first[lj] < j) {
                atomicMax(&smax[d.slot[j]], lj);
                dupseen = 1;
            }
        }
    }
    if (dupseen) sdup = 1;
    __syncthreads();
    if (threadIdx.x < kMaxSlots) d.tile_max[blockIdx.x * kMaxSlots + threadIdx.x] = smax[threadIdx.x];
    if (threadIdx.x == 0 && sdup) atomicOr(&d.ctrl->any_dup, 1);
}

// Exclusive max-scan of the tile maxima over tiles, per slot. One block.
__global__ void __launch_bounds__(1024) k_tilescan(Dev d) {
    __shared__ int lds[16];
    const int nslots = d.ctrl->nslots;
    for (int s = 0; s < nslots && s < kMaxSlots; s++) {
        int running = -1;
        for (int c0 = 0; c0 < d.ntiles; c0 += blockDim.x) {
            int t = c0 + threadIdx.x;
            int v = t < d.ntiles ? d.tile_max[t * kMaxSlots + s] : -1;
            int tot;
            int ex = block_exclusive_scan(v, MaxOp<int>(), -1, lds, &tot);
            if (t < d.ntiles) d.tile_pre[t * kMaxSlots + s] = ex > running ? ex : running;
            running = tot > running ? tot : running;
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_mark(Dev d) {
    __shared__ int lds[kBlock / 64];
    __shared__ int exl[kMaxSlots][kBlock];
    const int nslots = d.ctrl->nslots;
    const int base = blockIdx.x * kTile + threadIdx.x * kItems;
    int lo[kItems], sl[kItems], dup[kItems], fst[kItems];
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        int j = base + k;
        lo[k] = -1;
        sl[k] = 0;
        dup[k] = -1;
        fst[k] = 0;
        if (j < d.m) {
            lo[k] = d.low[j];
            sl[k] = d.slot[j];
            if (lo[k] > -1) {
                int f = d.first[lo[k]];
                fst[k] = (f == j);
                if (f < j) dup[k] = lo[k];
            }
        }
    }
#pragma unroll
    for (int s = 0; s < kMaxSlots; s++) {
        if (s < nslots) {
            int agg = -1;
#pragma unroll
            for (int k = 0; k < kItems; k++)
                if (sl[k] == s && dup[k] > agg) agg = dup[k];
            int tot;
            int ex = block_exclusive_scan(agg, MaxOp<int>(), -1, lds, &tot);
            int tp = d.tile_pre[blockIdx.x * kMaxSlots + s];
            exl[s][threadIdx.x] = ex > tp ? ex : tp;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        int j = base + k;
        if (j >= d.m || lo[k] < 0 || !fst[k]) continue;
        int c = exl[sl[k]][threadIdx.x];
#pragma unroll
        for (int q = 0; q < k; q++)
            if (sl[q] == sl[k] && dup[q] > c) c = dup[q];
        if (d.classes[j] == 0 && lo[k] > c) {
            int L = lo[k];
            atomicAdd(&d.ctrl->marks, 1);
            d.arglow[L] = j;
            d.classes[j] = -1;
            d.clear[L] = 1;
            d.ess[L] = 0;
            d.ess[j] = 0;
        }
    }
}

void launch_phase1_detect(const Dev &d, hipStream_t s) {
    int gm = (d.m + kBlock - 1) / kBlock;
    k_first<<<gm, kBlock, 0, s>>>(d);
    k_tile<<<d.ntiles, kBlock, 0, s>>>(d);
}

void launch_phase1_mark(const Dev &d, hipStream_t s) {
    k_tilescan<<<1, 1024, 0, s>>>(d);
    k_mark<<<d.ntiles, kBlock, 0, s>>>(d);
}

// ---------------------------------------------------------------------------
// Phase II (pms_functions.cu:87-112): reduce every unfinished column by the
// pivots known at the start of the phase.
//
// Pivots found here by the left rule (left[low_j] == j, the column is final)
// go to newpiv[] and are published to arglow[] after the phase, so every
// arglow[] entry read here names a column that is final and not written
// during the phase. clear[] / ess[] writes only ever store 1 / 0 and are
// facts about the matrix (a non-negative low of a partially reduced column
// is a positive, paired index), so their order does not matter.
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

__global__ void __launch_bounds__(kBlock) k_phase2_lane(Dev d) {
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
    bool dirty = false, ovf = false;
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
                    // the pivot's stored column does not end in its low:
                    // inconsistent state, stop instead of looping
                    atomicOr(&d.ctrl->err, kErrState);
                    active = false;
                } else if (lj < 0) {
                    active = false;
                } else {
                    if (d.left[lj] == j) {
                        d.newpiv[lj] = j;
                        d.classes[j] = -1;
                        active = false;
                    }
                    d.clear[lj] = 1;
                    if (active) {
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
        d.low[j] = lj;
    }
    if (ovf) d.ctrl->overflow = 1;
    // wave-reduce the counters, one atomic per wave
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

void launch_phase2(const Dev &d, int lanes, hipStream_t s) {
    (void)lanes;
    int gm = (d.m + kBlock - 1) / kBlock;
    k_phase2_lane<<<gm, kBlock, 0, s>>>(d);
}

// After Phase II: publish new pivots, clear_positives, ess_hat
// (kernels.cu:113-124) and update_classes (kernels.cu:93-101). Every thread
// touches its own column first, so one pass suffices.
__global__ void k_apply2(Dev d) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.m) return;
    int np = d.newpiv[t];
    if (np > -1) {
        d.arglow[t] = np;
        d.newpiv[t] = -1;
    }
    if (d.clear[t]) {
        d.low[t] = -1;
        d.classes[t] = 1;
        d.clear[t] = 0;
    }
    int lt = d.low[t];
    if (lt > -1) d.ess[lt] = 0;
    int a = d.arglow[t];
    if (a > -1) {
        d.ess[a] = 0;
        d.ess[t] = 0;
        d.classes[a] = -1;
        d.classes[t] = 1;
    }
}

void launch_apply2(const Dev &d, hipStream_t s) {
    int gm = (d.m + kBlock - 1) / kBlock;
    k_apply2<<<gm, kBlock, 0, s>>>(d);
}

// ---------------------------------------------------------------------------
// Convergence: ess_hat_final (kernels.cu:103-111) and the 1-based output
// (openph.cu:195-199).
// ---------------------------------------------------------------------------

__global__ void k_final_ess(Dev d) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.m) return;
    int lt = d.low[t];
    if (lt > -1) {
        d.ess[t] = 0;
        d.ess[lt] = 0;
    }
}

__global__ void k_output(Dev d, int *__restrict__ low1, int *__restrict__ ess) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.m) return;
    low1[t] = d.low[t] + 1;
    ess[t] = d.ess[t];
}

void launch_final(const Dev &d, int *low1_out, int *ess_out, hipStream_t s) {
    int gm = (d.m + kBlock - 1) / kBlock;
    if (!d.m) return;
    k_final_ess<<<gm, kBlock, 0, s>>>(d);
    k_output<<<gm, kBlock, 0, s>>>(d, low1_out, ess_out);
}

// ---------------------------------------------------------------------------
// Pool compaction: live columns (not cleared) are copied, in column order,
// to the other pool; cleared columns are dropped (they are never read
// again: not reduced, never a pivot).
// ---------------------------------------------------------------------------

__device__ __forceinline__ unsigned live_len(const Dev &d, int j) {
    return (j < d.m && d.classes[j] != 1) ? (unsigned)d.len[j] : 0u;
}

__global__ void __launch_bounds__(kBlock) k_scan_tiles(Dev d) {
    __shared__ unsigned long long lds[kBlock / 64];
    const int base = blockIdx.x * kTile + threadIdx.x * kItems;
    unsigned long long sum = 0;
#pragma unroll
    for (int k = 0; k < kItems; k++) sum += live_len(d, base + k);
    unsigned long long tot;
    block_exclusive_scan(sum, SumOp<unsigned long long>(), 0ull, lds, &tot);
    if (threadIdx.x == 0) d.scan_tmp[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(1024) k_scan_partials(Dev d) {
    __shared__ unsigned long long lds[16];
    unsigned long long running = 0;
    for (int c0 = 0; c0 < d.ntiles; c0 += blockDim.x) {
        int t = c0 + threadIdx.x;
        unsigned long long v = t < d.ntiles ? d.scan_tmp[t] : 0ull;
        unsigned long long tot;
        unsigned long long ex = block_exclusive_scan(v, SumOp<unsigned long long>(), 0ull, lds, &tot);
        __syncthreads();
        if (t < d.ntiles) d.scan_tmp[t] = ex + running;
        running += tot;
    }
    if (threadIdx.x == 0) {
        d.scan_tmp[d.ntiles] = running;
        d.ctrl->pool_top = running;
    }
}

__global__ void __launch_bounds__(kBlock) k_scan_apply(Dev d) {
    __shared__ unsigned long long lds[kBlock / 64];
    const int base = blockIdx.x * kTile + threadIdx.x * kItems;
    unsigned v[kItems];
    unsigned long long sum = 0;
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        v[k] = live_len(d, base + k);
        sum += v[k];
    }
    unsigned long long tot;
    unsigned long long ex = block_exclusive_scan(sum, SumOp<unsigned long long>(), 0ull, lds, &tot);
    ex += d.scan_tmp[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        int j = base + k;
        if (j < d.m) d.off2[j] = (unsigned)ex;
        ex += v[k];
    }
}

__global__ void k_copy_cols(Dev d) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d.m) return;
    unsigned n = live_len(d, j);
    const int *src = d.pool + d.off[j];
    int *dst = d.pool2 + d.off2[j];
    for (unsigned i = 0; i < n; i++) dst[i] = src[i];
    if (n == 0) d.len[j] = 0;
}

void launch_compact_scan(const Dev &d, hipStream_t s) {
    k_scan_tiles<<<d.ntiles, kBlock, 0, s>>>(d);
    k_scan_partials<<<1, 1024, 0, s>>>(d);
    k_scan_apply<<<d.ntiles, kBlock, 0, s>>>(d);
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

void launch_ctrl_init(const Dev &d, unsigned long long cap, hipStream_t s) {
    k_ctrl_init<<<1, 64, 0, s>>>(d.ctrl, cap);
}

__global__ void k_reset_flags(Ctrl *c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        c->any_dup = 0;
        c->overflow = 0;
    }
}

void launch_reset_flags(const Dev &d, hipStream_t s) {
    k_reset_flags<<<1, 64, 0, s>>>(d.ctrl);
}

}  // namespace oph
