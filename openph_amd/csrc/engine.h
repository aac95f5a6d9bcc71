// Internal interface between the host orchestration (openph.cpp) and the
// gfx950 kernels (kernels.hip). Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace oph {

constexpr int kMaxSlots = 16;      // distinct pseudo-dimensions supported
constexpr int kBlock = 256;        // threads per block for flat passes
constexpr int kItems = 4;          // columns per thread in the tiled scans
constexpr int kTile = kBlock * kItems;
constexpr int kIntMax = 0x7fffffff;
constexpr int kLongBlockLanes = 1024;  // lanes of the block kernels (long columns, serial)

// Device-side scalars shared by all kernels of one reduction. Lives in
// device memory; the host reads it back with one small copy.
constexpr int kStripes = 16;
// Iterations whose trackers and clock stamps a queued or replayed chunk keeps
// on the device (iteration i in entry i % kIterRing); a chunk holds at most
// half as many iterations, so its entries never collide.
constexpr int kIterRing = 128;

struct Ctrl {
    unsigned long long pool_top;   // next free pool entry
    unsigned long long pool_cap;   // pool capacity (entries)
    unsigned long long additions;  // column additions performed
    unsigned long long merge_elems;// sum |a|+|b|+|a xor b|
    int err;                       // validation / consistency error bits
    int any_dup;                   // a non-negative low occurs twice
    int overflow;                  // Phase II ran out of pool
    int nslots;                    // distinct pseudo-dimensions
    int maxlen;                    // longest input column
    int marks;                     // columns marked by Phase I (cumulative)
    int nwork;                     // columns queued for Phase II
    int spills;                    // Phase II columns that left LDS (cumulative)
    int nlong;                     // columns handed to the long-column pass
    int handoffs;                  // the same, cumulative
    // iteration control on the device, so that several iterations can be
    // queued (or replayed from a graph) between two reads of this block
    int halt;                      // kHalt* reason; 0 = keep iterating
    int iter;                      // current iteration (1 = the first)
    int max_iters;                 // iteration cap
    int stalls;                    // consecutive overflowing passes without progress
    int prev_marks;                // marks at the end of the previous iteration
    int overflows;                 // Phase II passes cut short by a full pool
    int suspended;                 // this iteration's Phase II was cut short and resumed
    unsigned long long prev_add;   // additions at the end of the previous iteration
    unsigned long long pool_peak;  // largest pool_top seen at an iteration end
    // trackers of one iteration (tracker.cu:92-134), as exact integers
    unsigned long long trk_l1;     // sum_j |low[j] - low_true[j]|
    int trk_linf;                  // max_j |low[j] - low_true[j]|
    int trk_unreduced;             // #{j : classes[j] == 0}
    long long trk_ess;             // #{j : ess[j] == 1}
    // Phase II blocks add their counters here, block b into stripe b % 16,
    // and the iteration's end (in k_apply2) folds the stripes into additions
    // / merge_elems / spills:
    // device-scope atomics on one address wait for each other
    unsigned long long st_add[kStripes];
    unsigned long long st_el[kStripes];
    int st_spill[kStripes];
};

// Why the iteration kernels stopped (Ctrl::halt). Every iteration kernel
// returns at once while halt is set; the host clears it to resume.
enum : int {
    kHaltNone = 0,
    kHaltConverged = 1,  // no low repeats: reduced (is_reduced, pms_functions.cu:7-27)
    kHaltCompact = 2,    // the pool overflowed or is 3/4 full: compact, then resume
    kHaltStuck = 3,      // lows repeat but nothing changed: internal fault
    kHaltIters = 4,      // iteration cap reached
    kHaltNoMem = 5,      // two overflowing passes in a row made no progress
    kHaltError = 6,      // err bits set
    kHaltSuspend = 7,    // Phase II ran out of pool: compact, finish the same pass, end the iteration
};

// Validation error bits (Ctrl::err).
enum : int {
    kErrRange = 1,      // index outside 1..m or row >= col
    kErrOrder = 2,      // columns decreasing
    kErrRows = 4,       // rows not strictly ascending inside a column
    kErrSlots = 8,      // more than kMaxSlots pseudo-dimensions
    kErrState = 16,     // internal consistency check failed
};

// All device buffers of one reduction. Sized for m columns and a pool.
struct Dev {
    int m = 0;
    int nnz = 0;
    int ntiles = 0;
    int slot0 = 8;      // minimum segment size of a column in the pool
    // Phase II worklists at most this long (0 = never) are taken by
    // 256-lane teams (k_phase2_wide) instead of the 1024-entry lane groups
    int wide_max = 0;
    // 1: the next iteration's k_first runs inside k_apply2 (k_apply2_first);
    // the loop is then started, and resumed after a halt, by launch_first
    int fuse_first = 0;
    // per column
    int *low = nullptr, *arglow = nullptr, *classes = nullptr, *ess = nullptr;
    int *left = nullptr, *beta = nullptr, *first = nullptr, *clear = nullptr;
    int *newpiv = nullptr, *slot = nullptr, *len = nullptr, *cap = nullptr;
    int *work = nullptr;                     // Phase II worklist
    int *longw = nullptr;                    // columns handed to the long-column pass
    int4 *pivdesc = nullptr;                 // per row: {pivot column, off, len, left}
    unsigned *off = nullptr, *off2 = nullptr;
    unsigned long long *scan_tmp = nullptr;  // tiled-scan partials
    int *slotmap = nullptr;                  // pseudo-dim + 1 -> slot
    int *tile_max = nullptr, *tile_pre = nullptr;
    int *pool = nullptr, *pool2 = nullptr;   // column pool (int32 rows)
    Ctrl *ctrl = nullptr;
    // a running batch of matrices (the server): column ranges of its members
    // [mstart[k], mstart[k+1]) and, per iteration, mdup[k] = 1 when a low of
    // member k repeats (null / 0 outside the server)
    int *mdup = nullptr;
    const int *mstart = nullptr;
    int nmem = 0;
    // per-iteration trackers and iteration times of the queued / graph
    // schedules, read once per chunk (null when not wanted)
    const int *ltrue = nullptr;              // 0-based low_true
    unsigned long long *trk_part = nullptr;  // k_track_part's per-block sums
    unsigned long long *trk_ring = nullptr;  // [kIterRing][4] per iteration
    unsigned long long *its = nullptr;       // [kIterRing][2] start / end clock
};

// ---- launchers (kernels.hip) ----
// Control block reset; for d.m > 0 also the length-presence map (slotmap).
void launch_ctrl_init(const Dev &d, unsigned long long cap, int max_iters, hipStream_t s);
// After the setup: pool capacity = max(requested, exact initial layout),
// at most `alloc`.
void launch_pool_cap(const Dev &d, unsigned long long requested, unsigned long long alloc,
                     hipStream_t s);
// COO -> per-column state, validation and pool layout (column j's segment:
// len + slot0 entries at its COO start + j * slot0), lengths present.
void launch_setup(const Dev &d, const int *rows1, const int *cols1, hipStream_t s);
// After launch_setup: malformed input makes every column empty and halts
// the loop; else left/beta and the slot numbers of the lengths present.
void launch_settle(const Dev &d, hipStream_t s);
// Slots of the columns, Phase 0 and its clears.
void launch_phase0(const Dev &d, hipStream_t s);
// iter + 1 and the per-iteration flags reset, unless halted. The pms loop
// does this in k_first; the serial algorithms use it to clear the overflow
// flag before they resume.
void launch_iter_begin(const Dev &d, hipStream_t s);
void launch_set_halt(const Dev &d, int value, hipStream_t s);
// Before a suspended Phase II pass runs again (kHaltSuspend, after the
// compaction): clears the halt, the overflow flag and the long-column list.
void launch_resume2(const Dev &d, hipStream_t s);
// Device clock (wall_clock64) of the current iteration into
// stamps[(2 * (iter - 1) + which) % cap], unless halted.
void launch_stamp(const Dev &d, unsigned long long *stamps, int which, int cap, hipStream_t s);
void launch_phase1_detect(const Dev &d, hipStream_t s);
// k_first alone: starts an iteration (Dev::fuse_first: the loop's first one,
// or the one resumed after a halt)
void launch_first(const Dev &d, hipStream_t s);
void launch_phase1_mark(const Dev &d, hipStream_t s);
void launch_apply1(const Dev &d, hipStream_t s);
// lanes: 1 = one lane per column over all columns; 4..64 = lane groups over
// the worklist with wlds LDS entries per buffer, followed by the long-column
// pass for the columns they hand off (wave teams when wlds < 1024, else
// 256-lane teams). grid = blocks of the groups; long_grid = CUs for the
// long-column pass (three of its blocks per CU).
void launch_phase2(const Dev &d, int lanes, int wlds, int grid, int long_grid, hipStream_t s);
// A running batch (the server): columns [a, d.m) of members appended to it,
// with their global 1-based COO entries [e0, e1): per-column state and pool
// layout after the current pool top, validation, left/beta, slots from the
// host's slot map (nslots of them), Phase 0 and its clears.
void launch_setup_range(const Dev &d, int a, int e0, int e1, const int *rows1, const int *cols1,
                        hipStream_t s);
void launch_slots_range(const Dev &d, int a, int nslots, hipStream_t s);
void launch_phase0_range(const Dev &d, int a, hipStream_t s);
// A member [a, b) whose lows no longer repeat: ess_hat_final on its range,
// its 1-based local low (low - a + 1, 0 = none) and ess into low1/ess, then
// its columns made inert (low -1, cleared, no pivots) so that later
// iterations and compactions skip them.
void launch_retire_range(const Dev &d, int a, int b, int *low1_out, int *ess_out, hipStream_t s);
// Batches: local 1-based COO of `count` matrices -> global block-diagonal
// COO (err gets bit 1 for an index outside its matrix), and global 1-based
// low -> local 1-based low.
void launch_batch_offsets(const int *rows1, const int *cols1, const int *noff, const int *moff,
                          int count, int nnz, int *rows_g, int *cols_g, int *err, hipStream_t s);
void launch_batch_local_low(int *low1, const int *moff, int count, int m, hipStream_t s);
bool phase2_shape_ok(int lanes, int wlds);
int phase2_default_lds(int lanes);
int phase2_blocks_per_cu(int lanes, int wlds);
// After Phase II: pivots published, clears applied, ess and classes; its
// first thread also ends the iteration (counters folded, Ctrl::halt set by
// the progress, stall, overflow and pool-fill checks).
void launch_apply2(const Dev &d, hipStream_t s);
// The serial algorithms (standard, twist): one workgroup reduces the columns
// order[0..n) in turn. On a full pool it stops with Ctrl::overflow = 1 and
// Ctrl::nwork = the position to resume from. stamps[k] gets the device clock
// after column k, *start (when given) before the first.
void launch_serial(const Dev &d, const int *order, int n, bool twist, unsigned long long *stamps,
                   unsigned long long *start, hipStream_t s);
// The row algorithms (ph_row, and ph_row_parallel with all_dims): lists of
// the columns of each low, then one workgroup over the rows from state[0]
// down. stamps[m-1-i] (may be null) gets the device clock after row i, start
// (may be null) the clock before the first row. A full pool stops the kernel
// with Ctrl::overflow = 1 and the resume point in state; Ctrl::nwork counts
// the additions and rows done by the launch.
struct RowLists {
    int *head, *next, *pivot, *plen, *cand, *state;
};
void launch_row_lists(const Dev &d, const RowLists &r, hipStream_t s);
void launch_ph_row(const Dev &d, const RowLists &r, bool all_dims, unsigned long long *stamps,
                   unsigned long long *start, hipStream_t s);
// ess = 1 except at the lows (ph_row_parallel's estimate).
void launch_ess_at_lows(const Dev &d, hipStream_t s);
// gated: only if the loop has converged (queued before the host knows)
void launch_final(const Dev &d, int *low1_out, int *ess_out, hipStream_t s, bool gated = false);
// 1-based low and ess out, with no final ess pass.
void launch_output(const Dev &d, int *low1_out, int *ess_out, hipStream_t s);
// ess_hat_final (kernels.cu:103-111) alone, and update_classes_final
// (kernels.cu:79-91) on classes[]: the state the reference tracks after its
// last iteration. Needs k_final_ess first; classes[] is not used afterwards.
void launch_final_ess(const Dev &d, hipStream_t s);
void launch_classes_final(const Dev &d, hipStream_t s);
// The trackers of the current state into Ctrl::trk_* (0-based low_true).
void launch_track(const Dev &d, const int *low_true0, hipStream_t s);
// Queued / graph schedules, after k_apply2: the end clock of the iteration
// (the first write after the host cleared the ring wins, so instances queued
// past a halt leave it alone) into its[], and the trackers of the state
// (d.ltrue) into trk_ring[iter % kIterRing]. Instances queued past a halt
// recompute the same entry from the same state. `again` overwrites the end
// clock: an iteration whose suspended Phase II pass has just been finished.
void launch_iter_end_stamp(const Dev &d, hipStream_t s, int again = 0);
void launch_track_iter(const Dev &d, hipStream_t s);
// Zeroes the clock ring before a chunk.
void launch_clear_stamps(const Dev &d, hipStream_t s);
int track_part_blocks(int m);
// Compaction, step 1: new offsets of the live columns in off2 and the live
// total in ctrl->pool_top (and scan_tmp[ntiles]).
void launch_compact_scan(const Dev &d, hipStream_t s);
// Compaction, step 2: copy into pool2, then swap pool/pool2 and off/off2.
void launch_compact_copy(Dev &d, hipStream_t s);

}  // namespace oph
