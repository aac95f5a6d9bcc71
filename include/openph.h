/*
 * openph - MI355X (gfx950) Z/2 boundary-matrix reduction, C ABI.
 *
 * Conventions shared with the reference (src/cuda/openph.cu:7-27):
 *   - the boundary matrix is an m x m Z/2 matrix given as 1-based COO
 *     triplets (rows1[e], cols1[e]), e < nnz, column-major with rows
 *     strictly ascending inside a column and columns non-decreasing
 *     (src/cuda/openph.cu:44, src/cuda/kernels.cu:27,42);
 *   - every entry satisfies 1 <= rows1[e] < cols1[e] <= m (a filtration
 *     order: the reference assumes it through its clearing rules and never
 *     checks it; here it is validated);
 *   - low is returned 1-based with 0 meaning "the reduced column is zero"
 *     (src/cuda/openph.cu:195-198), ess[j] = 1 iff cell j is essential
 *     (src/cuda/openph.cu:199).
 *
 * Results are unique for a given matrix: every algorithm string returns the
 * same low and ess arrays.
 */
#ifndef OPENPH_H
#define OPENPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes returned by the v2 entry points (0 = success). */
#define OPENPH_OK 0
#define OPENPH_EINVAL -1      /* bad argument or malformed matrix            */
#define OPENPH_EUNSORTED -2   /* COO not column-major / rows not ascending   */
#define OPENPH_EALGO -3       /* unknown algorithm string                    */
#define OPENPH_ENOMEM -4      /* device memory (or the pool budget) exhausted */
#define OPENPH_EDEVICE -5     /* HIP runtime error                           */
#define OPENPH_EITER -6       /* iteration cap reached                       */
#define OPENPH_EDIMS -7       /* more than OPENPH_MAX_DIMS pseudo-dimensions */
#define OPENPH_EBUSY -8       /* served queue full / admission refused       */
#define OPENPH_ESCRATCH -9    /* input beyond the static scratch of
                                 openph_pms_device_noalloc                   */

#define OPENPH_MAX_DIMS 16

/*
 * Drop-in replacement for the reference's host entry point
 * (src/cuda/openph.cu:29-35, called by the mex gateway
 * src/matlab/ph/ph.cu:84-87). Same parameter order and types.
 *
 *   algstr      "pms" | "standard" | "twist" | "ph_row" |
 *               "standard_parallel" | "twist_parallel" | "ph_row_parallel"
 *               (src/cuda/algorithms/algorithm_factory.cu:10-148). Each
 *               runs its own algorithm on the device: pms iterations;
 *               standard and twist one column at a time in the reference's
 *               order (a _parallel name runs the same order); ph_row and
 *               ph_row_parallel one row at a time, from the last. low is
 *               the same for all on a boundary matrix; ess is each
 *               algorithm's own estimate, as in the reference. For all but
 *               pms, num_iters is steps + 1, time_track[i] the device time
 *               of step i, and the err_* trackers are left as given.
 *   col_width   accepted and ignored: columns live in an adaptive device
 *               pool, there is no fixed p = col_width * max_nnz.
 *   h_low_true  1-based true low array (0 = none) used only by the
 *               trackers; NULL skips the err_* trackers.
 *   outputs     h_low[m], h_ess[m]; err_linf, err_lone, err_redu, err_ess,
 *               time_track: m+1 floats each (index = iteration);
 *               num_iters[0].
 *   trackers    for iterations i = 0 .. num_iters-1, as
 *               src/cuda/algorithms/tracker.cu:92-134 defines them (0-based
 *               low and low_true): err_lone[i] = sum_j |low_j - low_true_j|,
 *               err_linf[i] = max_j |low_j - low_true_j|, err_redu[i] =
 *               #{j : classes_j == 0} / m, err_ess[i] = #essential(low_true)
 *               / #{j : ess_j == 1} (1 when equal), time_track[i] = ms of
 *               iteration i (time_track[0] = 0). Then every one of the m+1
 *               entries of err_lone and err_linf is divided by the l1 and
 *               linf norms of h_low_true, as src/cuda/openph.cu:188-193 does
 *               (a zero norm gives IEEE inf / nan, as there). The sums are
 *               exact integers rounded once to float; the reference sums
 *               floats in cuBLAS's unspecified order.
 *
 * The reference has no error channel. Here errors never exit the process:
 * on failure num_iters[0] is set to the negative status code and h_low is
 * left as the caller initialised it (the mex gateway fills -1,
 * src/matlab/ph/ph.cu:72).
 */
void openph(char *algstr, int *h_rows_in, int *h_cols_in, int m, int col_width,
            int *h_low_true, int nnz, int *h_low, int *h_ess, float *err_linf,
            float *err_lone, float *err_redu, float *err_ess, float *time_track,
            int *num_iters, int gpu_number);

/* Options for the v2 / device entry points. Zero-initialise, then set. */
typedef struct openph_options {
    int device;              /* HIP device ordinal                           */
    int max_iters;           /* pms iteration cap (0 = 1,000,000)            */
    int64_t pool_elems;      /* initial column-pool size in int32 entries,
                                counted from the initial layout (every
                                column's entries, slot_elems at least): a
                                smaller value leaves no headroom
                                (0 = automatic)                              */
    int64_t pool_max_elems;  /* largest pool the reduction may grow to
                                (0 = 2^32 - 1 entries, 16 GiB)               */
    int lanes_per_column;    /* Phase II lanes per column: 0 = automatic,
                                else 1, 4, 8, 16, 32 or 64                   */
    int track;               /* 1 = fill the per-iteration trackers          */
    int slot_elems;          /* minimum pool segment per column (0 = 32)     */
    int group_lds;           /* LDS entries per lane-group buffer (0 = the
                                default of lanes_per_column); the compiled
                                shapes are {4: 32, 64}, {8: 32, 64, 72, 80, 96, 128},
                                {16: 64, 128, 256}, {32: 256, 512},
                                {64: 512, 1024}                              */
    int schedule;            /* how the host drives the pms iterations:
                                0 = automatic,
                                1 = read the device state after Phase I and
                                    after each iteration (Phase II grids
                                    sized by the work; used when trackers are
                                    filled),
                                2 = queue chunks of iterations on the stream
                                    and read the state once per chunk,
                                3 = the same chunks replayed from a cached
                                    hipGraph                                 */
    int chunk_iters;         /* iterations per chunk after the first one
                                (schedules 2 and 3; 0 = 3). The first chunk
                                is as long as the last reduction of a matrix
                                of the same size took.                       */
    int time_phase2;         /* 1 = time every Phase II pass into
                                openph_stats.phase2_ms: HIP events around it
                                (schedules 1 and 2) or device-clock stamps
                                (schedule 3). Each timed pass adds two
                                markers to the stream (about 5 us each), so
                                throughput runs leave it 0.                  */
} openph_options;

/* Statistics of one reduction. */
typedef struct openph_stats {
    int num_iters;           /* the reference's num_iters (iterations + 1)   */
    int compactions;         /* pool compactions performed                   */
    int64_t additions;       /* column additions performed                   */
    int64_t merge_elems;     /* sum over additions of |a| + |b| + |a xor b|  */
    int64_t pool_peak;       /* largest pool fill seen, in int32 entries     */
    float device_ms;         /* device time of the reduction: set-up, loop
                                and outputs                                  */
    float phase2_ms;         /* summed device time of the Phase II launches
                                (0 unless options.time_phase2)               */
    int phase2_launches;     /* number of Phase II launches                  */
    int m;                   /* columns reduced                              */
    int overflows;           /* Phase II passes cut short by a full pool     */
    int lanes;               /* Phase II lanes per column that were used     */
    int spills;              /* columns whose additions left LDS             */
    int group_lds;           /* LDS entries per lane-group buffer used       */
    int handoffs;            /* columns continued by the block kernel        */
    int schedule;            /* schedule used (1, 2 or 3, see openph_options) */
    int host_reads;          /* control-block reads in the iteration loop    */
} openph_stats;

/*
 * Status-returning variant of openph() on host buffers. low1 and ess are
 * m-element outputs (1-based low, 0/1 ess); stats may be NULL.
 * trackers (may be NULL) receives 5 x (m+1) floats laid out as
 * err_linf | err_lone | err_redu | err_ess | time_track and is only written
 * when opt->track is set: time_track always, the four err_* arrays when
 * low_true1 is given (same rules as openph()).
 */
int openph_v2(const char *algstr, const int *rows1, const int *cols1, int m, int nnz,
              const int *low_true1, const openph_options *opt, int *low1, int *ess,
              float *trackers, openph_stats *stats);

/*
 * Device-resident entry point: all array arguments are device pointers on
 * opt->device, the work is queued on `stream` (a hipStream_t, NULL = the
 * default stream) and the call returns when the results are in low1/ess.
 * Library workspace is allocated on first use and kept for later calls.
 */
int openph_pms_device(const int *rows1, const int *cols1, int m, int nnz,
                      int *low1, int *ess, void *stream, const openph_options *opt,
                      openph_stats *stats);

/*
 * openph_pms_device without any allocation: no counterpart in the reference
 * (whose openph() allocates and frees every buffer per call,
 * src/cuda/openph.cu:72-230). For callers that cannot allocate device memory:
 * code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc. Arguments as openph_pms_device (device pointers; the caller owns
 * inputs and outputs). Every working buffer lives in one static device array
 * of about 81 MB that each GPU holds from the library's load until the
 * process ends. Limits: m <= 196,608 columns and nnz <= 589,824 entries
 * (OPENPH_ESCRATCH beyond them, before any work), and column pools of
 * 8,388,608 entries (OPENPH_ENOMEM when the initial layout, nnz +
 * slot_elems per column, or the live columns do not fit). The iterations are
 * always queued (schedule 1 or 2; a request for graphs queues them), the
 * trackers are not available, and stats carry counters but no times. Calls on
 * one device share the array: they are serialised by a lock, and each returns
 * once low1/ess are written; the array keeps the last call's contents.
 */
int openph_pms_device_noalloc(const int *rows1, const int *cols1, int m, int nnz, int *low1,
                              int *ess, void *stream, const openph_options *opt,
                              openph_stats *stats);

/*
 * Batched reduction of `count` independent matrices in one pass of the
 * device engine (block-diagonal offsets). The inputs are concatenated:
 * matrix k has m_k = ms[k] columns and nnz_k = nnzs[k] entries and its COO
 * (1-based, local to the matrix) occupies rows1/cols1[nnz_off_k ..], with
 * nnz_off_k the prefix sum of nnzs. Outputs are concatenated the same way
 * over the m_k (low local and 1-based). Host pointers.
 */
int openph_pms_batch(int count, const int *ms, const int *nnzs, const int *rows1,
                     const int *cols1, const openph_options *opt, int *low1, int *ess,
                     openph_stats *stats);

/* Same as openph_pms_batch with device pointers for rows1/cols1/low1/ess
 * (ms and nnzs stay on the host). */
int openph_pms_batch_device(int count, const int *ms, const int *nnzs, const int *rows1,
                            const int *cols1, int *low1, int *ess, void *stream,
                            const openph_options *opt, openph_stats *stats);

/* One candidate measured by the autotuner. */
typedef struct openph_tune_row {
    int lanes_per_column;    /* Phase II lanes per column                    */
    int group_lds;           /* LDS entries per lane-group buffer            */
    int slot_elems;          /* minimum pool segment per column              */
    int schedule;            /* loop schedule (1, 2 or 3)                    */
    float wall_ms;           /* median wall time of one batched reduction    */
    int same;                /* 1: low/ess equal the reference candidate's   */
} openph_tune_row;

/*
 * Autotuner for the batched path (no counterpart in the reference, whose
 * threads per column and slot width p are fixed by the caller). Reduces the
 * batch held on the device (arguments as openph_pms_batch_device) with
 * candidate settings, one untimed run and `reps` timed runs each (median of
 * the host wall time, inputs resident), in three stages:
 *   1. Phase II shape: every compiled {lanes_per_column, group_lds} and one
 *      lane per column, with base's pool granularity and schedule;
 *   2. pool granularity: slot_elems in {8, 16, 32, 64} with the best shape;
 *   3. loop schedule: 1, 2 and 3 with the best shape and granularity.
 * Every candidate's low/ess must equal the first candidate's (base
 * options); a difference returns OPENPH_EDEVICE. *best receives base with
 * the fastest setting; pass it to later calls (or to a server's
 * openph_server_options.reduce). report (may be NULL) receives up to
 * max_report rows in the order measured, *nreport their number.
 * low1/ess end with the outputs of the last run.
 */
int openph_autotune_batch_device(int count, const int *ms, const int *nnzs, const int *rows1,
                                 const int *cols1, int *low1, int *ess, void *stream,
                                 const openph_options *base, int reps, openph_options *best,
                                 openph_tune_row *report, int max_report, int *nreport);

/*
 * Served reduction (no counterpart in the reference, which is a blocking
 * MATLAB call). A server owns `workers` device workspaces, each with its own
 * HIP stream and thread, sized at creation for worker_bytes. Each worker keeps
 * one running batch of requests that are reduced by the same pms iterations:
 * between chunks of two iterations it releases every request whose own lows
 * no longer repeat, and it admits queued requests of the batch's class
 * (input columns of at most 4 entries, or longer) while their columns, COO
 * entries and pool footprint fit its budget and the batch keeps at most
 * OPENPH_MAX_DIMS distinct column lengths. Batches of different workers run
 * concurrently on the device. A request too large for an empty batch runs
 * alone; the members of a batch that fails are run again alone.
 */
typedef struct openph_server openph_server;

typedef struct openph_server_options {
    int device;              /* HIP device ordinal                           */
    int workers;             /* concurrent batches (0 = 4, at most 16), each
                                on a stream of its own. HIP gives a process
                                4 hardware queues unless GPU_MAX_HW_QUEUES
                                (read when HIP starts) asks for more; streams
                                beyond them share queues and their kernels
                                serialise, so a serving process raises it
                                (16 for 4 workers beside other streams)      */
    int64_t worker_bytes;    /* device bytes a batch may claim (0 = 8 GiB)   */
    int max_batch;           /* requests per batch (0 = 1024)                */
    int max_queue;           /* queued requests before OPENPH_EBUSY
                                (0 = 65536)                                  */
    openph_options reduce;   /* options of each reduction                    */
    int64_t stage_bytes;     /* pinned host memory that holds the queued
                                requests' COO, validated and ready for one
                                DMA each (0 = 2 GiB, < 0 = none); requests
                                that do not fit are staged by the workers   */
} openph_server_options;

typedef struct openph_request_info {
    double submit_s;         /* seconds on the server's steady clock         */
    double start_s;          /* it joined a batch                            */
    double done_s;           /* its results were on the host                 */
    int batch_size;          /* requests that shared its batch up to its
                                release                                       */
    int worker;              /* worker that ran it                           */
    int num_iters;           /* iterations it took part in                   */
    int status;              /* 0 or a negative OPENPH_E* code               */
    int m;                   /* columns of the request                       */
} openph_request_info;

int openph_server_create(const openph_server_options *opt, openph_server **out);
/* Queue one matrix (host 1-based COO, copied); returns a ticket. Never waits
 * for the device; OPENPH_EBUSY when the queue is full, OPENPH_EDIMS when its
 * columns have more than OPENPH_MAX_DIMS distinct lengths. */
int openph_server_submit(openph_server *srv, const int *rows1, const int *cols1, int m, int nnz,
                         int64_t *ticket);
/* Queue a matrix the worker builds on the device from its input (host
 * arrays, copied): the VR 2-skeleton of n points in R^dim, or the cubical
 * lower-star filtration of an s0 x s1 x s2 grid (see openph_vr2_device and
 * openph_cubical_device). Non-finite points and NaN field values are
 * refused. */
int openph_server_submit_vr(openph_server *srv, const double *points, int n, int dim,
                            int64_t *ticket);
int openph_server_submit_cubical(openph_server *srv, const float *field, int s0, int s1, int s2,
                                 int64_t *ticket);
/* One request of any kind, with optional zero-copy input and caller-owned
 * outputs (openph_server_submit_request). The fields of the other kinds are
 * ignored. */
#define OPENPH_REQ_COO 0      /* rows1/cols1/m/nnz, as openph_server_submit  */
#define OPENPH_REQ_VR 1       /* points/n/dim, as openph_server_submit_vr    */
#define OPENPH_REQ_CUBICAL 2  /* field/s0/s1/s2, as openph_server_submit_cubical */
#define OPENPH_BORROW 1       /* flags: the input arrays are read in place    */
typedef struct openph_request {
    int kind;                /* OPENPH_REQ_*                                 */
    int flags;               /* 0 (inputs copied at submit) or OPENPH_BORROW */
    const int *rows1, *cols1;
    int m, nnz;
    const double *points;
    int n, dim;
    const float *field;
    int s0, s1, s2;
    int *low1, *ess;         /* NULL, or host buffers of the request's m
                                columns that the worker writes its outputs
                                into (both or neither)                       */
} openph_request;
/* Queue a request described by `req`; validation, errors and the queue are
 * those of the matching submit call. With OPENPH_BORROW nothing is copied at
 * submit: the input arrays are read by the worker when the request joins a
 * batch, so they must stay valid and unchanged until openph_server_wait has
 * returned for the ticket (or openph_server_destroy has). A lent COO is
 * validated after submit has returned, by the server's intake threads: a
 * malformed one completes with its error code, which wait returns. Output
 * buffers given in req->low1/ess receive the results before the request is
 * released and must stay valid for as long; wait may then pass NULL for its
 * own. */
int openph_server_submit_request(openph_server *srv, const openph_request *req, int64_t *ticket);
/* Block until the ticket's reduction is done; low1/ess (host, m entries) get
 * its outputs. A ticket can be waited for once: a second wait, also one that
 * starts while the first blocks, returns OPENPH_EINVAL. */
int openph_server_wait(openph_server *srv, int64_t ticket, int *low1, int *ess,
                       openph_request_info *info);
int openph_server_stats(openph_server *srv, int64_t *queued, int64_t *batches,
                        int64_t *completed);
/* Finishes queued requests, then stops the workers and frees the server. */
int openph_server_destroy(openph_server *srv);

/*
 * Filtrations built on the device (no counterpart in the reference, which
 * builds its matrices in Javaplex on the host,
 * numerics/utils/pointcloud_to_vr_stream.m, numerics/utils/stream2cmo.m:1-26).
 * They write the 1-based COO boundary matrix in filtration order, ready for
 * openph_pms_device, plus each column's filtration value and dimension
 * (values and dims may be NULL). All array arguments are device pointers on
 * `device`; the call returns when the outputs are written.
 *
 * VR 2-skeleton at full radius of n points in R^dim (dim <= 7, row-major
 * doubles): vertices, edges and triangles ordered by value (0, length,
 * longest edge), then dimension, then vertex tuple. Lengths are computed as
 * sqrt(sum_k (p_ik - p_jk)^2), summed from k = 0 with every operation
 * rounded once. m = n + C(n,2) + C(n,3), nnz = 2 C(n,2) + 3 C(n,3).
 */
int openph_vr2_size(int n, int64_t *m, int64_t *nnz);
int openph_vr2_device(const double *points, int n, int dim, int *rows1, int *cols1,
                      double *values, int *dims, void *stream, int device);

/*
 * Lower-star filtration of the cubical complex of an s0 x s1 x s2 grid of
 * float values (row-major): all (2 s0 - 1)(2 s1 - 1)(2 s2 - 1) cells, value
 * of a cell = max over its vertices, ordered by value, then dimension, then
 * linear cell index. m must stay below 2^28.
 */
int openph_cubical_size(int s0, int s1, int s2, int64_t *m, int64_t *nnz);
int openph_cubical_device(const float *field, int s0, int s1, int s2, int *rows1, int *cols1,
                          float *values, int *dims, void *stream, int device);

/* Human-readable message for a status code. */
const char *openph_strerror(int status);

/* Library version string. */
const char *openph_version(void);

#ifdef __cplusplus
}
#endif

#endif /* OPENPH_H */
