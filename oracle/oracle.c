/*
 * CPU oracle for the pms boundary-matrix reduction.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker that the GPU library is
 * compared against; it is linked only by tests/, by __graft_entry__.smoke()
 * and by bench.py's cpu_baseline leg, and never by the product library.
 *
 * All arrays are 0-based CSC: column j holds rows[colptr[j] .. colptr[j+1])
 * in ascending order. Outputs follow the reference's internal (0-based)
 * convention: low[j] = max row of reduced column j or -1, ess[j] in {0,1}.
 *
 * Functions and the reference code they restate:
 *   orc_standard  - src/cuda/algorithms/standard/standard.cu:18-54 with
 *                   reduce_col_host / left_to_right_host
 *                   (src/cuda/algorithms/kernels_host.cu:2-75) on the same
 *                   fixed-slot layout rows_mp[j*p + i], p = col_width*max_nnz
 *                   (src/cuda/openph.cu:41, src/cuda/kernels.cu:25-52).
 *   orc_twist     - src/cuda/algorithms/twist/twist.cu:2-87 (pseudo-dimension
 *                   descending, dims >= 1 only, clearing through low only),
 *                   pseudo-dimension = nnz(col)-1 (src/cuda/dimensions.cu:54-104).
 *   orc_twist_dyn - the same twist on growable columns (no slot limit).
 *   orc_pms       - the pms driver src/cuda/algorithms/pms/pms.cu:1-118 with
 *                   its phases src/cuda/algorithms/pms/pms_functions.cu:7-140,
 *                   src/cuda/algorithms/kernels.cu:79-124 and left/beta
 *                   src/cuda/algorithms/pms/beta.cu:2-15, executed serially
 *                   (Phase II columns in increasing order; pivots published
 *                   by a column are visible to later columns of the same pass,
 *                   which is one of the schedules the reference's racy
 *                   Phase II admits). Columns are growable arrays instead of
 *                   slots, so it has no overflow.
 *   orc_ph_row    - src/cuda/algorithms/ph_row/ph_row.cu:1-121 (row
 *                   algorithm, rows right to left, neighbours of a pivot
 *                   found by walking its pseudo-dimension list).
 *
 * How each is pinned (see tests/test_oracle.py): dense Z/2 brute force on
 * small complexes, union-find elder-rule dimension-0 pairs, closed-form pair
 * counts of full VR 2-skeleta and cubical grids, agreement of the four
 * algorithms with each other, and - in the build container only - the
 * reference's own kernels_host.cu compiled from /root/reference.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_OK 0
#define ORC_ENOMEM -1
#define ORC_EOVERFLOW -2
#define ORC_EINPUT -3

/* stats[0] = column additions, stats[1] = sum over additions of
 * |a| + |b| + |a xor b|, stats[2] = longest column produced,
 * stats[3] = iterations (pms) or processed columns. */
enum { ST_ADD = 0, ST_ELEMS = 1, ST_MAXLEN = 2, ST_ITERS = 3, ST_N = 4 };

static int max_col_nnz(int m, const int64_t *colptr) {
    int mx = 0;
    for (int j = 0; j < m; j++) {
        int c = (int)(colptr[j + 1] - colptr[j]);
        if (c > mx) mx = c;
    }
    return mx;
}

/* ---------------------------------------------------------------------- */
/* Fixed-slot store shared by standard and twist.                          */
/* ---------------------------------------------------------------------- */

typedef struct {
    int m, p;
    int32_t *slots; /* m*p, -1 padded */
    int32_t *len;   /* live entries per column */
    int32_t *tmp;   /* p scratch */
} slot_store;

static int slots_init(slot_store *s, int m, const int64_t *colptr,
                      const int32_t *rows, int col_width) {
    int mx = max_col_nnz(m, colptr);
    if (col_width < 1) col_width = 1;
    s->m = m;
    s->p = col_width * (mx > 0 ? mx : 1);
    s->slots = (int32_t *)malloc(sizeof(int32_t) * (size_t)m * s->p + 4);
    s->len = (int32_t *)calloc((size_t)m + 1, sizeof(int32_t));
    s->tmp = (int32_t *)malloc(sizeof(int32_t) * (size_t)s->p + 4);
    if (!s->slots || !s->len || !s->tmp) return ORC_ENOMEM;
    for (size_t i = 0; i < (size_t)m * s->p; i++) s->slots[i] = -1;
    for (int j = 0; j < m; j++) {
        int c = (int)(colptr[j + 1] - colptr[j]);
        memcpy(s->slots + (size_t)j * s->p, rows + colptr[j], sizeof(int32_t) * c);
        s->len[j] = c;
    }
    return ORC_OK;
}

static void slots_free(slot_store *s) {
    free(s->slots);
    free(s->len);
    free(s->tmp);
}

/* col[dst] <- col[src] xor col[dst]; returns the new low of dst or -1.
 * Follows left_to_right_host (kernels_host.cu:2-52): merge into a scratch
 * region, then copy back over the destination's slots. */
static int slot_add(slot_store *s, int src, int dst, int64_t *st, int *err) {
    const int32_t *a = s->slots + (size_t)src * s->p;
    int32_t *b = s->slots + (size_t)dst * s->p;
    int na = s->len[src], nb = s->len[dst];
    int ia = 0, ib = 0, n = 0;
    int32_t *t = s->tmp;
    while (ia < na && ib < nb) {
        int32_t x = a[ia], y = b[ib];
        if (x < y) { t[n++] = x; ia++; }
        else if (y < x) { t[n++] = y; ib++; }
        else { ia++; ib++; }
        if (n >= s->p && (ia < na || ib < nb)) { *err = ORC_EOVERFLOW; return -1; }
    }
    while (ia < na) { if (n >= s->p) { *err = ORC_EOVERFLOW; return -1; } t[n++] = a[ia++]; }
    while (ib < nb) { if (n >= s->p) { *err = ORC_EOVERFLOW; return -1; } t[n++] = b[ib++]; }
    memcpy(b, t, sizeof(int32_t) * n);
    for (int i = n; i < nb; i++) b[i] = -1;
    s->len[dst] = n;
    if (st) {
        st[ST_ADD] += 1;
        st[ST_ELEMS] += na + nb + n;
        if (n > st[ST_MAXLEN]) st[ST_MAXLEN] = n;
    }
    return n ? t[n - 1] : -1;
}

/* reduce_col_host (kernels_host.cu:62-75) */
static int slot_reduce_col(slot_store *s, int j, int32_t *low, int32_t *arglow,
                           int32_t *ess, int64_t *st) {
    int err = ORC_OK;
    int lj = low[j];
    while (lj > -1 && arglow[lj] != -1) {
        ess[lj] = 0;
        lj = slot_add(s, arglow[lj], j, st, &err);
        if (err) return err;
        low[j] = lj;
    }
    if (lj > -1) arglow[lj] = j;
    return ORC_OK;
}

static void init_low(int m, const int64_t *colptr, const int32_t *rows, int32_t *low) {
    for (int j = 0; j < m; j++)
        low[j] = colptr[j + 1] > colptr[j] ? rows[colptr[j + 1] - 1] : -1;
}

int orc_standard(int m, const int64_t *colptr, const int32_t *rows, int col_width,
                 int32_t *low, int32_t *ess, int64_t *stats) {
    if (m < 0) return ORC_EINPUT;
    slot_store s;
    int rc = slots_init(&s, m, colptr, rows, col_width);
    int32_t *arglow = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m + 1));
    if (rc || !arglow) { slots_free(&s); free(arglow); return ORC_ENOMEM; }
    init_low(m, colptr, rows, low);
    for (int j = 0; j < m; j++) { arglow[j] = -1; ess[j] = 1; }
    int64_t st[ST_N] = {0, 0, 0, 0};
    for (int j = 0; j < m && rc == ORC_OK; j++) {
        rc = slot_reduce_col(&s, j, low, arglow, ess, st);
        if (low[j] > -1) { ess[j] = 0; ess[low[j]] = 0; }
    }
    st[ST_ITERS] = m;
    if (stats) memcpy(stats, st, sizeof(st));
    free(arglow);
    slots_free(&s);
    return rc;
}

/* Pseudo-dimension of every column (nnz - 1; empty columns are -1) and the
 * per-dimension lists in index order (dimensions.cu:54-104). */
static int pseudo_dims(int m, const int64_t *colptr, int32_t *pd, int *cdim_out) {
    int cd = -1;
    for (int j = 0; j < m; j++) {
        pd[j] = (int32_t)(colptr[j + 1] - colptr[j]) - 1;
        if (pd[j] > cd) cd = pd[j];
    }
    *cdim_out = cd;
    return ORC_OK;
}

int orc_twist(int m, const int64_t *colptr, const int32_t *rows, int col_width,
              int32_t *low, int32_t *ess, int64_t *stats) {
    if (m < 0) return ORC_EINPUT;
    slot_store s;
    int rc = slots_init(&s, m, colptr, rows, col_width);
    int32_t *arglow = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m + 1));
    int32_t *pd = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m + 1));
    if (rc || !arglow || !pd) { slots_free(&s); free(arglow); free(pd); return ORC_ENOMEM; }
    int cd;
    pseudo_dims(m, colptr, pd, &cd);
    init_low(m, colptr, rows, low);
    for (int j = 0; j < m; j++) { arglow[j] = -1; ess[j] = 1; }
    int64_t st[ST_N] = {0, 0, 0, 0};
    int64_t processed = 0;
    for (int d = cd; d > 0 && rc == ORC_OK; d--) {
        for (int j = 0; j < m && rc == ORC_OK; j++) {
            if (pd[j] != d) continue;
            rc = slot_reduce_col(&s, j, low, arglow, ess, st);
            int lj = low[j];
            if (lj > -1) low[lj] = -1; /* twist_step_host: clear via low only */
            if (lj > -1) { ess[j] = 0; ess[lj] = 0; }
            processed++;
        }
    }
    st[ST_ITERS] = processed;
    if (stats) memcpy(stats, st, sizeof(st));
    free(arglow);
    free(pd);
    slots_free(&s);
    return rc;
}

/* ---------------------------------------------------------------------- */
/* pms restatement with growable columns.                                  */
/* ---------------------------------------------------------------------- */

typedef struct {
    int32_t **col;
    int32_t *len, *cap;
    int32_t *tmp;
    int tmpcap;
} dyn_store;

static int dyn_init(dyn_store *s, int m, const int64_t *colptr, const int32_t *rows) {
    s->col = (int32_t **)calloc((size_t)m + 1, sizeof(int32_t *));
    s->len = (int32_t *)calloc((size_t)m + 1, sizeof(int32_t));
    s->cap = (int32_t *)calloc((size_t)m + 1, sizeof(int32_t));
    s->tmpcap = 64;
    s->tmp = (int32_t *)malloc(sizeof(int32_t) * s->tmpcap);
    if (!s->col || !s->len || !s->cap || !s->tmp) return ORC_ENOMEM;
    for (int j = 0; j < m; j++) {
        int c = (int)(colptr[j + 1] - colptr[j]);
        s->len[j] = c;
        s->cap[j] = c;
        if (c) {
            s->col[j] = (int32_t *)malloc(sizeof(int32_t) * c);
            if (!s->col[j]) return ORC_ENOMEM;
            memcpy(s->col[j], rows + colptr[j], sizeof(int32_t) * c);
        }
    }
    return ORC_OK;
}

static void dyn_free(dyn_store *s, int m) {
    if (s->col)
        for (int j = 0; j < m; j++) free(s->col[j]);
    free(s->col);
    free(s->len);
    free(s->cap);
    free(s->tmp);
}

/* col[dst] <- col[src] xor col[dst]; returns new low of dst. */
static int dyn_add(dyn_store *s, int src, int dst, int64_t *st, int *err) {
    int na = s->len[src], nb = s->len[dst];
    if (na + nb > s->tmpcap) {
        int nc = 2 * (na + nb);
        int32_t *t = (int32_t *)realloc(s->tmp, sizeof(int32_t) * nc);
        if (!t) { *err = ORC_ENOMEM; return -1; }
        s->tmp = t;
        s->tmpcap = nc;
    }
    const int32_t *a = s->col[src], *b = s->col[dst];
    int32_t *t = s->tmp;
    int ia = 0, ib = 0, n = 0;
    while (ia < na && ib < nb) {
        int32_t x = a[ia], y = b[ib];
        if (x < y) { t[n++] = x; ia++; }
        else if (y < x) { t[n++] = y; ib++; }
        else { ia++; ib++; }
    }
    while (ia < na) t[n++] = a[ia++];
    while (ib < nb) t[n++] = b[ib++];
    if (n > s->cap[dst]) {
        int32_t *c = (int32_t *)realloc(s->col[dst], sizeof(int32_t) * n);
        if (!c) { *err = ORC_ENOMEM; return -1; }
        s->col[dst] = c;
        s->cap[dst] = n;
    }
    memcpy(s->col[dst], t, sizeof(int32_t) * n);
    s->len[dst] = n;
    if (st) {
        st[ST_ADD] += 1;
        st[ST_ELEMS] += na + nb + n;
        if (n > st[ST_MAXLEN]) st[ST_MAXLEN] = n;
    }
    return n ? t[n - 1] : -1;
}

/* left[i] = leftmost column containing row i; beta[j] = max{i : left[i]=j}
 * (beta.cu:2-15). */
static void left_beta(int m, const int64_t *colptr, const int32_t *rows,
                      int32_t *left, int32_t *beta) {
    for (int i = 0; i < m; i++) { left[i] = -1; beta[i] = -1; }
    for (int j = 0; j < m; j++)
        for (int64_t e = colptr[j]; e < colptr[j + 1]; e++) {
            int r = rows[e];
            if (left[r] == -1) {
                left[r] = j;
                if (r > beta[j]) beta[j] = r;
            }
        }
}

/* iter_stats (optional): per iteration 5 int64 -
 *   additions, merge elements, columns marked in Phase I, columns that
 *   entered Phase II's loop, longest chain of additions of one column. */
/* The quantities track() reads at the end of an iteration
 * (tracker.cu:92-134), as exact integers: q[0] = sum_j |low_j - lt_j|,
 * q[1] = max_j |low_j - lt_j| (both 0-based, openph.cu:113), q[2] =
 * #{classes_j == 0}, q[3] = #{ess_j == 1}. */
static void orc_track(int m, const int32_t *low, const int32_t *lt, const int32_t *classes,
                      const int32_t *ess, int64_t *q) {
    int64_t l1 = 0, linf = 0, un = 0, es = 0;
    for (int j = 0; j < m; j++) {
        int64_t d = (int64_t)low[j] - (int64_t)lt[j];
        if (d < 0) d = -d;
        l1 += d;
        if (d > linf) linf = d;
        un += classes[j] == 0;
        es += ess[j] == 1;
    }
    q[0] = l1;
    q[1] = linf;
    q[2] = un;
    q[3] = es;
}

/* low_true0 (0-based, may be NULL) and trk (4 x (m+1) entries, may be NULL):
 * when both are given, trk[4*i .. 4*i+3] receives orc_track of the state
 * tracked at iteration i (pms.cu:16-19 before Phase 0, pms.cu:106-109 after
 * every iteration). */
int orc_pms(int m, const int64_t *colptr, const int32_t *rows, int32_t *low,
            int32_t *ess, int32_t *num_iters, int64_t *stats, int64_t *iter_stats,
            int max_iter_stats, int max_iters, int snapshot, const int32_t *low_true0,
            int64_t *trk) {
    if (m < 0) return ORC_EINPUT;
    int rc = ORC_OK;
    dyn_store s;
    size_t mm = (size_t)m + 1;
    int32_t *arglow = malloc(sizeof(int32_t) * mm), *classes = calloc(mm, sizeof(int32_t));
    int32_t *left = malloc(sizeof(int32_t) * mm), *beta = malloc(sizeof(int32_t) * mm);
    int32_t *pd = malloc(sizeof(int32_t) * mm), *visited = malloc(sizeof(int32_t) * mm);
    int32_t *clear = malloc(sizeof(int32_t) * mm);
    int32_t *newpiv = malloc(sizeof(int32_t) * mm);
    if (dyn_init(&s, m, colptr, rows) || !arglow || !classes || !left || !beta ||
        !pd || !visited || !clear || !newpiv) {
        rc = ORC_ENOMEM;
        goto done;
    }
    int cd;
    pseudo_dims(m, colptr, pd, &cd);
    int ndim = cd + 2; /* pseudo-dims -1 .. cd */
    int32_t *ceil_d = malloc(sizeof(int32_t) * (ndim > 0 ? ndim : 1));
    if (!ceil_d) { rc = ORC_ENOMEM; goto done; }
    init_low(m, colptr, rows, low);
    left_beta(m, colptr, rows, left, beta);
    for (int j = 0; j < m; j++) { arglow[j] = -1; classes[j] = 0; ess[j] = 1; }
    int64_t st[ST_N] = {0, 0, 0, 0};
    const int tracking = low_true0 && trk;
    if (tracking) orc_track(m, low, low_true0, classes, ess, trk);

    /* Phase 0: mark_pivots_and_clear (pms_functions.cu:33-53) */
    for (int j = 0; j < m; j++) {
        int lj = low[j];
        if (lj > -1 && lj == beta[j]) {
            classes[j] = -1;
            low[lj] = -1;
            classes[lj] = 1;
            arglow[lj] = j;
            ess[lj] = 0;
            ess[j] = 0;
        }
    }

    int iter = 1;
    int converged = 0;
    while (!converged) {
        if (max_iters > 0 && iter > max_iters) { rc = ORC_EOVERFLOW; break; }
        int64_t add0 = st[ST_ADD], el0 = st[ST_ELEMS], marked = 0, entered = 0, maxchain = 0;
        /* Phase I: transverse_dimensions (pms_functions.cu:55-85) */
        for (int j = 0; j < m; j++) { visited[j] = 0; clear[j] = 0; }
        for (int d = 0; d < ndim; d++) ceil_d[d] = -1;
        for (int j = 0; j < m; j++) { /* per-dimension walks interleaved; rows of
                                        different pseudo-dims are disjoint */
            int lj = low[j];
            if (lj < 0) continue;
            int dp = pd[j] + 1;
            if (!visited[lj]) {
                if (classes[j] == 0 && lj > ceil_d[dp]) {
                    arglow[lj] = j;
                    classes[j] = -1;
                    clear[lj] = 1;
                    ess[lj] = 0;
                    ess[j] = 0;
                    marked++;
                }
            } else if (lj > ceil_d[dp]) {
                ceil_d[dp] = lj;
            }
            visited[lj] = 1;
        }
        /* clear_positives (pms_functions.cu:131-140) */
        for (int t = 0; t < m; t++)
            if (clear[t]) { low[t] = -1; classes[t] = 1; }
        /* Phase II (pms_functions.cu:87-112). With snapshot != 0, pivots
         * found in this pass are published only after it (newpiv), as in a
         * schedule where every column reads the arglow of the pass start. */
        for (int t = 0; t < m; t++) { clear[t] = 0; newpiv[t] = -1; }
        for (int j = 0; j < m && rc == ORC_OK; j++) {
            int lj = low[j];
            int pivot = lj > -1 ? arglow[lj] : -1;
            if (-1 < pivot && pivot < j && classes[j] == 0) entered++;
            int64_t chain0 = st[ST_ADD];
            while (-1 < pivot && pivot < j && classes[j] == 0) {
                if (low[j] > -1) ess[low[j]] = 0;
                lj = dyn_add(&s, pivot, j, st, &rc);
                if (rc) break;
                low[j] = lj;
                if (lj > -1) {
                    if (left[lj] == j) {
                        if (snapshot) newpiv[lj] = j; else arglow[lj] = j;
                        classes[j] = -1;
                    }
                    clear[lj] = 1;
                    pivot = arglow[lj];
                } else {
                    pivot = -1;
                }
            }
            if (st[ST_ADD] - chain0 > maxchain) maxchain = st[ST_ADD] - chain0;
        }
        if (rc) break;
        for (int t = 0; t < m; t++) {
            if (newpiv[t] > -1) arglow[t] = newpiv[t];
            if (clear[t]) { low[t] = -1; classes[t] = 1; }
        }
        /* ess_hat (kernels.cu:113-124) */
        for (int j = 0; j < m; j++) {
            if (low[j] > -1) ess[low[j]] = 0;
            if (arglow[j] > -1) { ess[arglow[j]] = 0; ess[j] = 0; }
        }
        /* update_classes (kernels.cu:93-101) */
        for (int t = 0; t < m; t++)
            if (arglow[t] > -1) { classes[arglow[t]] = -1; classes[t] = 1; }
        /* is_reduced (pms_functions.cu:7-27) */
        for (int t = 0; t < m; t++) visited[t] = 0;
        converged = 1;
        for (int j = 0; j < m; j++) {
            int lj = low[j];
            if (lj > -1) {
                if (visited[lj]) converged = 0;
                visited[lj] = 1;
            }
        }
        if (converged) {
            /* ess_hat_final, update_classes_final (kernels.cu:79-111) */
            for (int j = 0; j < m; j++)
                if (low[j] > -1) { ess[j] = 0; ess[low[j]] = 0; }
            for (int t = 0; t < m; t++) {
                if (low[t] > -1) { classes[t] = -1; classes[low[t]] = 1; }
                else if (ess[t] == 1) classes[t] = 2;
            }
        }
        if (tracking && iter <= m) orc_track(m, low, low_true0, classes, ess, trk + 4 * (int64_t)iter);
        if (iter_stats && iter - 1 < max_iter_stats) {
            int64_t *q = iter_stats + 5 * (int64_t)(iter - 1);
            q[0] = st[ST_ADD] - add0;
            q[1] = st[ST_ELEMS] - el0;
            q[2] = marked;
            q[3] = entered;
            q[4] = maxchain;
        }
        iter++;
    }
    st[ST_ITERS] = iter - 1;
    if (num_iters) *num_iters = iter;
    if (stats) memcpy(stats, st, sizeof(st));
    free(ceil_d);
done:
    free(arglow); free(classes); free(left); free(beta); free(pd); free(visited); free(clear); free(newpiv);
    dyn_free(&s, m);
    return rc;
}

/* twist (twist.cu:2-87) on growable columns: the same order, clearing and
 * merge as orc_twist, with no slot limit, for matrices whose fixed-slot
 * layout m*p would not fit in host memory (cubical 64^3: p ~ 1,734). */
int orc_twist_dyn(int m, const int64_t *colptr, const int32_t *rows, int32_t *low,
                  int32_t *ess, int64_t *stats) {
    if (m < 0) return ORC_EINPUT;
    int rc = ORC_OK;
    dyn_store s;
    size_t mm = (size_t)m + 1;
    int32_t *arglow = malloc(sizeof(int32_t) * mm), *pd = malloc(sizeof(int32_t) * mm);
    if (dyn_init(&s, m, colptr, rows) || !arglow || !pd) { rc = ORC_ENOMEM; goto done; }
    int cd;
    pseudo_dims(m, colptr, pd, &cd);
    init_low(m, colptr, rows, low);
    for (int j = 0; j < m; j++) { arglow[j] = -1; ess[j] = 1; }
    int64_t st[ST_N] = {0, 0, 0, 0};
    int64_t processed = 0;
    for (int d = cd; d > 0 && rc == ORC_OK; d--) {
        for (int j = 0; j < m && rc == ORC_OK; j++) {
            if (pd[j] != d) continue;
            /* reduce_col_host (kernels_host.cu:62-75) */
            int lj = low[j];
            while (lj > -1 && arglow[lj] != -1 && rc == ORC_OK) {
                ess[lj] = 0;
                lj = dyn_add(&s, arglow[lj], j, st, &rc);
                low[j] = lj;
            }
            if (rc) break;
            if (lj > -1) {
                arglow[lj] = j;
                low[lj] = -1; /* twist_step_host: clear via low only */
                ess[j] = 0;
                ess[lj] = 0;
            }
            processed++;
        }
    }
    st[ST_ITERS] = processed;
    if (stats) memcpy(stats, st, sizeof(st));
done:
    free(arglow); free(pd);
    dyn_free(&s, m);
    return rc;
}

/* ---------------------------------------------------------------------- */
/* Row algorithm (ph_row.cu:1-121).                                         */
/* ---------------------------------------------------------------------- */

/* all_dims = 0: ph_row (ph_row.cu:38-121), neighbours from the pivot's
 * pseudo-dimension list. all_dims = 1: ph_row_parallel
 * (ph_row_parallel.cu:20-84): every later column with the pivot's low
 * (mark_neighbours, ph_row_parallel.cu:2-10), and ess from ess_hat
 * (kernels.cu:113-124) after every row; arglow is never set by this
 * algorithm, so only the lows are marked, and there is no final pass. */
int orc_ph_row(int m, const int64_t *colptr, const int32_t *rows, int32_t *low,
               int32_t *ess, int64_t *stats, int all_dims) {
    if (m < 0) return ORC_EINPUT;
    int rc = ORC_OK;
    dyn_store s;
    size_t mm = (size_t)m + 1;
    int32_t *pivots = malloc(sizeof(int32_t) * mm), *pd = malloc(sizeof(int32_t) * mm);
    int32_t *nb = malloc(sizeof(int32_t) * mm), *next = malloc(sizeof(int32_t) * mm);
    int32_t *lastd = NULL;
    if (dyn_init(&s, m, colptr, rows) || !pivots || !pd || !nb || !next) { rc = ORC_ENOMEM; goto done; }
    int cd;
    pseudo_dims(m, colptr, pd, &cd);
    int ndim = cd + 2;
    lastd = malloc(sizeof(int32_t) * (ndim > 0 ? ndim : 1));
    if (!lastd) { rc = ORC_ENOMEM; goto done; }
    for (int d = 0; d < ndim; d++) lastd[d] = -1;
    for (int j = 0; j < m; j++) next[j] = -1;
    for (int j = 0; j < m; j++) {
        int dp = pd[j] + 1;
        if (lastd[dp] > -1) next[lastd[dp]] = j;
        lastd[dp] = j;
    }
    init_low(m, colptr, rows, low);
    for (int j = 0; j < m; j++) { ess[j] = 1; pivots[j] = -1; }
    for (int j = 0; j < m; j++)
        if (low[j] != -1 && (pivots[low[j]] == -1 || j < pivots[low[j]])) pivots[low[j]] = j;
    int64_t st[ST_N] = {0, 0, 0, 0};
    for (int i = m - 1; i > -1 && rc == ORC_OK; i--) {
        int pv = pivots[i];
        if (pv == -1) {
            if (all_dims) /* ess_hat runs after every row */
                for (int j = 0; j < m; j++)
                    if (low[j] > -1) ess[low[j]] = 0;
            continue;
        }
        /* neighbours: later columns of the pivot's pseudo-dimension whose
         * low equals low[pv] (get_neighbours_position_host, ph_row.cu:27-37) */
        int cnt = 0;
        if (all_dims) {
            for (int c = pv + 1; c < m; c++)
                if (low[c] == i) nb[cnt++] = c;
        } else {
            for (int c = next[pv]; c != -1; c = next[c])
                if (low[c] == low[pv]) nb[cnt++] = c;
        }
        for (int l = 0; l < cnt && rc == ORC_OK; l++) {
            int j = nb[l];
            if (!all_dims) {
                if (low[j] > -1) ess[low[j]] = 0;
                ess[pv] = 0;
                ess[low[pv]] = 0;
            }
            int lj = dyn_add(&s, pv, j, st, &rc);
            low[j] = lj;
            if (lj != -1 && (pivots[lj] == -1 || j < pivots[lj])) pivots[lj] = j;
            if (!all_dims && lj != -1 && pivots[lj] == j) { ess[j] = 0; ess[lj] = 0; }
        }
        if (all_dims)
            for (int j = 0; j < m; j++)
                if (low[j] > -1) ess[low[j]] = 0;
    }
    if (!all_dims)
        for (int j = 0; j < m; j++)
            if (low[j] > -1) { ess[j] = 0; ess[low[j]] = 0; }
    st[ST_ITERS] = m;
    if (stats) memcpy(stats, st, sizeof(st));
done:
    free(pivots); free(pd); free(nb); free(next); free(lastd);
    dyn_free(&s, m);
    return rc;
}
