// Test-only driver around the reference's own serial column reduction.
//
// Compiled by oracle/Makefile.ref with -I pointing at the reference's
// src/cuda directory; the reference sources are included where they lie and
// are not copied. The two included files need nothing beyond the C standard
// library. The outer loops below are this driver's own: the reference's
// standard() / twist() (src/cuda/algorithms/standard/standard.cu:1-57,
// src/cuda/algorithms/twist/twist.cu:10-87) also time each column with CUDA
// events and call the O(m) tracker, so they are not built here; the loops
// keep their order (all columns left to right; pseudo-dimensions from the
// top down to 1 with clearing through low) and their ess updates.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "functions_host.cu"
#include "algorithms/kernels_host.cu"

namespace {

struct Slots {
    int p = 0;
    std::vector<int> rows_mp, aux_mp;
};

// create_rows_mp + compute_low (src/cuda/kernels.cu:25-52) on the host.
void make_slots(Slots &s, int m, const int64_t *colptr, const int32_t *rows,
                int col_width, int *low) {
    int mx = 1;
    for (int j = 0; j < m; j++) {
        int c = (int)(colptr[j + 1] - colptr[j]);
        if (c > mx) mx = c;
    }
    s.p = col_width * mx;
    // one extra column of padding: left_to_right_host reads the slot one past
    // a full column before checking its bound (kernels_host.cu:10-11,34-35)
    size_t n = (size_t)(m + 1) * s.p;
    s.rows_mp.assign(n, -1);
    s.aux_mp.assign(n, -1);
    for (int j = 0; j < m; j++) {
        low[j] = -1;
        for (int64_t e = colptr[j]; e < colptr[j + 1]; e++) {
            s.rows_mp[(size_t)j * s.p + (e - colptr[j])] = rows[e];
            low[j] = rows[e];
        }
    }
}

}  // namespace

extern "C" int ref_standard(int m, const int64_t *colptr, const int32_t *rows,
                            int col_width, int32_t *low, int32_t *ess) {
    Slots s;
    make_slots(s, m, colptr, rows, col_width, low);
    std::vector<int> arglow(m + 1);
    fill_host(arglow.data(), -1, m);
    fill_host(ess, 1, m);
    for (int j = 0; j < m; j++) {
        reduce_col_host(j, s.rows_mp.data(), s.aux_mp.data(), low, arglow.data(), m, s.p, ess);
        if (low[j] > -1) {
            ess[j] = 0;
            ess[low[j]] = 0;
        }
    }
    return 0;
}

extern "C" int ref_twist(int m, const int64_t *colptr, const int32_t *rows,
                         int col_width, int32_t *low, int32_t *ess) {
    Slots s;
    make_slots(s, m, colptr, rows, col_width, low);
    std::vector<int> arglow(m + 1), pd(m + 1);
    fill_host(arglow.data(), -1, m);
    fill_host(ess, 1, m);
    int cd = -1;
    for (int j = 0; j < m; j++) {
        pd[j] = (int)(colptr[j + 1] - colptr[j]) - 1;
        if (pd[j] > cd) cd = pd[j];
    }
    for (int d = cd; d > 0; d--) {
        for (int j = 0; j < m; j++) {
            if (pd[j] != d) continue;
            reduce_col_host(j, s.rows_mp.data(), s.aux_mp.data(), low, arglow.data(), m, s.p, ess);
            int lj = low[j];
            if (lj > -1) {
                low[lj] = -1;
                ess[j] = 0;
                ess[lj] = 0;
            }
        }
    }
    return 0;
}
