// Filtrations built on the device, as 1-based COO boundary matrices in
// filtration order (the input the reduction takes):
//   * the Vietoris-Rips 2-skeleton of a point cloud at full radius;
//   * the lower-star filtration of the cubical complex of a 3-D grid.
// The reference builds its matrices in Javaplex on the host
// (numerics/utils/stream2cmo.m:1-26, numerics/utils/example_factory.m:17-122,
// numerics/utils/pointcloud_to_vr_stream.m). The order here: by filtration
// value, then dimension, then the lexicographic vertex tuple (VR) or the
// cell's linear index (cubical). The values are computed with the same IEEE
// operations, in the same order, as openph_amd/complexes.py, so both give
// identical matrices.
#include <hipcub/hipcub.hpp>

#include <cstring>

#include <algorithm>

#include "host.h"

namespace oph {

namespace {

__host__ __device__ inline long long choose2(long long n) { return n < 2 ? 0 : n * (n - 1) / 2; }
__host__ __device__ inline long long choose3(long long n) {
    return n < 3 ? 0 : n * (n - 1) / 2 * (n - 2) / 3;
}

// ---------------------------------------------------------------------------
// Vietoris-Rips 2-skeleton. Cell ids in their natural (secondary) order:
// vertices [0, n), edges [n, n + E) in (i, j) lexicographic order, triangles
// [n + E, m) in (a, b, c) lexicographic order. A stable sort by value then
// gives the filtration order (value, dimension, vertex tuple).
// ---------------------------------------------------------------------------

// Euclidean distance of points i and j: differences squared and summed from
// the first coordinate on, each operation rounded once (no fused
// multiply-add), then a correctly rounded square root.
__global__ void k_vr_dist(const double *__restrict__ p, int n, int d, double *__restrict__ D) {
#pragma clang fp contract(off)
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * n) return;
    const int i = (int)(t / n), j = (int)(t % n);
    double s = 0.0;
    for (int k = 0; k < d; k++) {
        const double a = p[(long long)i * d + k] - p[(long long)j * d + k];
        const double q = a * a;
        s = k == 0 ? q : s + q;
    }
    D[t] = __builtin_sqrt(s);
}

// first index of the pair block holding lexicographic pair number e among
// n items: the largest i with start(i) = i*n - i*(i+1)/2 <= e
__device__ inline int pair_row(long long e, int n) {
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        const long long st = (long long)mid * n - (long long)mid * (mid + 1) / 2;
        if (st <= e) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ inline void edge_of(long long e, int n, int &i, int &j) {
    i = pair_row(e, n);
    const long long st = (long long)i * n - (long long)i * (i + 1) / 2;
    j = (int)(e - st) + i + 1;
}

// triangles with first vertex < a come first: C(n,3) - C(n-a,3) of them
__device__ inline void tri_of(long long t, int n, int &a, int &b, int &c) {
    const long long all = choose3(n);
    int lo = 0, hi = n - 2;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (all - choose3(n - mid) <= t) lo = mid;
        else hi = mid;
    }
    a = lo;
    const long long local = t - (all - choose3(n - a));
    int jj, kk;
    edge_of(local, n - a - 1, jj, kk);
    b = jj + a + 1;
    c = kk + a + 1;
}

__global__ void k_vr_keys(const double *__restrict__ D, int n, long long E, long long m,
                          unsigned long long *__restrict__ key, unsigned *__restrict__ val) {
    for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < m;
         id += (long long)gridDim.x * blockDim.x) {
        double v = 0.0;
        if (id >= n && id < n + E) {
            int i, j;
            edge_of(id - n, n, i, j);
            v = D[(long long)i * n + j];
        } else if (id >= n + E) {
            int a, b, c;
            tri_of(id - n - E, n, a, b, c);
            const double ab = D[(long long)a * n + b], ac = D[(long long)a * n + c],
                         bc = D[(long long)b * n + c];
            v = fmax(fmax(ab, ac), bc);
        }
        // values are >= +0, so their bit patterns sort as the values do
        key[id] = (unsigned long long)__double_as_longlong(v);
        val[id] = (unsigned)id;
    }
}

__global__ void k_scatter_pos(const unsigned *__restrict__ sorted_id, long long m,
                              int *__restrict__ pos) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < m;
         k += (long long)gridDim.x * blockDim.x)
        pos[sorted_id[k]] = (int)k;
}

__global__ void k_vr_nnz(const unsigned *__restrict__ sorted_id, int n, long long E, long long m,
                         int *__restrict__ nnz) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < m;
         k += (long long)gridDim.x * blockDim.x) {
        const long long id = sorted_id[k];
        nnz[k] = id < n ? 0 : (id < n + E ? 2 : 3);
    }
}

__device__ inline long long edge_index(int i, int j, int n) {
    return (long long)i * n - (long long)i * (i + 1) / 2 + (j - i - 1);
}

__global__ void k_vr_coo(const unsigned *__restrict__ sorted_id,
                         const unsigned long long *__restrict__ sorted_key,
                         const int *__restrict__ pos, const int *__restrict__ off, int n,
                         long long E, long long m, int *__restrict__ rows1,
                         int *__restrict__ cols1, double *__restrict__ values,
                         int *__restrict__ dims) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < m;
         k += (long long)gridDim.x * blockDim.x) {
        const long long id = sorted_id[k];
        const int o = off[k];
        int dim = 0;
        if (id >= n && id < n + E) {
            int i, j;
            edge_of(id - n, n, i, j);
            int r0 = pos[i], r1 = pos[j];
            if (r0 > r1) {
                const int x = r0;
                r0 = r1;
                r1 = x;
            }
            rows1[o] = r0 + 1;
            rows1[o + 1] = r1 + 1;
            cols1[o] = cols1[o + 1] = (int)k + 1;
            dim = 1;
        } else if (id >= n + E) {
            int a, b, c;
            tri_of(id - n - E, n, a, b, c);
            int r[3] = {pos[n + edge_index(b, c, n)], pos[n + edge_index(a, c, n)],
                        pos[n + edge_index(a, b, n)]};
            if (r[0] > r[1]) { const int x = r[0]; r[0] = r[1]; r[1] = x; }
            if (r[1] > r[2]) { const int x = r[1]; r[1] = r[2]; r[2] = x; }
            if (r[0] > r[1]) { const int x = r[0]; r[0] = r[1]; r[1] = x; }
            for (int q = 0; q < 3; q++) {
                rows1[o + q] = r[q] + 1;
                cols1[o + q] = (int)k + 1;
            }
            dim = 2;
        }
        if (values) values[k] = __longlong_as_double((long long)sorted_key[k]);
        if (dims) dims[k] = dim;
    }
}

// ---------------------------------------------------------------------------
// Cubical lower-star filtration of an s0 x s1 x s2 grid. Cell (x, y, z) in
// [0, 2s-1)^3 (even coordinate = vertex position, odd = an edge along that
// axis), linear index (x * N1 + y) * N2 + z, dimension = number of odd
// coordinates, value = max of the field over the cell's vertices. One 64-bit
// key holds (value, dimension, index), so a sort of the keys is the order.
// ---------------------------------------------------------------------------

constexpr int kCubIndexBits = 28;

__device__ inline unsigned float_order(float v) {
    unsigned u = __float_as_uint(v == 0.0f ? 0.0f : v);  // -0 sorts with +0
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ void k_cub_keys(const float *__restrict__ f, int s0, int s1, int s2, long long m,
                           unsigned long long *__restrict__ key) {
    const int N1 = 2 * s1 - 1, N2 = 2 * s2 - 1;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < m;
         c += (long long)gridDim.x * blockDim.x) {
        const int z = (int)(c % N2), y = (int)((c / N2) % N1), x = (int)(c / ((long long)N1 * N2));
        const int x0 = x >> 1, x1 = (x + 1) >> 1, y0 = y >> 1, y1 = (y + 1) >> 1, z0 = z >> 1,
                  z1 = (z + 1) >> 1;
        float v = f[((long long)x0 * s1 + y0) * s2 + z0];
        for (int a = x0; a <= x1; a++)
            for (int b = y0; b <= y1; b++)
                for (int e = z0; e <= z1; e++) v = fmaxf(v, f[((long long)a * s1 + b) * s2 + e]);
        const unsigned dim = (unsigned)((x & 1) + (y & 1) + (z & 1));
        key[c] = ((unsigned long long)float_order(v) << 32) |
                 ((unsigned long long)dim << kCubIndexBits) | (unsigned long long)c;
    }
}

__global__ void k_cub_pos(const unsigned long long *__restrict__ sorted, long long m,
                          int *__restrict__ pos, int *__restrict__ nnz) {
    const unsigned long long mask = (1ull << kCubIndexBits) - 1;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < m;
         k += (long long)gridDim.x * blockDim.x) {
        const unsigned long long key = sorted[k];
        pos[key & mask] = (int)k;
        nnz[k] = 2 * (int)((key >> kCubIndexBits) & 3u);
    }
}

__global__ void k_cub_coo(const unsigned long long *__restrict__ sorted, const float *__restrict__ f,
                          int s0, int s1, int s2, long long m, const int *__restrict__ pos,
                          const int *__restrict__ off, int *__restrict__ rows1,
                          int *__restrict__ cols1, float *__restrict__ values,
                          int *__restrict__ dims) {
    const int N1 = 2 * s1 - 1, N2 = 2 * s2 - 1;
    const long long stride[3] = {(long long)N1 * N2, N2, 1};
    const unsigned long long mask = (1ull << kCubIndexBits) - 1;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < m;
         k += (long long)gridDim.x * blockDim.x) {
        const unsigned long long key = sorted[k];
        const long long c = (long long)(key & mask);
        const int co[3] = {(int)(c / stride[0]), (int)((c / N2) % N1), (int)(c % N2)};
        int r[6];
        int nr = 0;
        for (int ax = 0; ax < 3; ax++)
            if (co[ax] & 1) {
                r[nr++] = pos[c - stride[ax]];
                r[nr++] = pos[c + stride[ax]];
            }
        for (int a = 1; a < nr; a++) {  // insertion sort of at most 6 rows
            const int x = r[a];
            int b = a - 1;
            while (b >= 0 && r[b] > x) {
                r[b + 1] = r[b];
                b--;
            }
            r[b + 1] = x;
        }
        const int o = off[k];
        for (int q = 0; q < nr; q++) {
            rows1[o + q] = r[q] + 1;
            cols1[o + q] = (int)k + 1;
        }
        if (values) {
            // the value again from the field (the key holds it reordered)
            const int x = co[0], y = co[1], z = co[2];
            const int x0 = x >> 1, x1 = (x + 1) >> 1, y0 = y >> 1, y1 = (y + 1) >> 1,
                      z0 = z >> 1, z1 = (z + 1) >> 1;
            float v = f[((long long)x0 * s1 + y0) * s2 + z0];
            for (int a = x0; a <= x1; a++)
                for (int b = y0; b <= y1; b++)
                    for (int e = z0; e <= z1; e++) v = fmaxf(v, f[((long long)a * s1 + b) * s2 + e]);
            values[k] = v;
        }
        if (dims) dims[k] = nr / 2;
    }
}

int grid_for(long long m) {
    const long long g = (m + kBlock - 1) / kBlock;
    return (int)(g < 65536 ? (g < 1 ? 1 : g) : 65536);
}

// exclusive prefix sum of nnz[0..m) into off[0..m) (int32; the caller has
// checked that the total fits)
int scan_offsets(Context &c, const int *nnz, int *off, long long m, hipStream_t s) {
    size_t bytes = 0;
    OPH_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, nnz, off, (int)m, s));
    OPH_CHECK(grow(&c.f_tmp, c.c_ftmp, bytes + 1));
    OPH_CHECK(hipcub::DeviceScan::ExclusiveSum(c.f_tmp, bytes, nnz, off, (int)m, s));
    return OPENPH_OK;
}

}  // namespace

// Grows the builders' buffers for matrices of up to m columns (and VR
// complexes of up to n points), so that serving them does not allocate:
// hipFree waits for the whole device.
int reserve_builders(Context &c, size_t m, int n) {
    if (m == 0 || m > 0x7fffffffull) return OPENPH_OK;
    OPH_CHECK(grow(&c.f_dist, c.c_fdist, (size_t)n * n + 1));
    OPH_CHECK(grow(&c.f_key, c.c_fkey, m));
    OPH_CHECK(grow(&c.f_key2, c.c_fkey2, m));
    OPH_CHECK(grow(&c.f_val, c.c_fval, m));
    OPH_CHECK(grow(&c.f_val2, c.c_fval2, m));
    OPH_CHECK(grow(&c.f_pos, c.c_fpos, m));
    OPH_CHECK(grow(&c.f_nnz, c.c_fnnz, m));
    OPH_CHECK(grow(&c.f_off, c.c_foff, m));
    size_t b1 = 0, b2 = 0, b3 = 0;
    OPH_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, b1, c.f_key, c.f_key2, c.f_val, c.f_val2,
                                                 (int)m, 0, 64, (hipStream_t)0));
    OPH_CHECK(hipcub::DeviceRadixSort::SortKeys(nullptr, b2, c.f_key, c.f_key2, (int)m, 0, 64,
                                                (hipStream_t)0));
    OPH_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, b3, c.f_nnz, c.f_off, (int)m,
                                               (hipStream_t)0));
    OPH_CHECK(grow(&c.f_tmp, c.c_ftmp, std::max(b1, std::max(b2, b3)) + 1));
    return OPENPH_OK;
}

// The two builders on a given workspace and stream, without waiting for
// them (the server's workers build requests inside their batches).
int build_vr(Context &c, const double *points, int n, int dim, int *rows1, int *cols1,
             double *values, int *dims, hipStream_t s) {
    const long long E = choose2(n), T = choose3(n), m = n + E + T, nnz = 2 * E + 3 * T;
    if (n < 0 || dim < 1 || dim > 7 || (n > 0 && !points)) return OPENPH_EINVAL;
    if (m > 0x7fffffffLL || nnz > 0x7fffffffLL) return OPENPH_EINVAL;
    if (nnz > 0 && (!rows1 || !cols1)) return OPENPH_EINVAL;
    if (m == 0) return OPENPH_OK;
    OPH_CHECK(grow(&c.f_dist, c.c_fdist, (size_t)n * n));
    OPH_CHECK(grow(&c.f_key, c.c_fkey, (size_t)m));
    OPH_CHECK(grow(&c.f_key2, c.c_fkey2, (size_t)m));
    OPH_CHECK(grow(&c.f_val, c.c_fval, (size_t)m));
    OPH_CHECK(grow(&c.f_val2, c.c_fval2, (size_t)m));
    OPH_CHECK(grow(&c.f_pos, c.c_fpos, (size_t)m));
    OPH_CHECK(grow(&c.f_nnz, c.c_fnnz, (size_t)m));
    OPH_CHECK(grow(&c.f_off, c.c_foff, (size_t)m));
    k_vr_dist<<<grid_for((long long)n * n), kBlock, 0, s>>>(points, n, dim, c.f_dist);
    k_vr_keys<<<grid_for(m), kBlock, 0, s>>>(c.f_dist, n, E, m, c.f_key, c.f_val);
    size_t bytes = 0;
    OPH_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, c.f_key, c.f_key2, c.f_val,
                                                 c.f_val2, (int)m, 0, 64, s));
    OPH_CHECK(grow(&c.f_tmp, c.c_ftmp, bytes + 1));
    OPH_CHECK(hipcub::DeviceRadixSort::SortPairs(c.f_tmp, bytes, c.f_key, c.f_key2, c.f_val,
                                                 c.f_val2, (int)m, 0, 64, s));
    k_scatter_pos<<<grid_for(m), kBlock, 0, s>>>(c.f_val2, m, c.f_pos);
    k_vr_nnz<<<grid_for(m), kBlock, 0, s>>>(c.f_val2, n, E, m, c.f_nnz);
    int rc = scan_offsets(c, c.f_nnz, c.f_off, m, s);
    if (rc) return rc;
    k_vr_coo<<<grid_for(m), kBlock, 0, s>>>(c.f_val2, c.f_key2, c.f_pos, c.f_off, n, E, m, rows1,
                                             cols1, values, dims);
    OPH_CHECK(hipGetLastError());
    return OPENPH_OK;
}

int build_cubical(Context &c, const float *field, int s0, int s1, int s2, int *rows1, int *cols1,
                  float *values, int *dims, hipStream_t s) {
    int64_t m64 = 0, nnz64 = 0;
    if (s0 < 1 || s1 < 1 || s2 < 1 || !field) return OPENPH_EINVAL;
    const long long N0 = 2LL * s0 - 1, N1 = 2LL * s1 - 1, N2 = 2LL * s2 - 1;
    m64 = N0 * N1 * N2;
    nnz64 = 2 * ((s0 - 1) * N1 * N2 + N0 * (s1 - 1) * N2 + N0 * N1 * (s2 - 1));
    if (m64 >= (1LL << kCubIndexBits) || nnz64 > 0x7fffffffLL) return OPENPH_EINVAL;
    if (nnz64 > 0 && (!rows1 || !cols1)) return OPENPH_EINVAL;
    const long long m = m64;
    OPH_CHECK(grow(&c.f_key, c.c_fkey, (size_t)m));
    OPH_CHECK(grow(&c.f_key2, c.c_fkey2, (size_t)m));
    OPH_CHECK(grow(&c.f_pos, c.c_fpos, (size_t)m));
    OPH_CHECK(grow(&c.f_nnz, c.c_fnnz, (size_t)m));
    OPH_CHECK(grow(&c.f_off, c.c_foff, (size_t)m));
    k_cub_keys<<<grid_for(m), kBlock, 0, s>>>(field, s0, s1, s2, m, c.f_key);
    size_t bytes = 0;
    OPH_CHECK(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, c.f_key, c.f_key2, (int)m, 0, 64, s));
    OPH_CHECK(grow(&c.f_tmp, c.c_ftmp, bytes + 1));
    OPH_CHECK(hipcub::DeviceRadixSort::SortKeys(c.f_tmp, bytes, c.f_key, c.f_key2, (int)m, 0, 64, s));
    k_cub_pos<<<grid_for(m), kBlock, 0, s>>>(c.f_key2, m, c.f_pos, c.f_nnz);
    int rc = scan_offsets(c, c.f_nnz, c.f_off, m, s);
    if (rc) return rc;
    k_cub_coo<<<grid_for(m), kBlock, 0, s>>>(c.f_key2, field, s0, s1, s2, m, c.f_pos, c.f_off,
                                              rows1, cols1, values, dims);
    OPH_CHECK(hipGetLastError());
    return OPENPH_OK;
}

}  // namespace oph

using namespace oph;

extern "C" {

int openph_vr2_size(int n, int64_t *m, int64_t *nnz) {
    if (n < 0) return OPENPH_EINVAL;
    const long long E = choose2(n), T = choose3(n);
    if (m) *m = n + E + T;
    if (nnz) *nnz = 2 * E + 3 * T;
    return OPENPH_OK;
}

int openph_vr2_device(const double *points, int n, int dim, int *rows1, int *cols1, double *values,
                      int *dims, void *stream, int device) {
    if (n < 0 || dim < 1 || dim > 7) return OPENPH_EINVAL;
    if (n == 0) return OPENPH_OK;
    int rc = set_device(device);
    if (rc) return rc;
    Context *cp = context_for(device);
    std::lock_guard<std::mutex> g(cp->mu);
    hipStream_t s = (hipStream_t)stream;
    if ((rc = build_vr(*cp, points, n, dim, rows1, cols1, values, dims, s))) return rc;
    OPH_CHECK(hipStreamSynchronize(s));
    return OPENPH_OK;
}

int openph_cubical_size(int s0, int s1, int s2, int64_t *m, int64_t *nnz) {
    if (s0 < 1 || s1 < 1 || s2 < 1) return OPENPH_EINVAL;
    const long long N0 = 2LL * s0 - 1, N1 = 2LL * s1 - 1, N2 = 2LL * s2 - 1;
    if (m) *m = N0 * N1 * N2;
    // every odd coordinate of a cell gives it two facets
    if (nnz) *nnz = 2 * ((s0 - 1) * N1 * N2 + N0 * (s1 - 1) * N2 + N0 * N1 * (s2 - 1));
    return OPENPH_OK;
}

int openph_cubical_device(const float *field, int s0, int s1, int s2, int *rows1, int *cols1,
                          float *values, int *dims, void *stream, int device) {
    int rc = openph_cubical_size(s0, s1, s2, nullptr, nullptr);
    if (rc) return rc;
    if ((rc = set_device(device))) return rc;
    Context *cp = context_for(device);
    std::lock_guard<std::mutex> g(cp->mu);
    hipStream_t s = (hipStream_t)stream;
    if ((rc = build_cubical(*cp, field, s0, s1, s2, rows1, cols1, values, dims, s))) return rc;
    OPH_CHECK(hipStreamSynchronize(s));
    return OPENPH_OK;
}

}  // extern "C"
