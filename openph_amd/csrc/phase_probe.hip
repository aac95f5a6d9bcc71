// Cost of a phase boundary: what a persistent (cooperative) form of the pms
// iteration would trade against. The iteration is a chain of grid-wide
// passes over the columns (k_first, k_tile, k_mark, k_apply1, Phase II,
// k_apply2), each of which reads what the previous one wrote anywhere in the
// arrays, so passes are separated either by kernel boundaries or by grid
// barriers. This probe runs the same chain of `phases` passes over m columns
// (each pass: y[j] = x[j] + x[(j + m / 2) mod m] modulo 2^32, 4 columns per
// thread, the arrays swapped between passes; an even number of passes, so
// that a chain ends where the next starts) in four forms and reports the
// time:
//   0  one launch per pass, queued on a stream;
//   1  the same launches replayed from a hipGraph;
//   2  one cooperative launch per chain, a grid barrier between passes
//      (one arrival counter for the grid);
//   3  the same with a two-level barrier: one counter per group of blocks
//      (block index mod 8, the XCD a block is dealt to), then one for the
//      groups.
// Every form must leave the same array; the probe returns its checksum.
// Built as its own library (libopenph_probe.so); nothing in libopenph.so
// uses it.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace {

constexpr int kBlock = 256, kItems = 4, kTile = kBlock * kItems;
constexpr int kGroups = 8;
constexpr unsigned kSpinLimit = 1u << 24;  // about a second of s_sleep, then give up

__device__ __forceinline__ void pass_tile(const unsigned *__restrict__ x, unsigned *__restrict__ y,
                                          int m, int tile) {
    const int base = tile * kTile + threadIdx.x * kItems;
    const int half = m / 2;
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        const int j = base + k;
        if (j < m) {
            int o = j + half;
            o = o >= m ? o - m : o;
            y[j] = x[j] + x[o];
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_pass(const unsigned *x, unsigned *y, int m) {
    pass_tile(x, y, m, blockIdx.x);
}

// bar[0]: arrivals (grid), bar[32 * (1 + g)]: arrivals of group g, bar[16]:
// generation, bar[17]: set when a wait gave up. Counters only grow, so a
// barrier's target is (generation + 1) times the number of arrivals.
__device__ void grid_barrier(unsigned *bar, bool two_level, unsigned gen) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();  // this block's writes, before its arrival
        bool last = false;
        if (!two_level) {
            last = atomicAdd(&bar[0], 1u) == (gen + 1) * gridDim.x - 1;
        } else {
            const unsigned g = blockIdx.x % kGroups;
            const unsigned groups = gridDim.x < kGroups ? gridDim.x : kGroups;
            const unsigned gsize = (gridDim.x - g + kGroups - 1) / kGroups;
            if (atomicAdd(&bar[32 * (1 + g)], 1u) == (gen + 1) * gsize - 1)
                last = atomicAdd(&bar[0], 1u) == (gen + 1) * groups - 1;
        }
        if (last) {
            __hip_atomic_store(&bar[16], gen + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            unsigned spins = 0;
            while (__hip_atomic_load(&bar[16], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) <
                   gen + 1) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > kSpinLimit) {
                    atomicOr(&bar[17], 1u);
                    break;
                }
            }
        }
        __threadfence();  // the other blocks' writes, after the barrier
    }
    __syncthreads();
}

__global__ void __launch_bounds__(kBlock) k_passes(unsigned *x, unsigned *y, int m, int phases,
                                                   unsigned *bar, int two_level, unsigned gen0) {
    const int ntiles = (m + kTile - 1) / kTile;
    for (int p = 0; p < phases; p++) {
        const unsigned *src = (p & 1) ? y : x;
        unsigned *dst = (p & 1) ? x : y;
        for (int t = blockIdx.x; t < ntiles; t += gridDim.x) pass_tile(src, dst, m, t);
        if (p + 1 < phases) grid_barrier(bar, two_level != 0, gen0 + (unsigned)p);
    }
}

__global__ void k_init(unsigned *x, int m) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < m) x[j] = (j * 2654435761u) >> 20;
}

}  // namespace

extern "C" {

// Runs `reps` chains of `phases` passes over m columns in form `mode`
// (0 queued, 1 graph, 2 cooperative flat barrier, 3 cooperative two-level
// barrier; the cooperative grid is the resident blocks, at most max_blocks
// when that is positive); ms receives the time of all chains, blocks the
// grid, checksum the sum of the final array. Returns 0, a HIP
// error code, or -1 for bad arguments, -2 when a barrier wait gave up.
int openph_probe_phases(int m, int phases, int reps, int mode, int max_blocks, float *ms,
                        int *blocks, long long *checksum) {
    // an even chain ends in x, where the next one starts
    if (m < 2 || phases < 2 || (phases & 1) || reps < 1 || mode < 0 || mode > 3 || !ms) return -1;
    unsigned *x = nullptr, *y = nullptr, *bar = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipSuccess;
    bool gave_up = false;
    const int ntiles = (m + kTile - 1) / kTile;
    int grid = ntiles;
    auto chk = [&](hipError_t r) {
        if (e == hipSuccess && r != hipSuccess) e = r;
        return e == hipSuccess;
    };
    if (chk(hipMalloc((void **)&x, sizeof(unsigned) * m)) &&
        chk(hipMalloc((void **)&y, sizeof(unsigned) * m)) &&
        chk(hipMalloc((void **)&bar, sizeof(unsigned) * 32 * (kGroups + 1))) &&
        chk(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) && chk(hipEventCreate(&e0)) &&
        chk(hipEventCreate(&e1))) {
        k_init<<<(m + 255) / 256, 256, 0, s>>>(x, m);
        chk(hipGetLastError());
        chk(hipMemsetAsync(bar, 0, sizeof(unsigned) * 32 * (kGroups + 1), s));
        if (mode >= 2) {
            int dev = 0, per_cu = 0;
            hipDeviceProp_t prop;
            if (chk(hipGetDevice(&dev)) && chk(hipGetDeviceProperties(&prop, dev)) &&
                chk(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_passes, kBlock, 0))) {
                int resident = per_cu * prop.multiProcessorCount;
                if (max_blocks > 0 && max_blocks < resident) resident = max_blocks;
                grid = ntiles < resident ? ntiles : resident;
            }
        }
        if (mode == 1 && chk(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal))) {
            for (int p = 0; p < phases; p++)
                k_pass<<<ntiles, kBlock, 0, s>>>((p & 1) ? y : x, (p & 1) ? x : y, m);
            chk(hipStreamEndCapture(s, &graph)) && chk(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        }
        chk(hipStreamSynchronize(s));
        unsigned gen = 0;
        for (int r = 0; r < reps + 1 && e == hipSuccess; r++) {  // the first chain untimed
            if (r == 1) chk(hipEventRecord(e0, s));
            if (mode == 0) {
                for (int p = 0; p < phases; p++)
                    k_pass<<<ntiles, kBlock, 0, s>>>((p & 1) ? y : x, (p & 1) ? x : y, m);
                chk(hipGetLastError());
            } else if (mode == 1) {
                chk(hipGraphLaunch(exec, s));
            } else {
                int two = mode == 3 ? 1 : 0;
                void *args[] = {&x, &y, &m, &phases, &bar, &two, &gen};
                chk(hipLaunchCooperativeKernel((const void *)k_passes, dim3(grid), dim3(kBlock),
                                               args, 0, s));
                gen += (unsigned)(phases - 1);
            }
        }
        chk(hipEventRecord(e1, s));
        chk(hipStreamSynchronize(s));
        float t = 0.f;
        chk(hipEventElapsedTime(&t, e0, e1));
        *ms = t;
        if (blocks) *blocks = grid;
        unsigned flags[18] = {0};
        chk(hipMemcpy(flags, bar, sizeof(flags), hipMemcpyDeviceToHost));
        std::vector<unsigned> h((size_t)m);
        chk(hipMemcpy(h.data(), x, sizeof(unsigned) * m, hipMemcpyDeviceToHost));
        long long sum = 0;
        for (unsigned v : h) sum += v;
        if (checksum) *checksum = sum;
        gave_up = flags[17] != 0;
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (x) (void)hipFree(x);
    if (y) (void)hipFree(y);
    if (bar) (void)hipFree(bar);
    if (e != hipSuccess) return (int)e;
    return gave_up ? -2 : 0;
}

}  // extern "C"
