// Per-member density reductions for the steady-state test
// (simulation_controller.py:710-734):  sum rho, sum (rho - old)^2, sum old^2,
// accumulated in double, deterministic (fixed-order two-pass reduction).
#include "common.hpp"

#include <mutex>

namespace pgpe {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Partial sums of the allocation-free entry point: 3 doubles per tile of the member table,
// at most kScratchTiles tiles (1.5 MiB).  Each device holds its own copy until the process ends.
constexpr int kScratchTiles = 65536;
static __device__ __align__(256) double scratch[3 * kScratchTiles];

// `cap` bounds the tile index (tile_begin + local) that may be written or read in `partials`.
template <typename T>
__global__ void __launch_bounds__(256) density_delta_tiles(const c2<T>* __restrict__ psi, T* __restrict__ old,
                                                           int components, const pgpe_member* __restrict__ table,
                                                           const int32_t* __restrict__ tile_map, int num_tiles,
                                                           double* __restrict__ partials, int update, int64_t cap) {
    __shared__ double s_part[3][4];
    const int tile = xcd_tile(blockIdx.x, num_tiles);
    if (tile < 0) return;
    const int m = tile_map[2 * tile];
    const pgpe_member& mb = table[m];
    const int local = tile_map[2 * tile + 1];
    const int x = (local % mb.tiles_x) * kTileW + (threadIdx.x & 63);
    const int ybase = (local / mb.tiles_x) * mb.tile_h;
    const int64_t plane = (int64_t)mb.nx * mb.ny;
    double s_rho = 0.0, s_delta = 0.0, s_old = 0.0;
    if (x < mb.nx) {
        for (int r = threadIdx.x >> 6; r < mb.tile_h; r += 4) {
            const int y = ybase + r;
            if (y >= mb.ny) break;
            const int64_t cell = (int64_t)y * mb.nx + x;
            T rho = (T)0;
            for (int c = 0; c < components; ++c) {
                const c2<T> v = psi[mb.off * components + c * plane + cell];
                rho += v.x * v.x + v.y * v.y;
            }
            const T o = old[mb.off + cell];
            const double d = (double)rho - (double)o;
            s_rho += (double)rho;
            s_delta += d * d;
            s_old += (double)o * (double)o;
            if (update) old[mb.off + cell] = rho;
        }
    }
    s_rho = wave_sum(s_rho);
    s_delta = wave_sum(s_delta);
    s_old = wave_sum(s_old);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_part[0][w] = s_rho;
        s_part[1][w] = s_delta;
        s_part[2][w] = s_old;
    }
    __syncthreads();
    const int64_t slot = (int64_t)mb.tile_begin + local;
    if (threadIdx.x < 3 && slot < cap) {
        const int q = threadIdx.x;
        partials[3 * slot + q] = ((s_part[q][0] + s_part[q][1]) + s_part[q][2]) + s_part[q][3];
    }
}

// One wavefront per member sums its tiles' partials in a fixed order.
__global__ void density_delta_members(const pgpe_member* __restrict__ table, const double* __restrict__ partials,
                                      double* __restrict__ sums, int members, int64_t cap) {
    const int m = blockIdx.x;
    if (m >= members) return;
    const pgpe_member& mb = table[m];
    const int n = mb.tiles_x * mb.tiles_y;
    const int lane = threadIdx.x;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int t = lane; t < n; t += 64) {
        const int64_t tile = mb.tile_begin + t;
        if (tile >= cap) break;
        for (int q = 0; q < 3; ++q) acc[q] += partials[3 * tile + q];
    }
    for (int q = 0; q < 3; ++q) {
        const double v = wave_sum(acc[q]);
        if (lane == 0) sums[3 * m + q] = v;
    }
}

__global__ void zero_partials(double* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch[kMaxDevices] = {};
}  // namespace

// Device address of `scratch` on the current device, looked up once per device.  Host code
// uses only this address: the array's own name is a host-side stand-in.
int reduce_scratch_setup(double** out) {
    int dev = 0;
    PGPE_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(PGPE_ERR_ARG, "reduce scratch: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        PGPE_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(PGPE_ERR_HIP, "reduce scratch: hipGetSymbolAddress returned null");
        g_scratch[dev] = static_cast<double*>(p);
    }
    if (out) *out = g_scratch[dev];
    return PGPE_OK;
}

static int launch_density_delta(const void* psi, void* old, int precision, int components, const pgpe_member* table,
                                const int32_t* tile_map, int num_tiles, int members, double* partials, double* sums,
                                int update, int64_t cap, hipStream_t s) {
    const dim3 grid(xcd_grid(num_tiles));
    if (precision == 0)
        hipLaunchKernelGGL(density_delta_tiles<float>, grid, dim3(256), 0, s, (const float2*)psi, (float*)old,
                           components, table, tile_map, num_tiles, partials, update, cap);
    else if (precision == 1)
        hipLaunchKernelGGL(density_delta_tiles<double>, grid, dim3(256), 0, s, (const double2*)psi, (double*)old,
                           components, table, tile_map, num_tiles, partials, update, cap);
    else
        return fail(PGPE_ERR_ARG, "pgpe_density_delta: precision must be 0 or 1");
    PGPE_LAUNCHED("density_delta_tiles");
    hipLaunchKernelGGL(density_delta_members, dim3(members), dim3(64), 0, s, table, partials, sums, members, cap);
    PGPE_LAUNCHED("density_delta_members");
    return PGPE_OK;
}

}  // namespace pgpe

using namespace pgpe;

extern "C" int pgpe_density_delta(const void* psi, void* old, int precision, int components, const pgpe_member* table,
                                  const int32_t* tile_map, int num_tiles, int members, double* partials,
                                  double* sums, int update, void* stream) {
    if (!psi || !old || !table || !tile_map || !partials || !sums)
        return fail(PGPE_ERR_ARG, "pgpe_density_delta: null pointer");
    if (num_tiles <= 0 || members <= 0) return fail(PGPE_ERR_ARG, "pgpe_density_delta: empty table");
    if (components != 1 && components != 2) return fail(PGPE_ERR_ARG, "pgpe_density_delta: components must be 1 or 2");
    return launch_density_delta(psi, old, precision, components, table, tile_map, num_tiles, members, partials, sums,
                                update, INT64_MAX, (hipStream_t)stream);
}

extern "C" int pgpe_density_delta_noalloc(const void* psi, void* old, int precision, int components,
                                          const pgpe_member* table, const int32_t* tile_map, int num_tiles,
                                          int members, int total_tiles, double* sums, int update, void* stream) {
    if (!psi || !old || !table || !tile_map || !sums)
        return fail(PGPE_ERR_ARG, "pgpe_density_delta_noalloc: null pointer");
    if (num_tiles <= 0 || members <= 0 || total_tiles <= 0)
        return fail(PGPE_ERR_ARG, "pgpe_density_delta_noalloc: empty table");
    if (total_tiles > kScratchTiles)
        return fail(PGPE_ERR_ARG, "pgpe_density_delta_noalloc: the member table has " + std::to_string(total_tiles) +
                                      " tiles; the built-in scratch holds at most " + std::to_string(kScratchTiles));
    if (num_tiles > total_tiles)
        return fail(PGPE_ERR_ARG, "pgpe_density_delta_noalloc: more active tiles than the table holds");
    if (components != 1 && components != 2)
        return fail(PGPE_ERR_ARG, "pgpe_density_delta_noalloc: components must be 1 or 2");
    if (precision != 0 && precision != 1)
        return fail(PGPE_ERR_ARG, "pgpe_density_delta_noalloc: precision must be 0 or 1");
    double* partials = nullptr;
    int rc = reduce_scratch_setup(&partials);
    if (rc != PGPE_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    // members whose tiles are not in the map read zeros, not the previous call's partials
    const int64_t n = 3 * (int64_t)total_tiles;
    hipLaunchKernelGGL(zero_partials, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, partials, n);
    PGPE_LAUNCHED("zero_partials");
    return launch_density_delta(psi, old, precision, components, table, tile_map, num_tiles, members, partials, sums,
                                update, (int64_t)total_tiles, s);
}
