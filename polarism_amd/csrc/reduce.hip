// Per-member density reductions for the steady-state test
// (simulation_controller.py:710-734):  sum rho, sum (rho - old)^2, sum old^2,
// accumulated in double, deterministic (fixed-order two-pass reduction).
#include "common.hpp"

namespace pgpe {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <typename T>
__global__ void __launch_bounds__(256) density_delta_tiles(const c2<T>* __restrict__ psi, T* __restrict__ old,
                                                           int components, const pgpe_member* __restrict__ table,
                                                           const int32_t* __restrict__ tile_map, int num_tiles,
                                                           double* __restrict__ partials, int update) {
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
    if (threadIdx.x < 3) {
        const int q = threadIdx.x;
        partials[3 * ((int64_t)mb.tile_begin + local) + q] = ((s_part[q][0] + s_part[q][1]) + s_part[q][2]) + s_part[q][3];
    }
}

// One wavefront per member sums its tiles' partials in a fixed order.
__global__ void density_delta_members(const pgpe_member* __restrict__ table, const double* __restrict__ partials,
                                      double* __restrict__ sums, int members) {
    const int m = blockIdx.x;
    if (m >= members) return;
    const pgpe_member& mb = table[m];
    const int n = mb.tiles_x * mb.tiles_y;
    const int lane = threadIdx.x;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int t = lane; t < n; t += 64) {
        const int64_t tile = mb.tile_begin + t;
        for (int q = 0; q < 3; ++q) acc[q] += partials[3 * tile + q];
    }
    for (int q = 0; q < 3; ++q) {
        const double v = wave_sum(acc[q]);
        if (lane == 0) sums[3 * m + q] = v;
    }
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
    hipStream_t s = (hipStream_t)stream;
    const dim3 grid(xcd_grid(num_tiles));
    if (precision == 0)
        hipLaunchKernelGGL(density_delta_tiles<float>, grid, dim3(256), 0, s, (const float2*)psi, (float*)old,
                           components, table, tile_map, num_tiles, partials, update);
    else if (precision == 1)
        hipLaunchKernelGGL(density_delta_tiles<double>, grid, dim3(256), 0, s, (const double2*)psi, (double*)old,
                           components, table, tile_map, num_tiles, partials, update);
    else
        return fail(PGPE_ERR_ARG, "pgpe_density_delta: precision must be 0 or 1");
    PGPE_LAUNCHED("density_delta_tiles");
    hipLaunchKernelGGL(density_delta_members, dim3(members), dim3(64), 0, s, table, partials, sums, members);
    PGPE_LAUNCHED("density_delta_members");
    return PGPE_OK;
}
