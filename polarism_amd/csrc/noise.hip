// Philox4x32-10 words, Box-Muller pairs and the truncated-Wigner increment.
//
// Counter layout (cell, step_lo, step_hi, stream) and key = member seed follow
// noise/philox.py:30-38 and :86-99 so that every member/site/step draws the
// same normals as the reference generator.
#include "common.hpp"

namespace pgpe {

__global__ void philox_words_kernel(const uint4* __restrict__ ctr, uint4* __restrict__ out, uint64_t key,
                                    int64_t n) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint4 c = ctr[j];
    const u32x4 w = philox10(u32x4{c.x, c.y, c.z, c.w}, (uint32_t)key, (uint32_t)(key >> 32));
    out[j] = make_uint4(w.a, w.b, w.c, w.d);
}

template <typename T>
__global__ void normal_pairs_kernel(c2<T>* __restrict__ out, uint64_t key, uint64_t step, uint32_t stream_word,
                                    int64_t n) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const u32x4 w = philox10(u32x4{(uint32_t)j, (uint32_t)step, (uint32_t)(step >> 32), stream_word},
                             (uint32_t)key, (uint32_t)(key >> 32));
    T re, im;
    box_muller(w, re, im);
    out[j] = make_c<T>(re, im);
}

// One workgroup per 64 x tile_h tile of one member; each thread walks rows.
template <typename T>
__global__ void __launch_bounds__(256) wigner_add_kernel(c2<T>* __restrict__ psi, const T* __restrict__ active,
                                                         int active_components, const pgpe_member* __restrict__ table,
                                                         const int32_t* __restrict__ tile_map, int num_tiles,
                                                         uint64_t step) {
    const int tile = xcd_tile(blockIdx.x, num_tiles);
    if (tile < 0) return;
    const int m = tile_map[2 * tile];
    const pgpe_member& mb = table[m];
    const int local = tile_map[2 * tile + 1];
    const int tx = local % mb.tiles_x;
    const int ty = local / mb.tiles_x;
    const int x = tx * kTileW + (threadIdx.x & 63);
    if (x >= mb.nx) return;
    const int64_t plane = (int64_t)mb.nx * mb.ny;
    const T strength = (T)mb.noise_strength, gamma = (T)mb.noise_gamma, scat = (T)mb.noise_R;
    const T dt4v = (T)mb.noise_dt_over_4v;
    for (int c = 0; c < mb.components; ++c) {
        c2<T>* p = psi + mb.off * mb.components + c * plane;
        const T* a = active ? active + mb.off * active_components + (active_components == 2 ? c : 0) * plane
                            : nullptr;
        for (int r = threadIdx.x >> 6; r < mb.tile_h; r += 4) {
            const int y = ty * mb.tile_h + r;
            if (y >= mb.ny) break;
            const int64_t cell = (int64_t)y * mb.nx + x;
            c2<T> v = p[cell];
            const T dens = a ? a[cell] : (T)0;
            wigner_site<T>(v.x, v.y, mb.seed, (uint32_t)cell, step, 1u + (uint32_t)c, strength, gamma, scat,
                           dens, dt4v);
            p[cell] = v;
        }
    }
}

}  // namespace pgpe

using namespace pgpe;

extern "C" {

int pgpe_philox_words(const uint32_t* counters, uint32_t* out, uint64_t key, int64_t n, void* stream) {
    if (n < 0) return fail(PGPE_ERR_ARG, "pgpe_philox_words: negative n");
    if (n == 0) return PGPE_OK;
    if (!counters || !out) return fail(PGPE_ERR_ARG, "pgpe_philox_words: null pointer");
    const int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(philox_words_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       (const uint4*)counters, (uint4*)out, key, n);
    PGPE_LAUNCHED("philox_words_kernel");
    return PGPE_OK;
}

int pgpe_normal_pairs(void* out, int precision, uint64_t key, uint64_t step, uint32_t stream_word, int64_t n,
                      void* stream) {
    if (n < 0) return fail(PGPE_ERR_ARG, "pgpe_normal_pairs: negative n");
    if (n == 0) return PGPE_OK;
    if (!out) return fail(PGPE_ERR_ARG, "pgpe_normal_pairs: null pointer");
    const int64_t blocks = (n + 255) / 256;
    if (precision == 0) {
        hipLaunchKernelGGL(normal_pairs_kernel<float>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                           (float2*)out, key, step, stream_word, n);
    } else if (precision == 1) {
        hipLaunchKernelGGL(normal_pairs_kernel<double>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                           (double2*)out, key, step, stream_word, n);
    } else {
        return fail(PGPE_ERR_ARG, "pgpe_normal_pairs: precision must be 0 or 1");
    }
    PGPE_LAUNCHED("normal_pairs_kernel");
    return PGPE_OK;
}

int pgpe_wigner_add(void* psi, const void* active, int active_components, const pgpe_member* table,
                    const int32_t* tile_map, int num_tiles, int members, int precision, uint64_t step,
                    void* stream) {
    if (!psi || !table || !tile_map) return fail(PGPE_ERR_ARG, "pgpe_wigner_add: null pointer");
    if (num_tiles <= 0 || members <= 0) return fail(PGPE_ERR_ARG, "pgpe_wigner_add: empty table");
    if (active && active_components != 1 && active_components != 2)
        return fail(PGPE_ERR_ARG, "pgpe_wigner_add: active_components must be 1 or 2");
    const dim3 grid(xcd_grid(num_tiles));
    if (precision == 0) {
        hipLaunchKernelGGL(wigner_add_kernel<float>, grid, dim3(256), 0, (hipStream_t)stream, (float2*)psi,
                           (const float*)active, active_components, table, tile_map, num_tiles, step);
    } else if (precision == 1) {
        hipLaunchKernelGGL(wigner_add_kernel<double>, grid, dim3(256), 0, (hipStream_t)stream, (double2*)psi,
                           (const double*)active, active_components, table, tile_map, num_tiles, step);
    } else {
        return fail(PGPE_ERR_ARG, "pgpe_wigner_add: precision must be 0 or 1");
    }
    PGPE_LAUNCHED("wigner_add_kernel");
    return PGPE_OK;
}

}  // extern "C"
