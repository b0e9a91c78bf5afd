// Shared device/host helpers for libpgpe_hip.so (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/pgpe.h"

namespace pgpe {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t err, const char* what);

// Looks up (once per device, with that device current) the device address of the
// allocation-free reduction's scratch array; `out` may be null.
int reduce_scratch_setup(double** out);

#define PGPE_HIP(call)                                          \
    do {                                                        \
        hipError_t e_ = (call);                                 \
        if (e_ != hipSuccess) return ::pgpe::hip_fail(e_, #call); \
    } while (0)

// Check the launch that was just issued.
#define PGPE_LAUNCHED(name)                                     \
    do {                                                        \
        hipError_t e_ = hipGetLastError();                      \
        if (e_ != hipSuccess) return ::pgpe::hip_fail(e_, name); \
    } while (0)

// ---------------------------------------------------------------- constants
constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kTileW = 64;         // stencil tiles are one wavefront wide
constexpr int kXcds = 8;           // MI355X: 8 XCDs, workgroups dispatched round-robin

// Map the hardware workgroup id to a logical tile so that each XCD (own L2)
// receives a contiguous range of tiles: vertical neighbours of a tile, which
// re-read its halo rows, then usually run on the same XCD.  Returns -1 for
// the padding workgroups of a grid rounded up to a multiple of kXcds.
__device__ __forceinline__ int xcd_tile(int hw_block, int num_tiles) {
    const int per = (num_tiles + kXcds - 1) / kXcds;
    const int logical = (hw_block % kXcds) * per + hw_block / kXcds;
    return logical < num_tiles ? logical : -1;
}

inline int xcd_grid(int num_tiles) {
    const int per = (num_tiles + kXcds - 1) / kXcds;
    return per * kXcds;
}

// ---------------------------------------------------------------- complex
template <typename T> struct vec2;
template <> struct vec2<float> { using type = float2; };
template <> struct vec2<double> { using type = double2; };
template <typename T> using c2 = typename vec2<T>::type;

template <typename T>
__host__ __device__ __forceinline__ c2<T> make_c(T re, T im) {
    c2<T> v;
    v.x = re;
    v.y = im;
    return v;
}

// Load of one complex value as a native two-element vector: a plain copy of a double2 goes
// through a 16-byte memcpy, which keeps a struct that receives it (and arrays of such structs)
// in scratch memory.
template <typename T>
__device__ __forceinline__ c2<T> ld_c(const c2<T>* p) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 t = *reinterpret_cast<const v2*>(p);
    return make_c<T>(t.x, t.y);
}

__device__ __forceinline__ void sin_cos(float x, float* s, float* c) { sincosf(x, s, c); }
__device__ __forceinline__ void sin_cos(double x, double* s, double* c) { sincos(x, s, c); }

// ---------------------------------------------------------------- Philox4x32-10
// Counter (cell, step_lo, step_hi, stream), key = 64-bit seed split lo/hi
// (noise/philox.py:7-38).  Multipliers and Weyl increments are the published
// Random123 constants.
struct u32x4 { uint32_t a, b, c, d; };

__device__ __forceinline__ u32x4 philox10(u32x4 x, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = x.a * 0xD2511F53u;
        const uint32_t hi0 = __umulhi(x.a, 0xD2511F53u);
        const uint32_t lo1 = x.c * 0xCD9E8D57u;
        const uint32_t hi1 = __umulhi(x.c, 0xCD9E8D57u);
        u32x4 y;
        y.a = hi1 ^ x.b ^ k0;
        y.b = lo1;
        y.c = hi0 ^ x.d ^ k1;
        y.d = lo0;
        x = y;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return x;
}

// Box-Muller normals from the words, precision-specific uniforms (philox.py:41-56).
__device__ __forceinline__ void box_muller(const u32x4& w, float& n_re, float& n_im) {
    const float u1 = (float)((w.a >> 8) + 1u) * 5.9604644775390625e-08f;
    const float u2 = (float)((w.b >> 8) + 1u) * 5.9604644775390625e-08f;
    const float r = sqrtf(-2.0f * logf(u1));
    const float ang = 6.2831853071795864769f * u2;
    float s, c;
    sincosf(ang, &s, &c);
    n_re = r * c;
    n_im = r * s;
}

__device__ __forceinline__ void box_muller(const u32x4& w, double& n_re, double& n_im) {
    const double u1 = ((double)(w.a >> 5) * 67108864.0 + (double)(w.b >> 6) + 1.0) *
                      1.1102230246251565404236316680908203125e-16;
    const double u2 = ((double)(w.c >> 5) * 67108864.0 + (double)(w.d >> 6) + 1.0) *
                      1.1102230246251565404236316680908203125e-16;
    const double r = sqrt(-2.0 * log(u1));
    const double ang = 6.283185307179586476925286766559 * u2;
    double s, c;
    sincos(ang, &s, &c);
    n_re = r * c;
    n_im = r * s;
}

// Wigner increment for one site: psi += sigma (N1 + i N2) with
// sigma = sqrt(strength (gamma + R n) dt/(4 dx dy)) (noise/philox.py:133-152).
template <typename T>
__device__ __forceinline__ void wigner_site(T& re, T& im, uint64_t seed, uint32_t cell,
                                            uint64_t step, uint32_t stream_word, T strength,
                                            T gamma, T scatter, T active, T dt_over_4v) {
    u32x4 x{cell, (uint32_t)step, (uint32_t)(step >> 32), stream_word};
    const u32x4 w = philox10(x, (uint32_t)seed, (uint32_t)(seed >> 32));
    T n_re, n_im;
    box_muller(w, n_re, n_im);
    const T sigma = sqrt(strength * (gamma + scatter * active) * dt_over_4v);
    re += sigma * n_re;
    im += sigma * n_im;
}

}  // namespace pgpe
