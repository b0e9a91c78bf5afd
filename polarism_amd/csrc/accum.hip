// In-run observables accumulated on the device in double precision
// (results/accumulators.py:94-270): the per-mode power |psi_k|^2 of the momentum
// distribution, the density moments rho, rho^2 of g2, and the pieces of the time
// spectra (line cuts into a ring, the windowed block in time order, the power of its
// transform).  Elementwise and HBM-bound; grid-stride loops.
#include "common.hpp"

namespace pgpe {

// sum[i] += re^2 + im^2, the components widened to double first (accumulators.py:106-108).
template <typename T>
__global__ void __launch_bounds__(256) accumulate_power_kernel(const c2<T>* __restrict__ f, double* __restrict__ sum,
                                                               int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const c2<T> v = f[i];
        const double re = (double)v.x, im = (double)v.y;
        sum[i] += re * re + im * im;
    }
}

// rho = re^2 + im^2 in double; sum_rho += rho, sum_rho2 += rho * rho (accumulators.py:147-152).
template <typename T>
__global__ void __launch_bounds__(256) accumulate_moments_kernel(const c2<T>* __restrict__ f, double* __restrict__ s1,
                                                                 double* __restrict__ s2, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const c2<T> v = f[i];
        const double re = (double)v.x, im = (double)v.y;
        const double rho = re * re + im * im;
        s1[i] += rho;
        s2[i] += rho * rho;
    }
}

// Line cut of every plane into one ring slot: mode 0 copies row `row` of the field
// (spectrum-x); mode 1 takes row ky = 0 of the field's 2-D transform, times `scale`, in
// fftshift order along x (spectrum-k: FFT_x of the y-sum is that row, accumulators.py:206-215).
template <typename T>
__global__ void spectrum_line_kernel(const c2<T>* __restrict__ f, c2<T>* __restrict__ slot, int planes, int nx, int ny,
                                     int row, int mode, T scale) {
    const int64_t n = (int64_t)planes * nx;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int plane = (int)(i / nx), x = (int)(i - (int64_t)plane * nx);
        if (mode == 0) {
            slot[i] = f[((int64_t)plane * ny + row) * nx + x];
        } else {
            c2<T> v = f[(int64_t)plane * ny * nx + x];
            v.x *= scale;
            v.y *= scale;
            slot[(int64_t)plane * nx + (x + nx / 2) % nx] = v;
        }
    }
}

// block[t][j] = ring[(position + t) mod segment][j] * window[t] in double (time order, oldest
// first; the reference multiplies the complex ring by a float64 window, accumulators.py:217-219).
template <typename T>
__global__ void spectrum_block_kernel(const c2<T>* __restrict__ ring, double2* __restrict__ block,
                                      const double* __restrict__ window, int segment, int64_t width, int position) {
    const int64_t n = (int64_t)segment * width;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int t = (int)(i / width);
        const int64_t j = i - (int64_t)t * width;
        int src = position + t;
        if (src >= segment) src -= segment;
        const c2<T> v = ring[(int64_t)src * width + j];
        const double w = window[t];
        block[i] = make_double2((double)v.x * w, (double)v.y * w);
    }
}

// psd[(line * segment + f) * nx + x] += |B[f][line * nx + x]|^2 (the power moved to the
// (..., E, x) layout of accumulators.py:220-222).
__global__ void spectrum_power_kernel(const double2* __restrict__ bk, double* __restrict__ psd, int segment,
                                      int64_t width, int nx) {
    const int64_t n = (int64_t)segment * width;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int f = (int)(i / width);
        const int64_t j = i - (int64_t)f * width;
        const int64_t line = j / nx, x = j - line * nx;
        const double2 v = bk[i];
        psd[(line * segment + f) * nx + x] += v.x * v.x + v.y * v.y;
    }
}

// Result-frame observables of a field laid out (B, C, plane) (simulation_controller.py:63-89):
// density |psi_c|^2 per component; for two components the Stokes vector in the circular basis
// (S0, S1, S2, S3) = (|p|^2 + |m|^2, 2 Re(conj(p) m), 2 Im(conj(p) m), |p|^2 - |m|^2), laid out
// (B, 4, plane), and the degree of polarisation |(S1, S2, S3)| / S0 (0 where S0 = 0), (B, plane).
// Null outputs are skipped.
template <typename T>
__global__ void __launch_bounds__(256) frame_fields_kernel(const c2<T>* __restrict__ psi, int comps, int64_t plane,
                                                           int64_t n, T* __restrict__ density,
                                                           T* __restrict__ stokes, T* __restrict__ dop) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int64_t b = i / plane, s = i - b * plane;
        const c2<T> p = psi[b * comps * plane + s];
        const T dp = p.x * p.x + p.y * p.y;
        if (comps == 1) {
            if (density) density[i] = dp;
            continue;
        }
        const c2<T> m = psi[(b * 2 + 1) * plane + s];
        const T dm = m.x * m.x + m.y * m.y;
        if (density) {
            density[b * 2 * plane + s] = dp;
            density[(b * 2 + 1) * plane + s] = dm;
        }
        const T s0 = dp + dm;
        const T s1 = (T)2 * (p.x * m.x + p.y * m.y);
        const T s2 = (T)2 * (p.x * m.y - p.y * m.x);
        const T s3 = dp - dm;
        if (stokes) {
            T* o = stokes + b * 4 * plane + s;
            o[0] = s0;
            o[plane] = s1;
            o[2 * plane] = s2;
            o[3 * plane] = s3;
        }
        if (dop) {
            const T pol = sqrt(s1 * s1 + s2 * s2 + s3 * s3);
            dop[i] = s0 > (T)0 ? pol / (s0 > (T)1e-30 ? s0 : (T)1e-30) : (T)0;
        }
    }
}

static dim3 accum_grid(int64_t n) {
    const int64_t blocks = (n + 4 * 256 - 1) / (4 * 256);
    return dim3((unsigned)(blocks < 1 ? 1 : (blocks > 65535 ? 65535 : blocks)));
}

}  // namespace pgpe

using namespace pgpe;

extern "C" int pgpe_accumulate_power(const void* field, double* sum, int64_t n, int precision, void* stream) {
    if (!field || !sum) return fail(PGPE_ERR_ARG, "pgpe_accumulate_power: null pointer");
    if (n <= 0) return fail(PGPE_ERR_ARG, "pgpe_accumulate_power: n must be positive");
    hipStream_t s = (hipStream_t)stream;
    if (precision == 0)
        hipLaunchKernelGGL(accumulate_power_kernel<float>, accum_grid(n), dim3(256), 0, s, (const float2*)field, sum, n);
    else if (precision == 1)
        hipLaunchKernelGGL(accumulate_power_kernel<double>, accum_grid(n), dim3(256), 0, s, (const double2*)field, sum,
                           n);
    else
        return fail(PGPE_ERR_ARG, "pgpe_accumulate_power: precision must be 0 or 1");
    PGPE_LAUNCHED("accumulate_power_kernel");
    return PGPE_OK;
}

extern "C" int pgpe_accumulate_moments(const void* field, double* sum_rho, double* sum_rho2, int64_t n, int precision,
                                       void* stream) {
    if (!field || !sum_rho || !sum_rho2) return fail(PGPE_ERR_ARG, "pgpe_accumulate_moments: null pointer");
    if (n <= 0) return fail(PGPE_ERR_ARG, "pgpe_accumulate_moments: n must be positive");
    hipStream_t s = (hipStream_t)stream;
    if (precision == 0)
        hipLaunchKernelGGL(accumulate_moments_kernel<float>, accum_grid(n), dim3(256), 0, s, (const float2*)field,
                           sum_rho, sum_rho2, n);
    else if (precision == 1)
        hipLaunchKernelGGL(accumulate_moments_kernel<double>, accum_grid(n), dim3(256), 0, s, (const double2*)field,
                           sum_rho, sum_rho2, n);
    else
        return fail(PGPE_ERR_ARG, "pgpe_accumulate_moments: precision must be 0 or 1");
    PGPE_LAUNCHED("accumulate_moments_kernel");
    return PGPE_OK;
}

extern "C" int pgpe_spectrum_line(const void* field, void* slot, int planes, int nx, int ny, int row, int mode,
                                  double scale, int precision, void* stream) {
    if (!field || !slot) return fail(PGPE_ERR_ARG, "pgpe_spectrum_line: null pointer");
    if (planes <= 0 || nx <= 0 || ny <= 0 || (mode != 0 && mode != 1)) return fail(PGPE_ERR_ARG, "pgpe_spectrum_line: bad size or mode");
    if (mode == 0 && (row < 0 || row >= ny)) return fail(PGPE_ERR_ARG, "pgpe_spectrum_line: row out of range");
    hipStream_t s = (hipStream_t)stream;
    const dim3 grid = accum_grid((int64_t)planes * nx);
    if (precision == 0)
        hipLaunchKernelGGL(spectrum_line_kernel<float>, grid, dim3(256), 0, s, (const float2*)field, (float2*)slot,
                           planes, nx, ny, row, mode, (float)scale);
    else if (precision == 1)
        hipLaunchKernelGGL(spectrum_line_kernel<double>, grid, dim3(256), 0, s, (const double2*)field, (double2*)slot,
                           planes, nx, ny, row, mode, scale);
    else
        return fail(PGPE_ERR_ARG, "pgpe_spectrum_line: precision must be 0 or 1");
    PGPE_LAUNCHED("spectrum_line_kernel");
    return PGPE_OK;
}

extern "C" int pgpe_spectrum_block(const void* ring, void* block, const double* window, int segment, int64_t width,
                                   int position, int precision, void* stream) {
    if (!ring || !block || !window) return fail(PGPE_ERR_ARG, "pgpe_spectrum_block: null pointer");
    if (segment <= 0 || width <= 0 || position < 0 || position >= segment)
        return fail(PGPE_ERR_ARG, "pgpe_spectrum_block: bad size or position");
    hipStream_t s = (hipStream_t)stream;
    const dim3 grid = accum_grid((int64_t)segment * width);
    if (precision == 0)
        hipLaunchKernelGGL(spectrum_block_kernel<float>, grid, dim3(256), 0, s, (const float2*)ring, (double2*)block,
                           window, segment, width, position);
    else if (precision == 1)
        hipLaunchKernelGGL(spectrum_block_kernel<double>, grid, dim3(256), 0, s, (const double2*)ring,
                           (double2*)block, window, segment, width, position);
    else
        return fail(PGPE_ERR_ARG, "pgpe_spectrum_block: precision must be 0 or 1");
    PGPE_LAUNCHED("spectrum_block_kernel");
    return PGPE_OK;
}

extern "C" int pgpe_spectrum_power(const void* block_k, double* psd, int segment, int64_t width, int nx, void* stream) {
    if (!block_k || !psd) return fail(PGPE_ERR_ARG, "pgpe_spectrum_power: null pointer");
    if (segment <= 0 || width <= 0 || nx <= 0 || width % nx) return fail(PGPE_ERR_ARG, "pgpe_spectrum_power: bad size");
    hipLaunchKernelGGL(spectrum_power_kernel, accum_grid((int64_t)segment * width), dim3(256), 0, (hipStream_t)stream,
                       (const double2*)block_k, psd, segment, width, nx);
    PGPE_LAUNCHED("spectrum_power_kernel");
    return PGPE_OK;
}

extern "C" int pgpe_frame_fields(const void* psi, int precision, int components, int64_t plane, int64_t members,
                                 void* density, void* stokes, void* dop, void* stream) {
    if (!psi) return fail(PGPE_ERR_ARG, "pgpe_frame_fields: null field");
    if (components != 1 && components != 2) return fail(PGPE_ERR_ARG, "pgpe_frame_fields: components must be 1 or 2");
    if (components == 1 && (stokes || dop))
        return fail(PGPE_ERR_ARG, "pgpe_frame_fields: Stokes fields need two components");
    if (plane <= 0 || members <= 0) return fail(PGPE_ERR_ARG, "pgpe_frame_fields: bad size");
    const int64_t n = plane * members;
    hipStream_t s = (hipStream_t)stream;
    if (precision == 0)
        hipLaunchKernelGGL(frame_fields_kernel<float>, accum_grid(n), dim3(256), 0, s, (const float2*)psi, components,
                           plane, n, (float*)density, (float*)stokes, (float*)dop);
    else if (precision == 1)
        hipLaunchKernelGGL(frame_fields_kernel<double>, accum_grid(n), dim3(256), 0, s, (const double2*)psi,
                           components, plane, n, (double*)density, (double*)stokes, (double*)dop);
    else
        return fail(PGPE_ERR_ARG, "pgpe_frame_fields: precision must be 0 or 1");
    PGPE_LAUNCHED("frame_fields_kernel");
    return PGPE_OK;
}
