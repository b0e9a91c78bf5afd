// In-run observables accumulated on the device in double precision
// (results/accumulators.py:94-188): the per-mode power |psi_k|^2 of the momentum
// distribution and the density moments rho, rho^2 of g2.  Elementwise and
// HBM-bound: each launch reads the field (c bytes per point) and reads and writes
// its FP64 sums (16 or 32 bytes per point); 4 points per thread, grid-stride.
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
