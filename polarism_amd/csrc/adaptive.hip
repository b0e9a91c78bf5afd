// On-device control of adaptive Lawson trials.
//
// The reference decides every trial on the host: the embedded error is reduced to
// a Python float (ifrk_fft_cuda_solver.py:293), the PI controller updates h there
// (solver/step_controller.py:48-81) and the next trial is launched from Python
// (abstract_solver.py:86-96).  Here the controller state lives in device memory and
// three small kernels bracket each trial, so a run of trials can be replayed from one
// captured graph and the host looks at the state once per graph launch:
//
//   propose  h = h_next, clamped to land exactly on the target tick; underflow guard
//   decide   member norms from the FP64 error sums, max over members, accept when <= 1,
//            PI update of h_next, time and counters; done once t reaches the target
//   commit   copies the trial's new state over the current one when it was accepted
//
// The kernels of the trial read h from the state (pgpe_lawson_args.h_dev) and return at
// once when the state is done (skip_dev), so trials past the target cost only launches.
// All arithmetic is FP64 with the controller's constants (0.9, -0.7/k, 0.4/k, [0.2, 5],
// max(0.2, 0.9 err^(-1/k)), err_prev >= 1e-4, 16 eps max(1, |t|)) in the reference's
// operation order, so trial sequences match the host controller's.
#include "common.hpp"

#include <cfloat>

namespace pgpe {

__global__ void adaptive_propose_kernel(pgpe_adaptive_state* st) {
    if (threadIdx.x != 0 || st->done) return;
    const double t = st->t, target = st->target;
    const double prop = st->h_next;
    const int clamped = prop >= target - t;
    const double h = clamped ? target - t : prop;
    st->prop = prop;
    st->clamped = clamped;
    st->h = h;
    st->commit = 0;
    if (h < 16.0 * DBL_EPSILON * fmax(1.0, fabs(t))) {
        st->status = PGPE_ADAPTIVE_UNDERFLOW;
        st->done = 1;
    }
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ int wave_any(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
    return v;
}

// One wavefront: lanes stride over the members' (err5, err3) sums.
__global__ void adaptive_decide_kernel(pgpe_adaptive_state* st, const double* __restrict__ sums) {
    const int lane = threadIdx.x;
    if (st->done) {
        if (lane == 0) st->commit = 0;
        return;
    }
    const int members = st->members;
    const double count = st->count;
    const int blend = st->blend;
    double worst = 0.0;
    int nan_seen = 0;
    for (int m = lane; m < members; m += 64) {
        const double e5 = sums[2 * m], e3 = sums[2 * m + 1];
        double v;
        if (blend) {
            // DOP853: err5 / sqrt((err5 + 0.01 err3) count), 0 when the denominator vanishes
            const double den = e5 + 0.01 * e3;
            v = den > 0 ? e5 / sqrt(den * count) : 0.0;
        } else {
            v = sqrt(e5 / count);  // DP5
        }
        if (v != v) nan_seen = 1;
        else worst = fmax(worst, v);
    }
    worst = wave_max(worst);
    nan_seen = wave_any(nan_seen);
    if (lane != 0) return;
    const double err = nan_seen ? (double)NAN : worst;
    const int accepted = err <= 1.0;
    const double error = (err >= 0.0 && err < INFINITY) ? err : INFINITY;
    const double h = st->h, k = st->k;
    double factor;
    if (accepted) {
        st->t = st->clamped ? st->target : st->t + h;
        st->accepted += 1;
        st->dt_min = fmin(st->dt_min, h);
        st->dt_max_used = fmax(st->dt_max_used, h);
        st->dt_last = h;
        factor = error == 0.0 ? 5.0 : 0.9 * pow(error, -0.7 / k) * pow(st->err_prev, 0.4 / k);
        factor = fmin(5.0, fmax(0.2, factor));
        if (st->no_growth) factor = fmin(1.0, factor);
        st->err_prev = fmax(error, 1e-4);
        st->no_growth = 0;
    } else {
        st->rejected += 1;
        factor = fmin(1.0, fmax(0.2, 0.9 * pow(error, -1.0 / k)));
        st->no_growth = 1;
    }
    double h_next = h * factor;
    if (accepted && st->clamped) h_next = fmax(h_next, st->prop);
    if (st->dt_max > 0.0) h_next = fmin(h_next, st->dt_max);
    st->h_next = h_next;
    st->error = err;
    st->commit = accepted;
    if (!(st->t < st->target)) st->done = 1;
}

// Copies of the accepted state: 16-byte words, then the 4-byte tail of each region.
__global__ void __launch_bounds__(256) adaptive_commit_kernel(const pgpe_adaptive_state* st, pgpe_commit_list list) {
    if (!st->commit) return;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int r = 0; r < list.count; ++r) {
        const int64_t n16 = (int64_t)(list.bytes[r] / 16);
        const uint4* src = static_cast<const uint4*>(list.src[r]);
        uint4* dst = static_cast<uint4*>(list.dst[r]);
        for (int64_t i = tid; i < n16; i += stride) dst[i] = src[i];
        const int64_t words = (int64_t)(list.bytes[r] % 16) / 4;
        if (tid < words) {
            const uint32_t* s4 = reinterpret_cast<const uint32_t*>(src + n16);
            uint32_t* d4 = reinterpret_cast<uint32_t*>(dst + n16);
            d4[tid] = s4[tid];
        }
    }
}

}  // namespace pgpe

using namespace pgpe;

extern "C" int pgpe_adaptive_propose(pgpe_adaptive_state* state, void* stream) {
    if (!state) return fail(PGPE_ERR_ARG, "pgpe_adaptive_propose: null state");
    hipLaunchKernelGGL(adaptive_propose_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, state);
    PGPE_LAUNCHED("adaptive_propose_kernel");
    return PGPE_OK;
}

extern "C" int pgpe_adaptive_decide(pgpe_adaptive_state* state, const double* error_sums, void* stream) {
    if (!state || !error_sums) return fail(PGPE_ERR_ARG, "pgpe_adaptive_decide: null pointer");
    hipLaunchKernelGGL(adaptive_decide_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, state, error_sums);
    PGPE_LAUNCHED("adaptive_decide_kernel");
    return PGPE_OK;
}

extern "C" int pgpe_adaptive_commit(const pgpe_adaptive_state* state, const pgpe_commit_list* list, void* stream) {
    if (!state || !list) return fail(PGPE_ERR_ARG, "pgpe_adaptive_commit: null pointer");
    if (list->count < 0 || list->count > PGPE_COMMIT_MAX) return fail(PGPE_ERR_ARG, "pgpe_adaptive_commit: bad count");
    uint64_t total = 0;
    for (int r = 0; r < list->count; ++r) {
        if (!list->dst[r] || !list->src[r]) return fail(PGPE_ERR_ARG, "pgpe_adaptive_commit: null region");
        if (list->bytes[r] % 4) return fail(PGPE_ERR_ARG, "pgpe_adaptive_commit: sizes must be multiples of 4 bytes");
        if ((reinterpret_cast<uintptr_t>(list->dst[r]) | reinterpret_cast<uintptr_t>(list->src[r])) % 16)
            return fail(PGPE_ERR_ARG, "pgpe_adaptive_commit: regions must be 16-byte aligned");
        total += list->bytes[r];
    }
    const uint64_t blocks = (total / 16 + 255) / 256;
    const unsigned grid = (unsigned)(blocks < 1 ? 1 : (blocks > 4096 ? 4096 : blocks));
    hipLaunchKernelGGL(adaptive_commit_kernel, dim3(grid), dim3(256), 0, (hipStream_t)stream, state, *list);
    PGPE_LAUNCHED("adaptive_commit_kernel");
    return PGPE_OK;
}
