// Shared by the FFT / split-step translation units: the split-step kernel parameters, the
// local reservoir and nonlinear steps of the split-step schemes, and the column group shapes.
#pragma once

#include <cmath>

#include "dispersion.hpp"
#include "fft_core.hpp"
#include "fft_plan.hpp"

namespace pgpe {

enum RowMode {
    ROW_FWD = 0,
    ROW_INV = 1,
    ROW_SS_FIRST = 2,
    ROW_SS_MID = 3,
    ROW_SS_LAST = 4,
    ROW_KO_MID = 5,
    ROW_SS_JOIN = 6  // CPU order, between chained steps: SS_LAST's work, then the next step's SS_FIRST
};
enum ColMode { COL_FWD = 0, COL_INV = 1, COL_PROP = 2 };

template <typename T>
struct SplitParams {
    const pgpe_member* table;
    FftDims dims;
    const c2<T>* tw;
    const double* kx;
    const double* ky;
    c2<T>* in;
    c2<T>* out;
    int nx, ny, planes;
    int reservoir;        // pgpe_reservoir_kind
    int noise;
    T dt;
    double prop_h;        // propagator step (dt/2 for a half step)
    T scale;              // inverse-transform normalisation
    T* res[2];
    const T* pump;
    const T* coef;        // coef[0*members + m]: first half-step, coef[1*members + m]: second
    int members;
    int comps;            // 2: spinor planes (psi+, psi-) per member
    int res_comps;        // spinor reservoirs: 2 per component, 1 shared
    uint64_t step;
    const uint64_t* step_base;
};

template <typename T, int RES>
__device__ __forceinline__ void res_rhs(const pgpe_member& mb, T rho, T P, const T n[2], T d[2]) {
    if constexpr (RES == PGPE_RES_SINGLE) {
        d[0] = P - ((T)mb.res_a + (T)mb.R * rho) * n[0];
    } else if constexpr (RES == PGPE_RES_DOUBLE) {
        d[1] = (P - (T)mb.res_a * n[1]) + (T)mb.res_b * n[0];
        d[0] = (T)mb.res_c * n[1] - ((T)mb.res_d + (T)mb.R * rho) * n[0];
    } else if constexpr (RES == PGPE_RES_QDOUBLE) {
        const T transfer = (T)mb.res_a * n[1] * n[1];
        d[0] = (transfer - (T)mb.res_b * n[0]) - (T)mb.R * n[0] * rho;
        d[1] = (P - transfer) - (T)mb.res_c * n[1];
    }
}

// Midpoint RK2 with psi frozen, then clamp at zero (single_reservoir.py:100-111).
template <typename T, int RES>
__device__ __forceinline__ void res_half_step(const pgpe_member& mb, T rho, T P, T h, T n[2]) {
    constexpr int NF = RES == 0 ? 0 : (RES == 1 ? 1 : 2);
    if constexpr (NF == 0) return;
    T k1[2] = {0, 0}, k2[2] = {0, 0}, mid[2] = {0, 0};
    res_rhs<T, RES>(mb, rho, P, n, k1);
    const T hh = (T)0.5 * h;
#pragma unroll
    for (int f = 0; f < NF; ++f) mid[f] = n[f] + hh * k1[f];
    res_rhs<T, RES>(mb, rho, P, mid, k2);
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        const T v = n[f] + h * k2[f];
        n[f] = v > (T)0 ? v : (T)0;
    }
}

// Exact local gain/phase step (split_step_fft_solver.py:124-140), zero potential.
template <typename T>
__device__ __forceinline__ void nonlinear_step(const pgpe_member& mb, T nA, T nI, T dt, c2<T>& v) {
    const T rho = v.x * v.x + v.y * v.y;
    const T gain = ((T)mb.R * nA - (T)mb.gamma_C) / (T)2;
    const T x = (T)2 * gain * dt;
    const T phase_time = fabs(x) < (T)1e-6 ? dt * (((T)1 + x / (T)2) + x * x / (T)6) : dt * (expm1(x) / x);
    const T phase = (((T)mb.g_R * nA + (T)mb.g_I * nI) * dt + (T)mb.g_C * rho * phase_time) / (T)mb.hbar;
    const T amp = exp(gain * dt);
    T s, c;
    sin_cos(phase, &s, &c);
    // exp(gain dt - i phase) = amp (cos phase - i sin phase)
    const c2<T> f = make_c<T>(amp * c, -(amp * s));
    v = cmul<T>(v, f);
}

template <typename T, int RES>
__device__ __forceinline__ void load_res(const SplitParams<T>& p, int64_t g, T n[2]) {
    constexpr int NF = RES == 0 ? 0 : (RES == 1 ? 1 : 2);
#pragma unroll
    for (int f = 0; f < NF; ++f) n[f] = p.res[f][g];
}

template <typename T, int RES>
__device__ __forceinline__ void store_res(const SplitParams<T>& p, int64_t g, const T n[2]) {
    constexpr int NF = RES == 0 ? 0 : (RES == 1 ? 1 : 2);
#pragma unroll
    for (int f = 0; f < NF; ++f) p.res[f][g] = n[f];
}

// Column group shapes (threads x elements per thread), in the order they are tried.
constexpr int kColShapes[][2] = {{64, 4}, {128, 4}, {256, 4}, {256, 8}, {256, 16}, {256, 32}, {512, 32}};

// Kernel parameters of one split step (pgpe_split_args, the plan's wave numbers).
template <typename T>
inline SplitParams<T> split_params(const pgpe_split_args* a) {
    const pgpe_fft_plan* plan = (const pgpe_fft_plan*)a->plan;
    SplitParams<T> p{};
    p.table = a->table;
    p.kx = plan->kx;
    p.ky = plan->ky;
    p.nx = plan->nx;
    p.ny = plan->ny;
    p.planes = a->members;
    p.members = a->members;
    p.reservoir = a->reservoir;
    p.noise = a->noise;
    p.dt = (T)a->dt;
    p.prop_h = a->dt / 2.0;
    p.scale = (T)(1.0 / ((double)plan->nx * (double)plan->ny));
    p.res[0] = (T*)a->res[0];
    p.res[1] = (T*)a->res[1];
    p.pump = (const T*)a->pump;
    p.coef = (const T*)a->coef;
    p.step = a->step;
    p.step_base = a->step_base;
    p.comps = a->components == 2 ? 2 : 1;
    p.res_comps = a->res_components == 2 ? 2 : 1;
    return p;
}

// One spinor split step (fft_spinor.hip).
int split_spinor_step(const pgpe_split_args* a, hipStream_t s);

}  // namespace pgpe
