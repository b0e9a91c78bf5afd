// Spinor split step on gfx950 (split_step_fft_solver.py:46-53, 81-94, 124-145 and the
// K-outer order of split_step_fft_cuda_solver.py:106-184): the five passes of the scalar step
// over row pairs and column groups that hold psi+ and psi- together.
#include "split_common.hpp"

namespace pgpe {

// ---------------------------------------------------------------- spinor
// Spinor split-step (split_step_fft_solver.py:46-53, 81-94, 124-145): member m holds the planes
// psi+ (2m) and psi- (2m + 1).  A row group transforms row y of both components (sequences 2j
// and 2j + 1 for its j-th row pair), so the local physics sees both at every point: reservoirs
// per component (density |psi_c|^2, pump P w_c) or one shared reservoir (density |psi+|^2 +
// |psi-|^2, pump P); the exact gain/phase step per component with the cross-spin phase
// alpha_2 |psi_o|^2 tau_o / hbar of the other component (tau = phase time); Wigner noise with
// stream 1 + c and the active density of the component's reservoir.

// Exact local gain/phase step of both components (split_step_fft_solver.py:124-145).
template <typename T>
__device__ __forceinline__ void nonlinear_step_spinor(const pgpe_member& mb, const T nA[2], const T nI[2], T dt,
                                                      c2<T> v[2]) {
    T rho[2], gain[2], pt[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        rho[c] = v[c].x * v[c].x + v[c].y * v[c].y;
        gain[c] = ((T)mb.R * nA[c] - (T)mb.gamma_C) / (T)2;
        const T x = (T)2 * gain[c] * dt;
        pt[c] = fabs(x) < (T)1e-6 ? dt * (((T)1 + x / (T)2) + x * x / (T)6) : dt * (expm1(x) / x);
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int o = 1 - c;
        T phase = (((T)mb.g_R * nA[c] + (T)mb.g_I * nI[c]) * dt + (T)mb.g_C * rho[c] * pt[c]) / (T)mb.hbar;
        phase = phase + (T)mb.alpha_2 * rho[o] * pt[o] / (T)mb.hbar;
        const T amp = exp(gain[c] * dt);
        T sn, cs;
        sin_cos(phase, &sn, &cs);
        v[c] = cmul<T>(v[c], make_c<T>(amp * cs, -(amp * sn)));
    }
}

// Reservoir midpoint half-steps of one point: per component or shared (split_step_fft_solver.py:81-94).
template <typename T, int RES>
__device__ __forceinline__ void res_half_step_spinor(const pgpe_member& mb, int res_comps, const c2<T> v[2], T P,
                                                     T h, T n[2][2]) {
    const T rho0 = v[0].x * v[0].x + v[0].y * v[0].y;
    const T rho1 = v[1].x * v[1].x + v[1].y * v[1].y;
    if (res_comps == 2) {
        res_half_step<T, RES>(mb, rho0, P * (T)mb.pump_w_plus, h, n[0]);
        res_half_step<T, RES>(mb, rho1, P * (T)mb.pump_w_minus, h, n[1]);
    } else {
        res_half_step<T, RES>(mb, rho0 + rho1, P, h, n[0]);
    }
}

template <typename T, int V, int MODE, int RES, int NT>
__global__ void __launch_bounds__(NT) fft_rows_spinor_kernel(const SplitParams<T> p) {
    extern __shared__ unsigned char smem_raw[];
    c2<T>* buf = reinterpret_cast<c2<T>*>(smem_raw);
    constexpr int E = NT * V;
    constexpr int NF = RES == 0 ? 0 : (RES == 1 ? 1 : 2);
    const int n = p.nx;
    const int pairs_per = E / (2 * n);
    const int ld = seq_stride<T>(n);
    c2<T>* s_tw = buf + 2 * pairs_per * ld;
    const int total_pairs = p.members * p.ny;
    const int pair0 = blockIdx.x * pairs_per;
    const int count = min(pairs_per, total_pairs - pair0);
    if (count <= 0) return;
    stage_twiddles<T, NT>(s_tw, p.tw, n);
    const FastDiv by_n((uint32_t)n);
    const int elems = count * n;
    const int RC = p.res_comps == 2 ? 2 : 1;
    const int64_t plane = (int64_t)p.ny * n;
    uint64_t step = p.step;
    if (p.step_base) step += *p.step_base;
    // pair element e: row pair j = e / n, column k; J = pair0 + j -> member m, row y
    struct At {
        int m, j;
        int64_t cell;  // y * n + k inside the member's plane
        int slot;      // LDS slot of component 0 (component 1 is ld further... see slot1)
    };
    auto locate = [&](int e) {
        const int j = by_n.div(e);
        const int k = e - j * n;
        const int J = pair0 + j;
        const int m = J / p.ny;
        At a;
        a.m = m;
        a.j = j;
        a.cell = (int64_t)(J - m * p.ny) * n + k;
        a.slot = 2 * j * ld + pad_at(k);
        return a;
    };
    auto psi_at = [&](const At& a, int c) { return ((int64_t)2 * a.m + c) * plane + a.cell; };
    auto res_at = [&](const At& a, int rc) { return ((int64_t)a.m * RC + rc) * plane + a.cell; };
    // reservoir loops are unrolled over both slots (guarded by RC) so the arrays stay in registers
    auto load_n = [&](const At& a, T nn[2][2]) {
#pragma unroll
        for (int rc = 0; rc < 2; ++rc)
#pragma unroll
            for (int f = 0; f < 2; ++f) nn[rc][f] = (T)0;
#pragma unroll
        for (int rc = 0; rc < 2; ++rc)
            if (rc < RC) {
#pragma unroll
                for (int f = 0; f < NF; ++f) nn[rc][f] = p.res[f][res_at(a, rc)];
            }
    };
    auto store_n = [&](const At& a, const T nn[2][2]) {
#pragma unroll
        for (int rc = 0; rc < 2; ++rc)
            if (rc < RC) {
#pragma unroll
                for (int f = 0; f < NF; ++f) p.res[f][res_at(a, rc)] = nn[rc][f];
            }
    };
    constexpr int CH = stage_chunk<T, V>() < 8 ? stage_chunk<T, V>() : 8;
    struct Pt { c2<T> v[2]; T nn[2][2]; T P1, P2; };

    // ---- load (+ reservoir half-step of the CPU order) ----
    staged<Pt, CH, NT>(
        elems,
        [&](int e) {
            const At a = locate(e);
            Pt q;
            q.v[0] = ld_c<T>(&p.in[psi_at(a, 0)]);
            q.v[1] = ld_c<T>(&p.in[psi_at(a, 1)]);
            if constexpr (MODE == ROW_SS_FIRST) {
                load_n(a, q.nn);
                q.P1 = p.pump[p.table[a.m].pump_off + a.cell] * p.coef[a.m];
            }
            return q;
        },
        [&](int e, Pt q) {
            const At a = locate(e);
            if constexpr (MODE == ROW_SS_FIRST) {
                if constexpr (NF > 0) {
                    res_half_step_spinor<T, RES>(p.table[a.m], p.res_comps, q.v, q.P1, p.dt / (T)2, q.nn);
                    store_n(a, q.nn);
                }
            }
            buf[a.slot] = q.v[0];
            buf[a.slot + ld] = q.v[1];
        });
    __syncthreads();
    constexpr bool first_inverse = MODE == ROW_INV || MODE == ROW_SS_MID || MODE == ROW_SS_LAST || MODE == ROW_KO_MID;
    lds_fft<T, V, NT>(buf, p.dims, ld, 2 * count, s_tw, first_inverse);

    if constexpr (MODE == ROW_SS_MID || MODE == ROW_KO_MID) {
        staged<Pt, CH, NT>(
            elems,
            [&](int e) {
                const At a = locate(e);
                Pt q;
                q.v[0] = ld_c<T>(&buf[a.slot]);
                q.v[1] = ld_c<T>(&buf[a.slot + ld]);
                load_n(a, q.nn);
                if constexpr (MODE == ROW_KO_MID) {
                    const T env = p.pump[p.table[a.m].pump_off + a.cell];
                    q.P1 = env * p.coef[a.m];
                    q.P2 = env * p.coef[p.members + a.m];
                }
                return q;
            },
            [&](int e, Pt q) {
                const At a = locate(e);
                const pgpe_member& mb = p.table[a.m];
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    q.v[c].x *= p.scale;
                    q.v[c].y *= p.scale;
                }
                if constexpr (MODE == ROW_KO_MID && NF > 0)
                    res_half_step_spinor<T, RES>(mb, p.res_comps, q.v, q.P1, p.dt / (T)2, q.nn);
                T nA[2], nI[2];
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    const bool own = RC == 2 && c == 1;  // component 1's own reservoir, else slot 0
                    nA[c] = NF > 0 ? (own ? q.nn[1][0] : q.nn[0][0]) : (T)0;
                    nI[c] = NF > 1 ? (own ? q.nn[1][1] : q.nn[0][1]) : (T)0;
                }
                nonlinear_step_spinor<T>(mb, nA, nI, p.dt, q.v);
                if constexpr (MODE == ROW_KO_MID && NF > 0) {
                    res_half_step_spinor<T, RES>(mb, p.res_comps, q.v, q.P2, p.dt / (T)2, q.nn);
                    store_n(a, q.nn);
                }
                if (p.noise) {
#pragma unroll
                    for (int c = 0; c < 2; ++c) {
                        const T act = NF > 0 ? ((RC == 2 && c == 1) ? q.nn[1][0] : q.nn[0][0]) : (T)0;
                        wigner_site<T>(q.v[c].x, q.v[c].y, mb.seed, (uint32_t)a.cell, step, 1u + (uint32_t)c,
                                       (T)mb.noise_strength, (T)mb.noise_gamma, (T)mb.noise_R, act,
                                       (T)mb.noise_dt_over_4v);
                    }
                }
                buf[a.slot] = q.v[0];
                buf[a.slot + ld] = q.v[1];
            });
        __syncthreads();
        lds_fft<T, V, NT>(buf, p.dims, ld, 2 * count, s_tw, false);
    }

    // ---- store (+ reservoir half-step of the CPU order) ----
    staged<Pt, CH, NT>(
        elems,
        [&](int e) {
            const At a = locate(e);
            Pt q;
            q.v[0] = ld_c<T>(&buf[a.slot]);
            q.v[1] = ld_c<T>(&buf[a.slot + ld]);
            if constexpr (MODE == ROW_SS_LAST) {
                load_n(a, q.nn);
                q.P1 = p.pump[p.table[a.m].pump_off + a.cell] * p.coef[p.members + a.m];
            }
            return q;
        },
        [&](int e, Pt q) {
            const At a = locate(e);
            if constexpr (MODE == ROW_SS_LAST || MODE == ROW_INV) {
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    q.v[c].x *= p.scale;
                    q.v[c].y *= p.scale;
                }
            }
            if constexpr (MODE == ROW_SS_LAST) {
                if constexpr (NF > 0) {
                    res_half_step_spinor<T, RES>(p.table[a.m], p.res_comps, q.v, q.P1, p.dt / (T)2, q.nn);
                    store_n(a, q.nn);
                }
            }
            p.out[psi_at(a, 0)] = q.v[0];
            p.out[psi_at(a, 1)] = q.v[1];
        });
}

// Column groups of a spinor member: the same columns of psi+ and psi- (sequences c * count + i),
// forward along y, exp(-i H(k) dt / (2 hbar)) at every k (dispersion.py:66-83, in double), inverse.
template <typename T, int V, int NT>
__global__ void __launch_bounds__(NT) fft_cols_spinor_kernel(const SplitParams<T> p) {
    extern __shared__ unsigned char smem_raw[];
    c2<T>* buf = reinterpret_cast<c2<T>*>(smem_raw);
    constexpr int E = NT * V;
    const int n = p.ny;
    const int cols_per = E / (2 * n);
    const int ld = seq_stride<T>(n);
    c2<T>* s_tw = buf + 2 * cols_per * ld;
    const int groups_per_member = (p.nx + cols_per - 1) / cols_per;
    const int g = xcd_tile(blockIdx.x, groups_per_member * p.members);
    if (g < 0) return;
    const int m = g / groups_per_member;
    const int x0 = (g - m * groups_per_member) * cols_per;
    const int count = min(cols_per, p.nx - x0);
    stage_twiddles<T, NT>(s_tw, p.tw, n);
    const FastDiv by_count((uint32_t)count), by_n((uint32_t)n);
    constexpr int CH = stage_chunk<T, V>();
    struct Pt { c2<T> v; int at; };
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int64_t base = ((int64_t)2 * m + c) * n * p.nx;
        staged<Pt, CH, NT>(
            count * n,
            [&](int i) {
                const int y = by_count.div(i);
                const int col = i - y * count;
                return Pt{ld_c<T>(&p.in[base + (int64_t)y * p.nx + x0 + col]), (c * count + col) * ld + pad_at(y)};
            },
            [&](int, Pt q) { buf[q.at] = q.v; });
    }
    __syncthreads();
    lds_fft<T, V, NT>(buf, p.dims, ld, 2 * count, s_tw, false);
    const pgpe_member& mb = p.table[m];
    struct K { double kx, ky; };
    // the wave-number pairs are 16 bytes each: a shorter batch keeps them out of scratch at V = 32
    constexpr int CK = CH > 8 ? 8 : CH;
    staged<K, CK, NT>(
        count * n, [&](int i) { const int col = by_n.div(i); return K{p.kx[x0 + col], p.ky[i - col * n]}; },
        [&](int i, K q) {
            const int col = by_n.div(i);
            const int y = i - col * n;
            const int s0 = col * ld + pad_at(y), s1 = (count + col) * ld + pad_at(y);
            c2<T> v[2] = {buf[s0], buf[s1]}, o[2];
            spinor_propagate<T>(mb, q.kx, q.ky, p.prop_h, v, o);
            buf[s0] = o[0];
            buf[s1] = o[1];
        });
    __syncthreads();
    lds_fft<T, V, NT>(buf, p.dims, ld, 2 * count, s_tw, true);
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int64_t base = ((int64_t)2 * m + c) * n * p.nx;
        staged<Pt, CH, NT>(
            count * n,
            [&](int i) {
                const int y = by_count.div(i);
                const int col = i - y * count;
                return Pt{buf[(c * count + col) * ld + pad_at(y)], y * p.nx + x0 + col};
            },
            [&](int, Pt q) { p.out[base + q.at] = q.v; });
    }
}

template <typename T, int V, int MODE, int RES, int NT>
static int launch_rows_spinor_v(const pgpe_fft_plan* plan, const SplitParams<T>& p0, hipStream_t s) {
    constexpr int E = NT * V;
    SplitParams<T> p = p0;
    p.dims = plan->dims_x[radix_class<V>()];
    p.tw = (const c2<T>*)plan->tw_x[radix_class<V>()];
    const int pairs_per = E / (2 * plan->nx);
    const int total = p.members * p.ny;
    const size_t shm = (size_t)(2 * pairs_per * seq_stride<T>(plan->nx) + tw_slots<T>(plan->nx)) * sizeof(c2<T>);
    if (pairs_per < 1 || shm > 163840) return fail(PGPE_ERR_UNSUPPORTED, "split-step: spinor rows too long");
    auto kern = fft_rows_spinor_kernel<T, V, MODE, RES, NT>;
    if (shm > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute(fft_rows_spinor_kernel)");
    }
    hipLaunchKernelGGL(kern, dim3((total + pairs_per - 1) / pairs_per), dim3(NT), shm, s, p);
    PGPE_LAUNCHED("fft_rows_spinor_kernel");
    return PGPE_OK;
}

// A spinor row group holds one row of both components at least: 256 threads x 4, 8 or 16
// elements (several row pairs per group on short rows), 12 with a radix 3.
template <typename T, int MODE, int RES>
static int launch_rows_spinor(const pgpe_fft_plan* plan, const SplitParams<T>& p, hipStream_t s) {
    const int n2 = 2 * plan->nx;
    if (plan->nx % 3 == 0) {
        if (n2 > 256 * 12) return fail(PGPE_ERR_UNSUPPORTED, "split-step: spinor rows longer than 1536");
        return launch_rows_spinor_v<T, 12, MODE, RES, 256>(plan, p, s);
    }
    if (n2 <= 1024) return launch_rows_spinor_v<T, 4, MODE, RES, 256>(plan, p, s);
    if (n2 <= 2048) return launch_rows_spinor_v<T, 8, MODE, RES, 256>(plan, p, s);
    if (n2 <= 4096) return launch_rows_spinor_v<T, 16, MODE, RES, 256>(plan, p, s);
    return fail(PGPE_ERR_UNSUPPORTED, "split-step: spinor rows longer than 2048");
}

template <typename T, int V, int NT>
static int launch_cols_spinor_v(const pgpe_fft_plan* plan, const SplitParams<T>& p0, hipStream_t s) {
    constexpr int E = NT * V;
    SplitParams<T> p = p0;
    p.dims = plan->dims_y[radix_class<V>()];
    p.tw = (const c2<T>*)plan->tw_y[radix_class<V>()];
    const int cols_per = E / (2 * plan->ny);
    const int groups = (plan->nx + cols_per - 1) / cols_per;
    const size_t shm = (size_t)(2 * cols_per * seq_stride<T>(plan->ny) + tw_slots<T>(plan->ny)) * sizeof(c2<T>);
    auto kern = fft_cols_spinor_kernel<T, V, NT>;
    if (shm > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute(fft_cols_spinor_kernel)");
    }
    hipLaunchKernelGGL(kern, dim3(xcd_grid(groups * p.members)), dim3(NT), shm, s, p);
    PGPE_LAUNCHED("fft_cols_spinor_kernel");
    return PGPE_OK;
}

// Spinor column groups: the column shapes with room for both components' columns (a group
// holds cols_per columns of psi+ and of psi-), chosen as for scalar columns.
template <typename T>
static int launch_cols_spinor(const pgpe_fft_plan* plan, const SplitParams<T>& p, hipStream_t s) {
    const int n = plan->ny;
    auto lds = [&](int cols) { return (size_t)(2 * cols * seq_stride<T>(n) + tw_slots<T>(n)) * sizeof(c2<T>); };
    if (n % 3 == 0) {
        if (256 * 12 < 2 * n || lds(256 * 12 / (2 * n)) > 163840)
            return fail(PGPE_ERR_UNSUPPORTED, "split-step: spinor columns longer than 1536");
        return launch_cols_spinor_v<T, 12, 256>(plan, p, s);
    }
    const int cols_total = plan->nx * p.members;
    int want = sizeof(T) == 4 ? 8 : 4;
    while (want > 1 && cols_total / want < 512) want /= 2;
    int best = -1;
    for (int k = 0; k < 7; ++k) {
        const int cols = kColShapes[k][0] * kColShapes[k][1] / (2 * n);
        if (cols < 1) continue;
        if (lds(cols) > 163840) break;
        best = k;
        if (cols >= want) break;
    }
    switch (best) {
        case 0: return launch_cols_spinor_v<T, 4, 64>(plan, p, s);
        case 1: return launch_cols_spinor_v<T, 4, 128>(plan, p, s);
        case 2: return launch_cols_spinor_v<T, 4, 256>(plan, p, s);
        case 3: return launch_cols_spinor_v<T, 8, 256>(plan, p, s);
        case 4: return launch_cols_spinor_v<T, 16, 256>(plan, p, s);
        case 5: return launch_cols_spinor_v<T, 32, 256>(plan, p, s);
        case 6: return launch_cols_spinor_v<T, 32, 512>(plan, p, s);
        default: return fail(PGPE_ERR_UNSUPPORTED, "split-step: spinor columns too tall for one LDS group");
    }
}

// Spinor step: the five passes of the scalar step, every one on row pairs / column groups that
// hold both components (the plain transforms of the K-outer order included).
template <typename T, int RES>
static int split_spinor(const pgpe_split_args* a, hipStream_t s) {
    const pgpe_fft_plan* plan = (const pgpe_fft_plan*)a->plan;
    SplitParams<T> p = split_params<T>(a);
    c2<T>* psi = (c2<T>*)a->psi;
    c2<T>* tmp = (c2<T>*)a->tmp;
    int rc;
    p.in = psi;
    p.out = tmp;
    if (a->ordering == 0) rc = launch_rows_spinor<T, ROW_SS_FIRST, RES>(plan, p, s);
    else rc = launch_rows_spinor<T, ROW_FWD, 0>(plan, p, s);
    if (rc) return rc;
    p.in = tmp;
    p.out = tmp;
    if ((rc = launch_cols_spinor<T>(plan, p, s))) return rc;
    if (a->ordering == 0) rc = launch_rows_spinor<T, ROW_SS_MID, RES>(plan, p, s);
    else rc = launch_rows_spinor<T, ROW_KO_MID, RES>(plan, p, s);
    if (rc) return rc;
    if ((rc = launch_cols_spinor<T>(plan, p, s))) return rc;
    p.out = psi;
    if (a->ordering == 0) return launch_rows_spinor<T, ROW_SS_LAST, RES>(plan, p, s);
    return launch_rows_spinor<T, ROW_INV, 0>(plan, p, s);
}

int split_spinor_step(const pgpe_split_args* a, hipStream_t s) {
    const bool dbl = a->precision == 1;
    switch (a->reservoir) {
        case PGPE_RES_NONE: return dbl ? split_spinor<double, 0>(a, s) : split_spinor<float, 0>(a, s);
        case PGPE_RES_SINGLE: return dbl ? split_spinor<double, 1>(a, s) : split_spinor<float, 1>(a, s);
        case PGPE_RES_DOUBLE: return dbl ? split_spinor<double, 2>(a, s) : split_spinor<float, 2>(a, s);
        case PGPE_RES_QDOUBLE: return dbl ? split_spinor<double, 3>(a, s) : split_spinor<float, 3>(a, s);
        default: return fail(PGPE_ERR_ARG, "pgpe_split_step: unknown reservoir kind");
    }
}

}  // namespace pgpe
