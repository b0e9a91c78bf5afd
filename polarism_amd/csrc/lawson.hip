// Lawson integrating-factor Runge-Kutta (interaction picture) stages on gfx950.
//
// Restates solver/ifrk_fft_cuda_solver.py:96-293 (generic Lawson step and the
// embedded error) with the transforms done by the LDS Stockham kernels and
// the pointwise work fused into them.  For stage i of an s-stage tableau:
//
//   column kernel  psi_i^k = L(c_i h) psi0^k + sum_j (h a_ij) L((c_i - c_j) h) f_j
//                  (and -k^2 psi_i^k when kinetic relaxation is on), inverse along y
//   row kernel     inverse along x -> psi_i; n_i = n0 + sum_j (h a_ij) dn_j;
//                  N(psi_i, n_i) and dn_i; forward along x of N
//   column kernel  forward along y -> f_i
//
// L(tau) is exp(i (-hbar/2m) tau k^2) for a scalar field and the closed-form
// 2x2 exponential exp(-i H(k) tau / hbar) of the TE-TM / Zeeman dispersion for
// a spinor (dispersion.py:34-83).  The final kernels form psi_new^k (and the
// embedded error fields b - b_hat, b - b_hat3 for adaptive steps), transform
// back, clamp the reservoir, add the Wigner increment of fixed steps and
// reduce the scaled error per member in double precision.
#include <cmath>

#include "dispersion.hpp"
#include "fft_core.hpp"
#include "fft_plan.hpp"

namespace pgpe {

constexpr int kMaxStages = 12;

// LDS of a group: `seqs` padded sequences of length n plus the staged twiddle table.
template <typename T>
static size_t lds_bytes(int seqs, int n) {
    return (size_t)(seqs * seq_stride<T>(n) + tw_slots<T>(n)) * sizeof(c2<T>);
}

enum LawsonColMode { LC_FWD = 0, LC_COMBINE = 1, LC_FINAL = 2 };
enum LawsonRowMode { LR_RHS = 0, LR_RHS_REAL = 1, LR_FINAL = 2 };

template <typename T>
struct LawsonParams {
    const pgpe_member* table;
    FftDims dims;
    const c2<T>* tw;
    const double* kx;
    const double* ky;
    int nx, ny, members, comps, res_comps, nres;
    T scale;
    double h;
    // k-space combination
    const c2<T>* psi0k;
    const c2<T>* f[kMaxStages];
    double hc[kMaxStages];      // h * a_ij (stage) or h * b_j (final)
    double fr[kMaxStages];      // c_i - c_j (stage) or 1 - c_j (final)
    int nterms;
    double frac0;               // c_i (stage) or 1 (final)
    int lap;                    // produce -k^2 psi^k as an extra sequence
    int nerr;                   // final: number of error sequences (0, 1, 2)
    double he[2][kMaxStages];   // h * (b_j - b_hat_j), h * (b_j - b_hat3_j)
    c2<T>* psik_out;            // final: psi_new^k, or null
    // buffers (main sequence in/out; aux holds lap or error sequences)
    const c2<T>* in;
    c2<T>* out;
    c2<T>* aux;
    // stage physics
    const c2<T>* psi_real;
    const T* res0[2];
    const T* dres[kMaxStages][2];
    double rc[kMaxStages];      // reservoir stage (or final) coefficients, h * a_ij / h * b_j
    int nr;
    T* dres_out[2];
    const T* pump;
    const T* coef;
    const c2<T>* potential;
    // final
    c2<T>* psi_out;
    T* res_out[2];
    const c2<T>* psi_start;
    double atol, rtol;
    double* partials;           // 2 per row workgroup
    int noise;
    uint64_t step;
    const uint64_t* step_base;
    // device-driven adaptive trials: the trial's h is read from h_dev (the h-dependent
    // coefficients above then hold the bare tableau weights), and every kernel of the trial
    // returns at once while *skip is nonzero (the controller has reached its target)
    const double* h_dev;
    const int32_t* skip;
};

// The trial's step size, and the factor that turns the coefficient arrays (hc, he, rc)
// into h-weighted ones: 1 when the host folded h in already.  The products are the same
// IEEE operations as the host's h * a_ij, so both modes give identical coefficients.
template <typename T>
__device__ __forceinline__ double trial_h(const LawsonParams<T>& p) { return p.h_dev ? *p.h_dev : p.h; }
template <typename T>
__device__ __forceinline__ double coef_h(const LawsonParams<T>& p) { return p.h_dev ? *p.h_dev : 1.0; }
template <typename T>
__device__ __forceinline__ bool trial_skipped(const LawsonParams<T>& p) { return p.skip && *p.skip; }

// Linear propagator applied to one k point.  For the scalar field the angle
// follows ifrk_fft_cuda_solver.py:108 ((-hbar/2m) * (h f)) * k^2 in T; for the
// spinor the 2x2 closed form of dispersion.py:66-83 (spinor_propagate).
template <typename T>
__device__ __forceinline__ void apply_linear(const pgpe_member& mb, int comps, double kxv, double kyv, double tau,
                                             const c2<T>* v, c2<T>* out) {
    // Both branches give values and the stores follow the merge: a store to out[0] in one
    // branch and to out[1] in the other would be merged into one store through a selected
    // pointer, which keeps the caller's output array in scratch memory.
    c2<T> r0, r1 = make_c<T>((T)0, (T)0);
    if (comps == 1) {
        const T rate = (T)(-mb.hbar / (2.0 * mb.m_eff));
        const T k2 = (T)(kxv * kxv + kyv * kyv);
        const T angle = (rate * (T)tau) * k2;
        T s, c;
        sin_cos(angle, &s, &c);
        r0 = cmul<T>(make_c<T>(c, s), v[0]);
    } else {
        c2<T> o[2];
        spinor_propagate<T>(mb, kxv, kyv, tau, v, o);
        r0 = o[0];
        r1 = o[1];
    }
    out[0] = r0;
    out[1] = r1;
}

template <typename T, int RES>
__device__ __forceinline__ void lawson_res_rhs(const pgpe_member& mb, T rho, T P, const T n[2], T d[2]) {
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

// ------------------------------------------------------------------ columns
// Sequences per column group in LDS: kind a (0 main, 1.. aux) x component c x column i.
template <typename T, int V, int MODE>
__global__ void __launch_bounds__(256) lawson_cols_kernel(const LawsonParams<T> p) {
    if (trial_skipped(p)) return;
    const double hh = trial_h(p), hm = coef_h(p);
    extern __shared__ unsigned char smem_raw[];
    c2<T>* buf = reinterpret_cast<c2<T>*>(smem_raw);
    constexpr int E = 256 * V;
    const int n = p.ny;
    const int ld = seq_stride<T>(n);
    const int kinds = 1 + (MODE == LC_COMBINE ? p.lap : (MODE == LC_FINAL ? p.nerr : 0));
    const int seqs_per_col = kinds * p.comps;
    const int cols_per = E / (n * seqs_per_col);
    c2<T>* s_tw = buf + cols_per * seqs_per_col * ld;  // twiddles staged once per group
    const int groups = (p.nx + cols_per - 1) / cols_per;
    // XCD-aware: the groups one XCD runs side by side are neighbours along x, so the
    // narrow row segments they read share cache lines through the same L2
    const int gi = xcd_tile(blockIdx.x, groups * p.members);
    if (gi < 0) return;
    const int m = gi / groups;
    const int x0 = (gi - m * groups) * cols_per;
    const int cnt = min(cols_per, p.nx - x0);
    const int64_t plane = (int64_t)p.nx * n;
    const int64_t planes_total = (int64_t)p.members * p.comps;
    const pgpe_member& mb = p.table[m];
    auto seq = [&](int a, int c, int i) { return ((a * p.comps + c) * cnt + i) * ld; };
    stage_twiddles<T>(s_tw, p.tw, n);
    constexpr int CH = stage_chunk<T, V>();
    struct CPt { c2<T> v; int at; };

    if constexpr (MODE == LC_FWD) {
        for (int c = 0; c < p.comps; ++c) {
            const c2<T>* src = p.in + (m * p.comps + c) * plane;
            staged<CPt, CH>(
                cnt * n,
                [&](int e) {
                    const int y = e / cnt, i = e - y * cnt;
                    return CPt{src[(int64_t)y * p.nx + x0 + i], seq(0, c, i) + pad_at(y)};
                },
                [&](int, CPt q) { buf[q.at] = q.v; });
        }
        __syncthreads();
        lds_fft<T, V>(buf, p.dims, ld, cnt * p.comps, s_tw, false);
        if (p.frac0 != 0.0) {
            // stage values are kept as g_i = L(-c_i h) FFT(N_i): every later combination then
            // needs one propagator instead of one per term
            __syncthreads();
            for (int e = threadIdx.x; e < cnt * n; e += 256) {
                const int y = e / cnt, i = e - y * cnt;
                c2<T> v[2], o[2];
                for (int c = 0; c < 2; ++c) if (c < p.comps) v[c] = buf[seq(0, c, i) + pad_at(y)];
                apply_linear<T>(mb, p.comps, p.kx[x0 + i], p.ky[y], -p.frac0 * hh, v, o);
                for (int c = 0; c < 2; ++c) if (c < p.comps) buf[seq(0, c, i) + pad_at(y)] = o[c];
            }
            __syncthreads();
        }
    } else {
        // build the combination for every (k point, component) of the group
        for (int e = threadIdx.x; e < cnt * n; e += 256) {
            const int y = e / cnt, i = e - y * cnt;
            const int64_t g = (int64_t)y * p.nx + x0 + i;
            const double kxv = p.kx[x0 + i], kyv = p.ky[y];
            // psi_i = L(c_i h) (psi0 + sum_j hc_j g_j) with g_j = L(-c_j h) f_j, i.e.
            // L(c_i h) psi0 + sum_j hc_j L((c_i - c_j) h) f_j (ifrk_fft_cuda_solver.py:171-173);
            // the error fields of the final step (DP5: one, DOP853: two) likewise
            c2<T> sum[2];
            for (int c = 0; c < 2; ++c) if (c < p.comps) sum[c] = p.psi0k[(m * p.comps + c) * plane + g];
            constexpr int QE = MODE == LC_FINAL ? 2 : 0;
            c2<T> er[QE > 0 ? QE : 1][2];
            for (int q = 0; q < QE; ++q) er[q][0] = er[q][1] = make_c<T>(0, 0);
            // the stage terms of the k point, gathered first so that their loads overlap
            c2<T> vt[kMaxStages][2];
            bool used[kMaxStages];
#pragma unroll
            for (int j = 0; j < kMaxStages; ++j) {
                used[j] = j < p.nterms && p.hc[j] != 0.0;
                for (int q = 0; q < QE; ++q) used[j] |= j < p.nterms && q < p.nerr && p.he[q][j] != 0.0;
                if (used[j])
                    for (int c = 0; c < 2; ++c) if (c < p.comps) vt[j][c] = p.f[j][(m * p.comps + c) * plane + g];
            }
#pragma unroll
            for (int j = 0; j < kMaxStages; ++j) {
                if (!used[j]) continue;
                if (p.hc[j] != 0.0) {
                    const T w = (T)(p.hc[j] * hm);
                    for (int c = 0; c < 2; ++c) if (c < p.comps) {
                        sum[c].x = sum[c].x + w * vt[j][c].x;
                        sum[c].y = sum[c].y + w * vt[j][c].y;
                    }
                }
                for (int q = 0; q < QE; ++q)
                    if (q < p.nerr && p.he[q][j] != 0.0) {
                        const T w = (T)(p.he[q][j] * hm);
                        for (int c = 0; c < 2; ++c) if (c < p.comps) {
                            er[q][c].x = er[q][c].x + w * vt[j][c].x;
                            er[q][c].y = er[q][c].y + w * vt[j][c].y;
                        }
                    }
            }
            c2<T> acc[2];
            apply_linear<T>(mb, p.comps, kxv, kyv, p.frac0 * hh, sum, acc);
            for (int q = 0; q < QE; ++q)
                if (q < p.nerr) {
                    c2<T> o[2] = {make_c<T>(0, 0), make_c<T>(0, 0)};
                    apply_linear<T>(mb, p.comps, kxv, kyv, p.frac0 * hh, er[q], o);
                    er[q][0] = o[0];
                    er[q][1] = o[1];
                }
            for (int c = 0; c < 2; ++c) if (c < p.comps) {
                buf[seq(0, c, i) + pad_at(y)] = acc[c];
                if (MODE == LC_FINAL && p.psik_out) p.psik_out[(m * p.comps + c) * plane + g] = acc[c];
                if (MODE == LC_COMBINE && p.lap) {
                    const T mk2 = -(T)(kxv * kxv + kyv * kyv);
                    buf[seq(1, c, i) + pad_at(y)] = make_c<T>(mk2 * acc[c].x, mk2 * acc[c].y);
                }
            }
            if constexpr (MODE == LC_FINAL) {
                for (int q = 0; q < QE; ++q)
                    if (q < p.nerr)
                        for (int c = 0; c < 2; ++c) if (c < p.comps) buf[seq(1 + q, c, i) + pad_at(y)] = er[q][c];
            }
        }
        __syncthreads();
        lds_fft<T, V>(buf, p.dims, ld, cnt * seqs_per_col, s_tw, true);
    }
    // store: main -> out, aux kinds -> aux
    for (int a = 0; a < kinds; ++a) {
        for (int c = 0; c < p.comps; ++c) {
            c2<T>* dst = (a == 0 ? p.out : p.aux + (int64_t)(a - 1) * planes_total * plane) + (m * p.comps + c) * plane;
            staged<CPt, CH>(
                cnt * n,
                [&](int e) {
                    const int y = e / cnt, i = e - y * cnt;
                    return CPt{buf[seq(a, c, i) + pad_at(y)], y * p.nx + x0 + i};
                },
                [&](int, CPt q) { dst[q.at] = q.v; });
        }
    }
}

// ------------------------------------------------------------------ rows
template <typename T>
__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <typename T, int V, int MODE, int RES>
__global__ void __launch_bounds__(256) lawson_rows_kernel(const LawsonParams<T> p, int rows_per) {
    if (trial_skipped(p)) return;
    const double hm = coef_h(p);
    extern __shared__ unsigned char smem_raw[];
    c2<T>* buf = reinterpret_cast<c2<T>*>(smem_raw);
    constexpr int NF = RES == 0 ? 0 : (RES == 1 ? 1 : 2);
    const int n = p.nx;
    const int ld = seq_stride<T>(n);
    const FastDiv by_n((uint32_t)n);
    auto slot = [&](int e) {  // LDS slot of element e of a sequence group (row e / n, column e % n)
        const int r = by_n.div(e);
        return r * ld + pad_at(e - r * n);
    };
    const int kinds = 1 + (MODE == LR_RHS ? p.lap : (MODE == LR_FINAL ? p.nerr : 0));
    const int groups_per_member = p.ny / rows_per;
    const int m = blockIdx.x / groups_per_member;
    if (m >= p.members) return;
    const int y0 = (blockIdx.x - m * groups_per_member) * rows_per;
    const int64_t plane = (int64_t)p.nx * p.ny;
    const int64_t planes_total = (int64_t)p.members * p.comps;
    const int C = p.comps;
    const pgpe_member& mb = p.table[m];
    auto seq = [&](int a, int c, int r) { return ((a * C + c) * rows_per + r) * ld; };
    c2<T>* s_tw = buf + kinds * C * rows_per * ld;  // twiddles staged once per group
    const int elems = rows_per * n;
    stage_twiddles<T>(s_tw, p.tw, n);

    // ---- load ----
    for (int a = 0; a < kinds; ++a)
        for (int c = 0; c < C; ++c) {
            const c2<T>* src = MODE == LR_RHS_REAL ? p.psi_real : (a == 0 ? p.in : p.aux + (int64_t)(a - 1) * planes_total * plane);
            src += (m * C + c) * plane + (int64_t)y0 * n;
            c2<T>* dst = buf + seq(a, c, 0);
            staged<c2<T>, stage_chunk<T, V>()>(elems, [&](int e) { return src[e]; }, [&](int e, c2<T> v) { dst[slot(e)] = v; });
        }
    __syncthreads();
    if (MODE != LR_RHS_REAL) lds_fft<T, V>(buf, p.dims, ld, rows_per * C * kinds, s_tw, true);

    // ---- pointwise physics ----
    const T gC = (T)mb.g_C, gR = (T)mb.g_R, gI = (T)mb.g_I, Rr = (T)mb.R, gammaC = (T)mb.gamma_C;
    const T inv_hbar_neg = -(T)mb.inv_hbar;   // imaginary part of -1j/hbar
    const T a2 = (T)mb.alpha_2;
    const T eta = (T)mb.eta, hb2m = (T)mb.hbar_over_2m;
    const T coef = p.coef ? p.coef[m] : (T)0;
    const T* env = p.pump + mb.pump_off;
    const int nres_c = p.res_comps;
    uint64_t step = p.step;
    if (p.step_base) step += *p.step_base;
    double acc5 = 0.0, acc3 = 0.0;

    for (int e = threadIdx.x; e < elems; e += 256) {
        const int64_t cell = (int64_t)y0 * n + e;
        c2<T> psi[2], lap[2] = {make_c<T>(0, 0), make_c<T>(0, 0)};
        T rho[2] = {0, 0};
        for (int c = 0; c < 2; ++c) if (c < C) {
            c2<T> v = buf[seq(0, c, 0) + slot(e)];
            if (MODE != LR_RHS_REAL) {
                v.x *= p.scale;
                v.y *= p.scale;
            }
            psi[c] = v;
            rho[c] = v.x * v.x + v.y * v.y;
            if (MODE == LR_RHS && p.lap) {
                const c2<T> l = buf[seq(1, c, 0) + slot(e)];
                lap[c] = make_c<T>(l.x * p.scale, l.y * p.scale);
            }
        }
        // reservoir values of this stage (or final) for each reservoir component; the stage
        // derivatives of the cell are gathered first so that their loads overlap
        T nres[2][2] = {{0, 0}, {0, 0}};
        for (int rcix = 0; rcix < 2; ++rcix) if (rcix < nres_c) {
            for (int f = 0; f < NF; ++f) {
                const int64_t g = (int64_t)(m * nres_c + rcix) * plane + cell;
                T dv[kMaxStages];
#pragma unroll
                for (int j = 0; j < kMaxStages; ++j) {
                    bool used = j < p.nr && p.rc[j] != 0.0;
                    if constexpr (MODE == LR_FINAL)
                        for (int q = 0; q < 2; ++q) used |= j < p.nr && q < p.nerr && p.he[q][j] != 0.0;
                    dv[j] = used ? p.dres[j][f][g] : (T)0;
                }
                const T r0 = p.res0[f][g];
                T v = r0;
#pragma unroll
                for (int j = 0; j < kMaxStages; ++j)
                    if (j < p.nr && p.rc[j] != 0.0) v += (T)(p.rc[j] * hm) * dv[j];
                if constexpr (MODE == LR_FINAL) {
                    // clamp, store the new state; reservoir error sum_j (h w_j) dn_j against
                    // atol + rtol max(|n0|, |n1|)
                    v = v > (T)0 ? v : (T)0;
                    p.res_out[f][g] = v;
                    for (int q = 0; q < p.nerr; ++q) {
                        T ev = 0;
#pragma unroll
                        for (int j = 0; j < kMaxStages; ++j)
                            if (j < p.nr && p.he[q][j] != 0.0) ev += (T)(p.he[q][j] * hm) * dv[j];
                        const double sc = p.atol + p.rtol * (double)fmax(fabs(r0), fabs(v));
                        const double mag = fabs((double)ev);
                        const double s = sc > 0 ? mag / sc : (mag == 0 ? 0.0 : INFINITY);
                        if (q == 0) acc5 += s * s; else acc3 += s * s;
                    }
                }
                nres[rcix][f] = v;
            }
        }
        if constexpr (MODE == LR_FINAL) {
            // new field, its error terms, noise
            for (int c = 0; c < 2; ++c) if (c < C) {
                c2<T> v = psi[c];
                const int64_t g = (m * C + c) * plane + cell;
                if (p.nerr) {
                    const c2<T> st = p.psi_start[g];
                    const double sc = p.atol + p.rtol * fmax(hypot((double)st.x, (double)st.y),
                                                             hypot((double)v.x, (double)v.y));
                    for (int q = 0; q < p.nerr; ++q) {
                        const c2<T> er = buf[seq(1 + q, c, 0) + slot(e)];
                        const double mag = hypot((double)er.x * (double)p.scale, (double)er.y * (double)p.scale);
                        const double s = sc > 0 ? mag / sc : (mag == 0 ? 0.0 : INFINITY);
                        if (q == 0) acc5 += s * s; else acc3 += s * s;
                    }
                }
                if (p.noise) {
                    const T act = NF ? (nres_c == 2 && c == 1 ? nres[1][0] : nres[0][0]) : (T)0;
                    wigner_site<T>(v.x, v.y, mb.seed, (uint32_t)cell, step, 1u + (uint32_t)c, (T)mb.noise_strength,
                                   (T)mb.noise_gamma, (T)mb.noise_R, act, (T)mb.noise_dt_over_4v);
                }
                p.psi_out[g] = v;
            }
        } else {
            // stage right-hand sides (ifrk_fft_cuda_solver.py:128-142) and reservoir derivatives
            for (int c = 0; c < 2; ++c) if (c < C) {
                const T nA = NF ? (nres_c == 2 && c == 1 ? nres[1][0] : nres[0][0]) : (T)0;
                const T nI = NF > 1 ? (nres_c == 2 && c == 1 ? nres[1][1] : nres[0][1]) : (T)0;
                T er = gC * rho[c] + gR * nA + gI * nI, ei = 0;
                if (p.potential) {
                    const c2<T> pot = p.potential[cell];
                    er = (((pot.x + gC * rho[c]) + gR * nA) + gI * nI);
                    ei = pot.y;
                }
                if (C == 2) er = er + a2 * rho[1 - c];
                const T gain = (T)0.5 * (Rr * nA - gammaC);
                // (-i/hbar) eff = (inv*ei, -inv*er); times psi; plus gain psi
                const T ar = -(inv_hbar_neg * ei), ai = inv_hbar_neg * er;
                c2<T> d;
                d.x = (ar * psi[c].x - ai * psi[c].y) + gain * psi[c].x;
                d.y = (ar * psi[c].y + ai * psi[c].x) + gain * psi[c].y;
                if (p.lap) {
                    const T s = (eta * nA) * hb2m;
                    d.x = d.x + s * lap[c].x;
                    d.y = d.y + s * lap[c].y;
                }
                buf[seq(0, c, 0) + slot(e)] = d;
            }
            for (int rcix = 0; rcix < 2; ++rcix) if (rcix < nres_c) {
                T r = nres_c == 2 ? rho[rcix] : (C == 2 ? rho[0] + rho[1] : rho[0]);
                T P = env[cell] * coef;
                if (C == 2 && nres_c == 2) P = P * (T)(rcix == 0 ? mb.pump_w_plus : mb.pump_w_minus);
                T d[2] = {0, 0};
                lawson_res_rhs<T, RES>(mb, r, P, nres[rcix], d);
                for (int f = 0; f < NF; ++f) p.dres_out[f][(int64_t)(m * nres_c + rcix) * plane + cell] = d[f];
            }
        }
    }
    if constexpr (MODE == LR_FINAL) {
        __shared__ double s_acc[2][4];
        acc5 = wsum<T>(acc5);
        acc3 = wsum<T>(acc3);
        if ((threadIdx.x & 63) == 0) {
            s_acc[0][threadIdx.x >> 6] = acc5;
            s_acc[1][threadIdx.x >> 6] = acc3;
        }
        __syncthreads();
        if (threadIdx.x < 2 && p.partials)
            p.partials[2 * blockIdx.x + threadIdx.x] =
                ((s_acc[threadIdx.x][0] + s_acc[threadIdx.x][1]) + s_acc[threadIdx.x][2]) + s_acc[threadIdx.x][3];
        return;
    }
    __syncthreads();
    lds_fft<T, V>(buf, p.dims, ld, rows_per * C, s_tw, false);
    for (int c = 0; c < C; ++c) {
        c2<T>* dst = p.out + (m * C + c) * plane + (int64_t)y0 * n;
        const c2<T>* src = buf + seq(0, c, 0);
        staged<c2<T>, stage_chunk<T, V>()>(elems, [&](int e) { return src[slot(e)]; }, [&](int e, c2<T> v) { dst[e] = v; });
    }
}

// Per-member sums of the row partials, fixed order.
__global__ void lawson_member_error(const double* __restrict__ partials, int groups_per_member, int members,
                                    double* __restrict__ sums, const int32_t* skip) {
    if (skip && *skip) return;
    const int m = blockIdx.x;
    if (m >= members) return;
    double a5 = 0.0, a3 = 0.0;
    for (int g = threadIdx.x; g < groups_per_member; g += 64) {
        a5 += partials[2 * ((int64_t)m * groups_per_member + g)];
        a3 += partials[2 * ((int64_t)m * groups_per_member + g) + 1];
    }
    a5 = wsum<double>(a5);
    a3 = wsum<double>(a3);
    if (threadIdx.x == 0) {
        sums[2 * m] = a5;
        sums[2 * m + 1] = a3;
    }
}

// Rows per group: the largest power of two dividing ny whose seqs x rows sequences fit the
// group's 256 x V elements (one per thread slot), with their padded LDS copy and the staged
// twiddles within the 160 KB of LDS; 0 when none does.
template <typename T>
static int rows_per_group(int V, int nx, int ny, int seqs) {
    const int r = 256 * V / (nx * seqs);
    int best = 0;
    for (int d = 1; d <= r; d *= 2)
        if (ny % d == 0 && lds_bytes<T>(d * seqs, nx) <= 163840) best = d;
    return best;
}

template <typename T, int V, int MODE>
static int launch_cols_v(const pgpe_fft_plan* plan, LawsonParams<T> p, int seqs, hipStream_t s) {
    p.dims = plan->dims_y[radix_class<V>()];
    p.tw = (const c2<T>*)plan->tw_y[radix_class<V>()];
    const int cols_per = 256 * V / (p.ny * seqs);
    const size_t shm = lds_bytes<T>(cols_per * seqs, p.ny);
    auto kern = lawson_cols_kernel<T, V, MODE>;
    if (shm > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute(lawson_cols_kernel)");
    }
    const int groups = (p.nx + cols_per - 1) / cols_per;
    hipLaunchKernelGGL(kern, dim3(xcd_grid(groups * p.members)), dim3(256), shm, s, p);
    PGPE_LAUNCHED("lawson_cols_kernel");
    return PGPE_OK;
}

// A column group holds kinds x comps sequences per column; it takes twice the elements per
// thread (V = 32 or 24) when one column's sequences exceed 256 x V or fill it alone.
template <typename T, int MODE>
static int launch_cols(const pgpe_fft_plan* plan, LawsonParams<T> p, hipStream_t s) {
    const int kinds = 1 + (MODE == LC_COMBINE ? p.lap : (MODE == LC_FINAL ? p.nerr : 0));
    const int seqs = kinds * p.comps;
    const int need = plan->ny * seqs;
    const int V = plan->vy;
    auto fits = [&](int v) { return 256 * v >= need && lds_bytes<T>(256 * v / need * seqs, plan->ny) <= 163840; };
    // a group of a single column reads 8- or 16-byte row segments: twice the elements per
    // thread gives two columns (measured: the DP5 final combination on 1024^2 spinor fields
    // 159 -> 128 us); groups of two or more columns keep V (wider ones measured slower)
    const bool single = fits(V) && 256 * V / need < 2;
    if (fits(V) && !(single && fits(2 * V)))
        return V == 16 ? launch_cols_v<T, 16, MODE>(plan, p, seqs, s) : launch_cols_v<T, 12, MODE>(plan, p, seqs, s);
    if (fits(2 * V))
        return V == 16 ? launch_cols_v<T, 32, MODE>(plan, p, seqs, s) : launch_cols_v<T, 24, MODE>(plan, p, seqs, s);
    return fail(PGPE_ERR_UNSUPPORTED, "lawson: grid too tall for one LDS column group");
}

template <typename T, int V, int MODE, int RES>
static int launch_rows_v(const pgpe_fft_plan* plan, LawsonParams<T> p, int rows_per, int seqs, hipStream_t s) {
    p.dims = plan->dims_x[radix_class<V>()];
    p.tw = (const c2<T>*)plan->tw_x[radix_class<V>()];
    const size_t shm = lds_bytes<T>(rows_per * seqs, p.nx);
    auto kern = lawson_rows_kernel<T, V, MODE, RES>;
    if (shm > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute(lawson_rows_kernel)");
    }
    const int gpm = p.ny / rows_per;
    hipLaunchKernelGGL(kern, dim3(gpm * p.members), dim3(256), shm, s, p, rows_per);
    PGPE_LAUNCHED("lawson_rows_kernel");
    return PGPE_OK;
}

// Rows of a group never straddle two members (rows_per divides ny), so the
// error partials of a group belong to one member.  Long rows with several
// sequences per row take twice the elements per thread (V = 32 or 24).
template <typename T, int MODE, int RES>
static int launch_rows(const pgpe_fft_plan* plan, LawsonParams<T> p, hipStream_t s, int* groups_per_member) {
    const int kinds = 1 + (MODE == LR_RHS ? p.lap : (MODE == LR_FINAL ? p.nerr : 0));
    const int seqs = kinds * p.comps;
    const int V = plan->vx;
    int rows_per = rows_per_group<T>(V, plan->nx, plan->ny, seqs);
    if (rows_per >= 1) {
        if (groups_per_member) *groups_per_member = plan->ny / rows_per;
        return V == 16 ? launch_rows_v<T, 16, MODE, RES>(plan, p, rows_per, seqs, s)
                       : launch_rows_v<T, 12, MODE, RES>(plan, p, rows_per, seqs, s);
    }
    rows_per = rows_per_group<T>(2 * V, plan->nx, plan->ny, seqs);
    if (rows_per < 1) return fail(PGPE_ERR_UNSUPPORTED, "lawson: rows too long for one LDS row group");
    if (groups_per_member) *groups_per_member = plan->ny / rows_per;
    return V == 16 ? launch_rows_v<T, 32, MODE, RES>(plan, p, rows_per, seqs, s)
                   : launch_rows_v<T, 24, MODE, RES>(plan, p, rows_per, seqs, s);
}

template <typename T>
static LawsonParams<T> make_params(const pgpe_lawson_args* a) {
    const pgpe_fft_plan* plan = (const pgpe_fft_plan*)a->plan;
    LawsonParams<T> p{};
    p.table = a->table;
    p.kx = plan->kx;
    p.ky = plan->ky;
    p.nx = plan->nx;
    p.ny = plan->ny;
    p.members = a->members;
    p.comps = a->components;
    p.res_comps = a->res_components;
    p.scale = (T)(1.0 / ((double)plan->nx * (double)plan->ny));
    p.h = a->h;
    p.psi0k = (const c2<T>*)a->psi0k;
    p.nterms = a->nterms;
    for (int j = 0; j < a->nterms && j < kMaxStages; ++j) {
        p.f[j] = (const c2<T>*)a->f[j];
        p.hc[j] = a->hc[j];
        p.fr[j] = a->fr[j];
        p.he[0][j] = a->he5[j];
        p.he[1][j] = a->he3[j];
        p.rc[j] = a->rc[j];
        for (int f = 0; f < 2; ++f) p.dres[j][f] = (const T*)a->dres[j][f];
    }
    p.nr = a->nterms;
    p.frac0 = a->frac0;
    p.lap = a->lap;
    p.nerr = a->nerr;
    p.psik_out = (c2<T>*)a->psik_out;
    p.in = (const c2<T>*)a->work;
    p.out = (c2<T>*)a->work;
    p.aux = (c2<T>*)a->aux;
    p.psi_real = (const c2<T>*)a->psi_real;
    for (int f = 0; f < 2; ++f) {
        p.res0[f] = (const T*)a->res0[f];
        p.dres_out[f] = (T*)a->dres_out[f];
        p.res_out[f] = (T*)a->res_out[f];
    }
    p.pump = (const T*)a->pump;
    p.coef = (const T*)a->coef;
    p.potential = (const c2<T>*)a->potential;
    p.psi_out = (c2<T>*)a->psi_out;
    p.psi_start = (const c2<T>*)a->psi_start;
    p.atol = a->atol;
    p.rtol = a->rtol;
    p.partials = a->partials;
    p.noise = a->noise;
    p.step = a->step;
    p.step_base = a->step_base;
    p.h_dev = a->h_dev;
    p.skip = a->skip_dev;
    return p;
}

template <typename T, int RES>
static int lawson_dispatch(const pgpe_lawson_args* a, hipStream_t s) {
    const pgpe_fft_plan* plan = (const pgpe_fft_plan*)a->plan;
    LawsonParams<T> p = make_params<T>(a);
    int rc = PGPE_OK;
    switch (a->op) {
        case PGPE_LAWSON_STAGE:
            // combination + inverse (columns), inverse + RHS + forward (rows), forward (columns) -> f_out
            if (a->stage_real) {
                if ((rc = launch_rows<T, LR_RHS_REAL, RES>(plan, p, s, nullptr))) return rc;
            } else {
                if ((rc = launch_cols<T, LC_COMBINE>(plan, p, s))) return rc;
                if ((rc = launch_rows<T, LR_RHS, RES>(plan, p, s, nullptr))) return rc;
            }
            p.in = (const c2<T>*)a->work;
            p.out = (c2<T>*)a->f_out;
            return launch_cols<T, LC_FWD>(plan, p, s);
        case PGPE_LAWSON_FINAL: {
            if ((rc = launch_cols<T, LC_FINAL>(plan, p, s))) return rc;
            int gpm = 0;
            if ((rc = launch_rows<T, LR_FINAL, RES>(plan, p, s, &gpm))) return rc;
            if (a->nerr && a->error_sums) {
                hipLaunchKernelGGL(lawson_member_error, dim3(a->members), dim3(64), 0, s, a->partials, gpm, a->members,
                                   a->error_sums, a->skip_dev);
                PGPE_LAUNCHED("lawson_member_error");
            }
            return PGPE_OK;
        }
        default:
            return fail(PGPE_ERR_ARG, "pgpe_lawson: unknown op");
    }
}

}  // namespace pgpe

using namespace pgpe;

extern "C" int pgpe_lawson(const pgpe_lawson_args* a, void* stream) {
    if (!a || !a->plan || !a->table || !a->work) return fail(PGPE_ERR_ARG, "pgpe_lawson: null argument");
    const pgpe_fft_plan* plan = (const pgpe_fft_plan*)a->plan;
    if (plan->precision != a->precision) return fail(PGPE_ERR_ARG, "pgpe_lawson: plan precision mismatch");
    if (a->members <= 0 || (a->components != 1 && a->components != 2))
        return fail(PGPE_ERR_ARG, "pgpe_lawson: members > 0 and components 1 or 2 required");
    if (a->res_components != 1 && a->res_components != 2)
        return fail(PGPE_ERR_ARG, "pgpe_lawson: res_components must be 1 or 2");
    if (a->nterms < 0 || a->nterms > kMaxStages) return fail(PGPE_ERR_ARG, "pgpe_lawson: too many stages");
    if (a->nerr < 0 || a->nerr > 2) return fail(PGPE_ERR_ARG, "pgpe_lawson: nerr must be 0..2");
    if ((a->lap || a->nerr) && !a->aux) return fail(PGPE_ERR_ARG, "pgpe_lawson: aux buffer required");
    if (a->op == PGPE_LAWSON_STAGE && (!a->f_out || !a->pump || !a->coef))
        return fail(PGPE_ERR_ARG, "pgpe_lawson: stage needs f_out, pump and coef");
    if (a->op == PGPE_LAWSON_STAGE && !a->stage_real && !a->psi0k)
        return fail(PGPE_ERR_ARG, "pgpe_lawson: stage needs psi0k");
    if (a->op == PGPE_LAWSON_STAGE && a->stage_real && (!a->psi_real || a->lap))
        return fail(PGPE_ERR_ARG, "pgpe_lawson: real-space stage needs psi_real and no kinetic relaxation");
    if (a->op == PGPE_LAWSON_FINAL && (!a->psi_out || !a->psi0k)) return fail(PGPE_ERR_ARG, "pgpe_lawson: final needs psi_out");
    if (a->op == PGPE_LAWSON_FINAL && a->nerr && (!a->psi_start || !a->partials))
        return fail(PGPE_ERR_ARG, "pgpe_lawson: error estimate needs psi_start and partials");
    hipStream_t s = (hipStream_t)stream;
    const bool dbl = a->precision == 1;
    switch (a->reservoir) {
        case PGPE_RES_NONE: return dbl ? lawson_dispatch<double, 0>(a, s) : lawson_dispatch<float, 0>(a, s);
        case PGPE_RES_SINGLE: return dbl ? lawson_dispatch<double, 1>(a, s) : lawson_dispatch<float, 1>(a, s);
        case PGPE_RES_DOUBLE: return dbl ? lawson_dispatch<double, 2>(a, s) : lawson_dispatch<float, 2>(a, s);
        case PGPE_RES_QDOUBLE: return dbl ? lawson_dispatch<double, 3>(a, s) : lawson_dispatch<float, 3>(a, s);
        default: return fail(PGPE_ERR_ARG, "pgpe_lawson: unknown reservoir kind");
    }
}
