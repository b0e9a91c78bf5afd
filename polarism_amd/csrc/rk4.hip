// Fused RK4-FDM stage kernels (scalar field, periodic 2-D grid).
//
// One launch evaluates one RK4 stage of the stage-coupled condensate +
// reservoir system for every member of a packed member table
// (rk4_fdm_solver.py:36-77 with Model.rhs model.py:89-127):
//
//   stage input   psi_s = psi0 + a_s dt k_{s-1},  n_s = n0 + a_s dt kn_{s-1}
//   condensate    k = (-i/hbar)[-hbar^2/2m lap psi_s + (gC|psi_s|^2 + gR nA + gI nI) psi_s]
//                     + (R nA - gammaC)/2 psi_s + eta nA hbar/2m lap psi_s
//   reservoir     kn = reservoir RHS (single / double / quadratic-double)
//   final stage   psi1 = psi0 + dt/6 (k1 + 2k2 + 2k3 + k4), n1 = max(n0 + ..., 0)
//                 then the Wigner increment of the step (noise/stochastic.py:224-253).
//
// Layout: a workgroup of 256 threads (4 wavefronts) owns a 64 x TH tile of
// one member.  The stage input psi_s of the tile plus a one-point halo is
// formed once in LDS from coalesced loads of psi0 and k_{s-1}, so HBM sees each
// array once per stage; the 5-point Laplacian then reads LDS.  Wavefront w
// handles rows w, w+4, ... so every global access of a wavefront is one
// contiguous 64-point row segment.  The right-hand side folds the kinetic
// term into stencil weights on neighbour sums (scalar_rhs), which agrees with the
// reference's operation order to rounding; the library is built with
// -ffp-contract=off, and the RK4 point arithmetic (stencil sums, right-hand side,
// stage combinations) opts into fused multiply-adds per source expression (clang
// fp contract(on)), so the stage and whole-step kernels fuse the same operations.
#include <algorithm>

#include "common.hpp"

namespace pgpe {

template <typename T>
struct Rk4Params {
    const pgpe_member* table;
    const int32_t* tile_map;
    int num_tiles;
    int noise;
    T a;          // stage-input scale: 0, dt/2, dt/2, dt
    T dt6;        // dt / 6
    const c2<T>* psi0;
    const c2<T>* kprev;
    const c2<T>* k1;
    const c2<T>* k2;
    c2<T>* kout;
    c2<T>* psi_out;
    const T* res0[2];
    const T* kres_prev[2];
    const T* kres1[2];
    const T* kres2[2];
    T* kres_out[2];
    T* res_out[2];
    const T* pump;
    const T* coef;
    uint64_t step;
    const uint64_t* step_base;
    const c2<T>* potential;   // complex V packed per member at pot_off, or null
    const int32_t* active_tiles;  // device count of tile_map entries to run, or null (num_tiles)
};

// Reservoir right-hand side with the condensate frozen.  Rate sums are
// pre-combined on the host in double, as the reference adds Python floats
// before touching the array (double_reservoir.py:268-269).
template <typename T, int RES>
__device__ __forceinline__ void reservoir_rhs(const pgpe_member& mb, T rho, T P, const T n[2], T d[2]) {
#pragma clang fp contract(on)
    if constexpr (RES == PGPE_RES_SINGLE) {
        // dnR = P - (gamma_R + R rho) nR
        d[0] = P - ((T)mb.res_a + (T)mb.R * rho) * n[0];
    } else if constexpr (RES == PGPE_RES_DOUBLE) {
        // state (nA, nI); res_a = gI + R_IA, res_b = R_AI, res_c = R_IA, res_d = gA + R_AI
        const T nA = n[0], nI = n[1];
        d[1] = (P - (T)mb.res_a * nI) + (T)mb.res_b * nA;
        d[0] = (T)mb.res_c * nI - ((T)mb.res_d + (T)mb.R * rho) * nA;
    } else if constexpr (RES == PGPE_RES_QDOUBLE) {
        // state (nR, nI); res_a = kappa, res_b = gamma_R, res_c = gamma_I
        const T nR = n[0], nI = n[1];
        const T transfer = (T)mb.res_a * nI * nI;
        d[0] = (transfer - (T)mb.res_b * nR) - (T)mb.R * nR * rho;
        d[1] = (P - transfer) - (T)mb.res_c * nI;
    }
}

// Streaming accesses of the stage kernels: arrays each point touches once per
// launch (reservoir terms, pump, outputs).  They carry the non-temporal hint so
// they do not displace the psi / k rows that neighbouring tiles re-read as halo
// (measured on the 8192^2 FP32 middle stage: 503 -> 473 us).  Building with
// PGPE_TEMPORAL_STREAMS drops the hint for comparison.
template <typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
#ifndef PGPE_TEMPORAL_STREAMS
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}
template <typename T>
__device__ __forceinline__ void st_stream(T* p, T v) {
#ifndef PGPE_TEMPORAL_STREAMS
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}
template <typename T>
__device__ __forceinline__ c2<T> ld_stream_c(const c2<T>* p) {
#ifndef PGPE_TEMPORAL_STREAMS
    return make_c<T>(__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y));
#else
    return *p;
#endif
}
template <typename T>
__device__ __forceinline__ void st_stream_c(c2<T>* p, const c2<T>& v) {
#ifndef PGPE_TEMPORAL_STREAMS
    __builtin_nontemporal_store(v.x, &p->x);
    __builtin_nontemporal_store(v.y, &p->y);
#else
    *p = v;
#endif
}

template <int RES> struct ResFields { static constexpr int n = (RES == 0) ? 0 : (RES == 1 ? 1 : 2); };

// A value known to be equal in every lane, read from the first active lane into scalar registers.
__device__ __forceinline__ float uniform_value(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double uniform_value(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Member constants of the scalar right-hand side in the working precision, with the kinetic
// term folded into stencil weights.  From the 5-point neighbour sums sx = psi_e + psi_w and
// sy = psi_n + psi_s,
//   (1/hbar)(-hbar^2/2m) lap psi = cy sy + cx sx - c0 psi,
//   cx = kin / (hbar dx^2), cy = kin / (hbar dy^2), c0 = 2 (cx + cy)   (kin = -hbar^2/2m);
// from the isotropic 9-point sums s4 = sy + sx and sd (the four diagonals, dx = dy = h),
//   c4 s4 + c1 sd - c20 psi,  c1 = kin / (6 hbar h^2), c4 = 4 c1, c20 = 20 c1.
// The couplings enter as g / hbar and the gain as R/2 nA - gammaC/2.  Every folded constant is
// formed in FP64 from the member row and rounded once.
template <typename T>
struct ScalarConsts {
    T cx, cy, c0, c1, c4, c20;
    T gCh, gRh, gIh, hR, hG, inv_hbar;
    T eta_h;      // eta hbar / 2m
    T lap_scale;  // hbar / kin: lap psi = (K - centre psi) hbar / kin
    __device__ __forceinline__ explicit ScalarConsts(const pgpe_member& mb) {
        const double kh = mb.kin * mb.inv_hbar;
        const double k9 = kh * mb.inv_dx2 / 6.0;
        cx = (T)(kh * mb.inv_dx2);
        cy = (T)(kh * mb.inv_dy2);
        c0 = (T)(2.0 * kh * (mb.inv_dx2 + mb.inv_dy2));
        c1 = (T)k9;
        c4 = (T)(4.0 * k9);
        c20 = (T)(20.0 * k9);
        gCh = (T)(mb.g_C * mb.inv_hbar);
        gRh = (T)(mb.g_R * mb.inv_hbar);
        gIh = (T)(mb.g_I * mb.inv_hbar);
        hR = (T)(0.5 * mb.R);
        hG = (T)(0.5 * mb.gamma_C);
        inv_hbar = (T)mb.inv_hbar;
        eta_h = (T)(mb.eta * mb.hbar_over_2m);
        lap_scale = mb.eta != 0.0 ? (T)(1.0 / kh) : (T)0;  // used by the eta term only
    }
    // The constants are converted from the member row's doubles by vector instructions, so they
    // land in vector registers; a kernel short of them moves the (uniform) values to scalar ones.
    __device__ __forceinline__ void to_scalar_registers() {
        T* f[] = {&cx, &cy, &c0, &c1, &c4, &c20, &gCh, &gRh, &gIh, &hR, &hG, &inv_hbar, &eta_h, &lap_scale};
        for (T* v : f) *v = uniform_value(*v);
    }
};

// Neighbour index along one axis: periodic wrap, or the closed interval's
// mirror psi[-k] = psi[k], psi[n-1+k] = psi[n-1-k] (model.py:227-248), which
// for the 5-point stencil gives the reference's 2 (psi[1] - psi[0]) / h^2 edge
// rows and stays exact for deeper halos (the mirrored field evolves as the
// mirror image).
__device__ __forceinline__ int map_index(int g, int n, int boundary) {
    if (boundary == PGPE_BC_CLOSED) {
        if (n == 1) return 0;
        const int period = 2 * (n - 1);
        int r = g % period;
        r += r < 0 ? period : 0;
        return r < n ? r : period - r;
    }
    if (g >= 0 && g < n) return g;
    if (g < 0 && g >= -n) return g + n;
    int r = g % n;
    return r < 0 ? r + n : r;
}

// Kinetic term (1/hbar)(-hbar^2/2m) lap psi without its centre part, from the 5-point sums
// sx = psi_e + psi_w, sy = psi_n + psi_s (the stencil of model.py:220-224, periodic or mirrored):
// K = cy sy + cx sx.
template <typename T>
__device__ __forceinline__ c2<T> kin5(const ScalarConsts<T>& k, const c2<T>& sx, const c2<T>& sy) {
#pragma clang fp contract(on)
    return make_c<T>(k.cy * sy.x + k.cx * sx.x, k.cy * sy.y + k.cx * sx.y);
}

// The same for the isotropic 9-point stencil of rk4_cuda_solver.py:760-770,
// (4 (n + s + e + w) + sw + se + nw + ne - 20 psi) / (6 h^2): K = c4 s4 + c1 sd with
// s4 = (n + s) + (e + w) and sd = ((sw + se) + nw) + ne, diagonals in the order
// (y-1, x-1), (y-1, x+1), (y+1, x-1), (y+1, x+1).
template <typename T>
__device__ __forceinline__ c2<T> kin9(const ScalarConsts<T>& k, const c2<T>& s4, const c2<T>& sd) {
#pragma clang fp contract(on)
    return make_c<T>(k.c4 * s4.x + k.c1 * sd.x, k.c4 * s4.y + k.c1 * sd.y);
}

template <typename T>
__device__ __forceinline__ c2<T> csum(const c2<T>& a, const c2<T>& b) {
    return make_c<T>(a.x + b.x, a.y + b.y);
}

// Scalar condensate and reservoir derivatives at one point (model.py:89-127):
//   k = (-i/hbar)[-hbar^2/2m lap psi + (V + gC|psi|^2 + gR nA + gI nI) psi]
//       + (R nA - gammaC)/2 psi + eta nA hbar/2m lap psi,
// from the stage input ps, the kinetic sums K (kin5 / kin9) with their centre weight
// (c0 / c20), the stage reservoir n and a complex potential V (used when has_v).  With
// E = (gC rho + gR nA + gI nI + Re V) / hbar - centre and G = R nA / 2 - gammaC / 2 + Im V / hbar,
//   k = (G psi_r + (E psi_i + K_i), G psi_i - (E psi_r + K_r)).
// The per-stage and the whole-step kernels all call this, so they perform the same operations in
// the same order.
template <typename T, int RES, bool ETA = true>
__device__ __forceinline__ void scalar_rhs(const pgpe_member& mb, const ScalarConsts<T>& k, const c2<T>& ps,
                                           const c2<T>& K, T centre, const T n[2], T P, bool has_v, c2<T> V,
                                           c2<T>& d, T dn[2]) {
#pragma clang fp contract(on)
    constexpr int NF = ResFields<RES>::n;
    const T rho = ps.x * ps.x + ps.y * ps.y;
    dn[0] = (T)0;
    dn[1] = (T)0;
    reservoir_rhs<T, RES>(mb, rho, P, n, dn);
    const T nA = (NF > 0) ? n[0] : (T)0;
    const T nI = (NF > 1) ? n[1] : (T)0;
    // terms of reservoir fields the model does not have are left out
    T E = NF > 1 ? k.gCh * rho + (k.gRh * nA + (k.gIh * nI - centre))
                 : (NF > 0 ? k.gCh * rho + (k.gRh * nA - centre) : k.gCh * rho - centre);
    T G = NF > 0 ? k.hR * nA - k.hG : -k.hG;
    if (has_v) {
        E = E + V.x * k.inv_hbar;
        G = G + V.y * k.inv_hbar;
    }
    d.x = G * ps.x + (E * ps.y + K.y);
    d.y = G * ps.y - (E * ps.x + K.x);
    if (ETA && k.eta_h != (T)0) {
        const T s = k.eta_h * nA;
        d.x = d.x + s * ((K.x - centre * ps.x) * k.lap_scale);
        d.y = d.y + s * ((K.y - centre * ps.y) * k.lap_scale);
    }
}

// KIND: 0 = first stage (input is psi0), 1 = middle stage, 2 = final stage.
template <typename T, int RES, int TH, int KIND>
__global__ void __launch_bounds__(256) rk4_stage_kernel(const Rk4Params<T> p) {
#pragma clang fp contract(on)
    constexpr int NF = ResFields<RES>::n;
    constexpr int W = kTileW + 2;
    __shared__ c2<T> s_psi[(TH + 2) * W];

    const int tile = xcd_tile(blockIdx.x, p.active_tiles ? *p.active_tiles : p.num_tiles);
    if (tile < 0) return;
    const int m = p.tile_map[2 * tile];
    // a copy of the member row: its fields stay in registers past the global stores
    const pgpe_member mb = p.table[m];
    const int nx = mb.nx, ny = mb.ny;
    const int local = p.tile_map[2 * tile + 1];
    const int x0 = (local % mb.tiles_x) * kTileW;
    const int y0 = (local / mb.tiles_x) * TH;
    const int64_t off = mb.off;
    const c2<T>* __restrict__ psi0 = p.psi0 + off;

    // ---- stage input with halo into LDS (periodic wrap or mirrored edges) ----
    // Wavefront w loads the interior rows w, w + 4, ... it later computes (one point per
    // lane, one coalesced 64-point segment per row) and keeps psi0 and k_prev of those rows
    // in registers for the final combination; waves 0 / 1 load the halo rows above / below
    // (lanes 0 and 1 also the corners), waves 2 / 3 the halo columns left / right.  The
    // reservoir and pump values of the wave's rows are loaded here too, so every stream's
    // loads are in flight before the first wait (rows past the member's edge load the
    // wrapped or mirrored point, in bounds and unused).
    constexpr int RPW = TH / 4;
    constexpr int NFA = NF > 0 ? NF : 1;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int bc = mb.boundary;
    const T* __restrict__ env = p.pump + mb.pump_off;
    c2<T> z[RPW], kp[RPW];
    T r0a[RPW][NFA], b3a[RPW][NFA], envq[RPW];
    {
        const int gxi = map_index(x0 + lane, nx, bc);
#pragma unroll
        for (int q = 0; q < RPW; ++q) {
            const int r = wave + 4 * q;
            const int64_t g = (int64_t)map_index(y0 + r, ny, bc) * nx + gxi;
            z[q] = psi0[g];
            c2<T> v = z[q];
            if constexpr (KIND != 0) {
                kp[q] = p.kprev[off + g];
                v.x = v.x + p.a * kp[q].x;
                v.y = v.y + p.a * kp[q].y;
            }
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                r0a[q][f] = ld_stream(&p.res0[f][off + g]);
                if constexpr (KIND != 0) b3a[q][f] = ld_stream(&p.kres_prev[f][off + g]);
            }
            envq[q] = ld_stream(&env[g]);
            s_psi[(r + 1) * W + lane + 1] = v;
        }
        auto halo = [&](int ly, int lx) {
            const int64_t g = (int64_t)map_index(y0 + ly, ny, bc) * nx + map_index(x0 + lx, nx, bc);
            c2<T> v = psi0[g];
            if constexpr (KIND != 0) {
                const c2<T> k = p.kprev[off + g];
                v.x = v.x + p.a * k.x;
                v.y = v.y + p.a * k.y;
            }
            s_psi[(ly + 1) * W + lx + 1] = v;
        };
        if (wave < 2) {
            const int ly = wave == 0 ? -1 : TH;
            halo(ly, lane);
            if (lane < 2) halo(ly, lane == 0 ? -1 : kTileW);
        } else if (lane < TH) {
            halo(lane, wave == 2 ? -1 : kTileW);
        }
    }
    __syncthreads();

    const int gx = x0 + lane;
    if (gx >= nx) return;
    const ScalarConsts<T> kc(mb);
    const bool nine = mb.laplacian == PGPE_LAP_NINE;
    const T centre = nine ? kc.c20 : kc.c0;
    const T coef = p.coef[m];
    uint64_t step = p.step;
    if (KIND == 2 && p.noise && p.step_base) step += *p.step_base;

#pragma unroll
    for (int q = 0; q < RPW; ++q) {
        const int r = wave + 4 * q;
        const int gy = y0 + r;
        if (gy >= ny) break;
        const int64_t idx = (int64_t)gy * nx + gx;
        const int c = (r + 1) * W + lane + 1;
        T n[2] = {(T)0, (T)0}, r0[2] = {(T)0, (T)0}, b3[2] = {(T)0, (T)0};
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            r0[f] = r0a[q][f];
            T v = r0[f];
            if constexpr (KIND != 0) {
                b3[f] = b3a[q][f];
                v = v + p.a * b3[f];
            }
            n[f] = v;
        }
        const T P = envq[q] * coef;
        const c2<T> ps = s_psi[c];
        const c2<T> sx = csum<T>(s_psi[c + 1], s_psi[c - 1]);
        const c2<T> sy = csum<T>(s_psi[c + W], s_psi[c - W]);
        const c2<T> K = nine ? kin9<T>(kc, csum<T>(sy, sx),
                                       csum<T>(csum<T>(csum<T>(s_psi[c - W - 1], s_psi[c - W + 1]), s_psi[c + W - 1]),
                                               s_psi[c + W + 1]))
                             : kin5<T>(kc, sx, sy);
        const bool has_v = p.potential != nullptr;
        const c2<T> vpot = has_v ? p.potential[mb.pot_off + idx] : make_c<T>((T)0, (T)0);
        c2<T> d;
        T dn[2];
        scalar_rhs<T, RES>(mb, kc, ps, K, centre, n, P, has_v, vpot, d, dn);

        if constexpr (KIND != 2) {
            st_stream_c<T>(&p.kout[off + idx], d);
#pragma unroll
            for (int f = 0; f < NF; ++f) st_stream(&p.kres_out[f][off + idx], dn[f]);
        } else {
            const c2<T> a1 = ld_stream_c<T>(&p.k1[off + idx]);
            const c2<T> a2 = ld_stream_c<T>(&p.k2[off + idx]);
            const c2<T> a3 = kp[q];
            c2<T> out;
            out.x = z[q].x + p.dt6 * (((a1.x + (T)2 * a2.x) + (T)2 * a3.x) + d.x);
            out.y = z[q].y + p.dt6 * (((a1.y + (T)2 * a2.y) + (T)2 * a3.y) + d.y);
            T n_new[2] = {(T)0, (T)0};
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                const T b1 = ld_stream(&p.kres1[f][off + idx]);
                const T b2 = ld_stream(&p.kres2[f][off + idx]);
                T v = r0[f] + p.dt6 * (((b1 + (T)2 * b2) + (T)2 * b3[f]) + dn[f]);
                v = v > (T)0 ? v : (T)0;  // set_state clamp (single_reservoir.py:70-73)
                n_new[f] = v;
                p.res_out[f][off + idx] = v;
            }
            if (p.noise) {
                wigner_site<T>(out.x, out.y, mb.seed, (uint32_t)idx, step, 1u, (T)mb.noise_strength,
                               (T)mb.noise_gamma, (T)mb.noise_R, NF > 0 ? n_new[0] : (T)0,
                               (T)mb.noise_dt_over_4v);
            }
            p.psi_out[off + idx] = out;
        }
    }
}

// ------------------------------------------------------------------ spinor
// Spinor derivatives at one point (model.py:129-208, the stencil of rk4_cuda_solver.py:1600-1637)
// from the stage input of both components in LDS (planes t0 = psi+, t1 = psi-, point ci, row
// stride W, the point's eight neighbours in place), the stage reservoir n (per component, or one
// shared in slot 0 when !two) and the pump P; d per component, dn per reservoir slot.  The
// spinor stage and whole-step kernels both call this, so they perform the same operations in the
// same order.
template <typename T, int RES>
__device__ __forceinline__ void spinor_rhs(const pgpe_member& mb, const c2<T>* t0, const c2<T>* t1, int ci, int W,
                                           bool two, const T (&n)[2][2], T P, c2<T> (&d)[2], T (&dn)[2][2]) {
    constexpr int NF = ResFields<RES>::n;
    const T inv_dx2 = (T)mb.inv_dx2, inv_dy2 = (T)mb.inv_dy2, inv4 = (T)mb.inv_4dxdy;
    const T kin = (T)mb.kin, inv_hbar = (T)mb.inv_hbar;
    const T gC = (T)mb.g_C, gR = (T)mb.g_R, gI = (T)mb.g_I, R = (T)mb.R, gammaC = (T)mb.gamma_C;
    const T a2 = (T)mb.alpha_2, zh = (T)mb.zeeman, beta = (T)mb.te_tm;
    const T eta_c = (T)mb.eta, hb2m = (T)mb.hbar_over_2m;
    c2<T> ps[2], lap[2], te[2];
    T rho[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const c2<T>* t = c == 0 ? t0 : t1;
        const c2<T> z = t[ci], nn = t[ci + W], so = t[ci - W], e = t[ci + 1], w = t[ci - 1];
        ps[c] = z;
        rho[c] = z.x * z.x + z.y * z.y;
        lap[c].x = ((nn.x - (T)2 * z.x) + so.x) * inv_dy2 + ((e.x - (T)2 * z.x) + w.x) * inv_dx2;
        lap[c].y = ((nn.y - (T)2 * z.y) + so.y) * inv_dy2 + ((e.y - (T)2 * z.y) + w.y) * inv_dx2;
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        te[c] = make_c<T>(0, 0);
        if (beta == (T)0) continue;
        const c2<T>* t = c == 0 ? t1 : t0;   // the other component
        const c2<T> z = t[ci], nn = t[ci + W], so = t[ci - W], e = t[ci + 1], w = t[ci - 1];
        const c2<T> ne = t[ci + W + 1], nw = t[ci + W - 1], se = t[ci - W + 1], sw = t[ci - W - 1];
        c2<T> dxx, dyy, dxy;
        dxx.x = ((e.x - (T)2 * z.x) + w.x) * inv_dx2;
        dxx.y = ((e.y - (T)2 * z.y) + w.y) * inv_dx2;
        dyy.x = ((nn.x - (T)2 * z.x) + so.x) * inv_dy2;
        dyy.y = ((nn.y - (T)2 * z.y) + so.y) * inv_dy2;
        dxy.x = (((ne.x - nw.x) - se.x) + sw.x) * inv4;
        dxy.y = (((ne.y - nw.y) - se.y) + sw.y) * inv4;
        // cross = -2i dxy (c = 0) or +2i dxy (c = 1)
        const c2<T> cross = c == 0 ? make_c<T>((T)2 * dxy.y, -(T)2 * dxy.x) : make_c<T>(-(T)2 * dxy.y, (T)2 * dxy.x);
        te[c].x = -beta * ((dxx.x - dyy.x) + cross.x);
        te[c].y = -beta * ((dxx.y - dyy.y) + cross.y);
    }
#pragma unroll
    for (int rc = 0; rc < 2; ++rc) {
        if (rc > 0 && !two) continue;
        const T rr = two ? rho[rc] : rho[0] + rho[1];
        const T Pc = two ? P * (T)(rc == 0 ? mb.pump_w_plus : mb.pump_w_minus) : P;
        reservoir_rhs<T, RES>(mb, rr, Pc, n[rc], dn[rc]);
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const T nA = NF > 0 ? (two ? n[c][0] : n[0][0]) : (T)0;
        const T nI = NF > 1 ? (two ? n[c][1] : n[0][1]) : (T)0;
        const T sgn = c == 0 ? (T)1 : (T)-1;
        const T eff = ((((gC * rho[c]) + a2 * rho[1 - c]) + zh * sgn / (T)2) + gR * nA) + gI * nI;
        const T gain = (R * nA - gammaC) * (T)0.5;
        const T Xr = (kin * lap[c].x + eff * ps[c].x) + te[c].x;
        const T Xi = (kin * lap[c].y + eff * ps[c].y) + te[c].y;
        d[c].x = inv_hbar * Xi + gain * ps[c].x;
        d[c].y = -(inv_hbar * Xr) + gain * ps[c].y;
        if (eta_c != (T)0) {
            const T s = (eta_c * nA) * hb2m;
            d[c].x = d[c].x + s * lap[c].x;
            d[c].y = d[c].y + s * lap[c].y;
        }
    }
}

// Circular-basis spinor stage (model.py:129-208): per component c, with o the
// other component,
//   X_c = -hbar^2/2m lap psi_c + (gC rho_c + a2 rho_o +- Z/2 + gR nA + gI nI) psi_c
//         - beta [(dxx - dyy) o -/+ 2i dxy o]
//   k_c = (-i/hbar) X_c + (R nA - gammaC)/2 psi_c + eta nA hbar/2m lap psi_c
// dxy is the 4-corner central difference, so the LDS tile keeps its corners.  Periodic grids
// wrap the halo; closed intervals mirror it about the wall nodes, which gives the Laplacian the
// edge rows 2 (psi[1] - psi[0]) / h^2 of model.py:227-248 and the TE-TM differences their
// reflect-padded values.
template <typename T, int RES, int TH, int KIND>
__global__ void __launch_bounds__(256) rk4_spinor_stage_kernel(const Rk4Params<T> p, int res_comps) {
    constexpr int NF = ResFields<RES>::n;
    constexpr int W = kTileW + 2;
    constexpr int PL = (TH + 2) * W;
    __shared__ c2<T> s_psi[2 * PL];

    const int tile = xcd_tile(blockIdx.x, p.active_tiles ? *p.active_tiles : p.num_tiles);
    if (tile < 0) return;
    const int m = p.tile_map[2 * tile];
    // a copy of the member row: its fields stay in registers past the global stores
    const pgpe_member mb = p.table[m];
    const int nx = mb.nx, ny = mb.ny;
    const int local = p.tile_map[2 * tile + 1];
    const int x0 = (local % mb.tiles_x) * kTileW;
    const int y0 = (local / mb.tiles_x) * TH;
    const int64_t plane = (int64_t)nx * ny;
    const int64_t poff = mb.off * 2;              // first psi plane of this member
    const int64_t roff = mb.off * res_comps;      // first reservoir plane

    for (int i = threadIdx.x; i < 2 * PL; i += 256) {
        const int c = i / PL;
        const int j = i - c * PL;
        const int ly = j / W;
        const int lx = j - ly * W;
        // periodic wrap, or the closed interval's mirror about the wall nodes (the one-point
        // reflect extension of rk4_cuda_solver.py:1563-1568 and model.py:196-205)
        const int gy = map_index(y0 + ly - 1, ny, mb.boundary);
        const int gx = map_index(x0 + lx - 1, nx, mb.boundary);
        const int64_t g = poff + c * plane + (int64_t)gy * nx + gx;
        c2<T> v = p.psi0[g];
        if constexpr (KIND != 0) {
            const c2<T> k = p.kprev[g];
            v.x = v.x + p.a * k.x;
            v.y = v.y + p.a * k.y;
        }
        s_psi[i] = v;
    }
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int gx = x0 + lane;
    if (gx >= nx) return;
    const T coef = p.coef[m];
    const T* __restrict__ env = p.pump + mb.pump_off;
    uint64_t step = p.step;
    if (KIND == 2 && p.noise && p.step_base) step += *p.step_base;
    const bool two = res_comps == 2;

    for (int r = threadIdx.x >> 6; r < TH; r += 4) {
        const int gy = y0 + r;
        if (gy >= ny) break;
        const int64_t cell = (int64_t)gy * nx + gx;
        const int ci = (r + 1) * W + lane + 1;
        // reservoir stage values; the component loops run over both slots with constant indices
        // (guarded by res_comps) so that these arrays stay in registers
        T n[2][2] = {{0, 0}, {0, 0}}, dn[2][2] = {{0, 0}, {0, 0}};
        const T P = env[cell] * coef;
#pragma unroll
        for (int rc = 0; rc < 2; ++rc) {
            if (rc >= res_comps) continue;
            const int64_t g = roff + rc * plane + cell;
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                T v = p.res0[f][g];
                if constexpr (KIND != 0) v = v + p.a * p.kres_prev[f][g];
                n[rc][f] = v;
            }
        }
        c2<T> d[2];
        spinor_rhs<T, RES>(mb, s_psi, s_psi + PL, ci, W, two, n, P, d, dn);
        if constexpr (KIND != 2) {
#pragma unroll
            for (int c = 0; c < 2; ++c) p.kout[poff + c * plane + cell] = d[c];
#pragma unroll
            for (int rc = 0; rc < 2; ++rc) {
                if (rc >= res_comps) continue;
#pragma unroll
                for (int f = 0; f < NF; ++f) p.kres_out[f][roff + rc * plane + cell] = dn[rc][f];
            }
        } else {
            T n_new[2][2] = {{0, 0}, {0, 0}};
#pragma unroll
            for (int rc = 0; rc < 2; ++rc) {
                if (rc >= res_comps) continue;
                const int64_t g = roff + rc * plane + cell;
#pragma unroll
                for (int f = 0; f < NF; ++f) {
                    const T r0 = p.res0[f][g];
                    T v = r0 + p.dt6 * (((p.kres1[f][g] + (T)2 * p.kres2[f][g]) + (T)2 * p.kres_prev[f][g]) + dn[rc][f]);
                    v = v > (T)0 ? v : (T)0;
                    n_new[rc][f] = v;
                    p.res_out[f][g] = v;
                }
            }
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const int64_t g = poff + c * plane + cell;
                const c2<T> a1 = p.k1[g], a2v = p.k2[g], a3 = p.kprev[g], z = p.psi0[g];
                c2<T> out;
                out.x = z.x + p.dt6 * (((a1.x + (T)2 * a2v.x) + (T)2 * a3.x) + d[c].x);
                out.y = z.y + p.dt6 * (((a1.y + (T)2 * a2v.y) + (T)2 * a3.y) + d[c].y);
                if (p.noise) {
                    const T act = NF > 0 ? (two ? n_new[c][0] : n_new[0][0]) : (T)0;
                    wigner_site<T>(out.x, out.y, mb.seed, (uint32_t)cell, step, 1u + (uint32_t)c,
                                   (T)mb.noise_strength, (T)mb.noise_gamma, (T)mb.noise_R, act,
                                   (T)mb.noise_dt_over_4v);
                }
                p.psi_out[g] = out;
            }
        }
    }
}

// ------------------------------------------------------------------ whole step
// All four stages of one RK4 step in one launch (temporal blocking).  A
// workgroup loads psi0, the reservoir and the pump of a 64 x TH output tile
// plus a 4-point halo (72 x (TH + 8) points, periodic wrap), keeps the
// per-point state in registers and only the stage input psi_s in LDS (double
// buffered), and recomputes the halo: after stage s the outer s rings are
// stale, so after four stages exactly the 64 x TH interior is exact.  HBM
// sees psi0, n0 and the pump once (times the halo overhead) and the new state
// once per step, instead of the 13 complex + 17 real arrays per point of the
// four stage launches.  Each point runs the same operations in the same order
// as the stage kernels (scalar_rhs; the final combination accumulates
// ((k1 + 2 k2) + 2 k3) + k4 left to right), so on periodic grids the two paths
// agree bit for bit.  On a closed interval the halo points beyond a wall are
// mirror images computed here rather than copied, and their stencil sums run
// in mirrored order, so the two paths agree to rounding there.
template <typename T>
struct Rk4StepParams {
    const pgpe_member* table;
    const int32_t* tile_map;
    int num_tiles;
    int noise;
    int members;
    T a_half, a_full, dt6;
    const c2<T>* psi0;
    c2<T>* psi_out;
    const T* res0[2];
    T* res_out[2];
    const T* pump;
    const T* coef;       // 4 x members: pump coefficient of stage s of member m at coef[s * members + m]
    uint64_t step;
    const uint64_t* step_base;
    const int32_t* active_tiles;  // device count of tile_map entries to run, or null (num_tiles)
    int res_comps;                // spinor members: reservoir planes per member (1 shared, 2 per component)
};

constexpr int kStepHalo = 4;

template <typename T, int RES, int TH>
__global__ void __launch_bounds__(256) rk4_step_kernel(const Rk4StepParams<T> p) {
#pragma clang fp contract(on)
    constexpr int NF = ResFields<RES>::n;
    constexpr int EW = kTileW + 2 * kStepHalo;
    constexpr int EH = TH + 2 * kStepHalo;
    constexpr int NP = EW * EH;
    constexpr int PPT = (NP + 255) / 256;
    // Each buffer has a one-row-plus-one pad on both sides, so every stencil
    // neighbour f +- EW +- 1 is an LDS address without clamping; the pads and
    // the outer ring hold stale values that only reach stale halo points.
    constexpr int PAD = EW + 1;
    extern __shared__ unsigned char smem_raw[];
    c2<T>* buf0 = reinterpret_cast<c2<T>*>(smem_raw) + PAD;
    c2<T>* buf1 = buf0 + NP + 2 * PAD;

    const int tile = xcd_tile(blockIdx.x, p.active_tiles ? *p.active_tiles : p.num_tiles);
    if (tile < 0) return;
    const int m = p.tile_map[2 * tile];
    // a copy of the member row: its fields stay in registers past the global stores
    const pgpe_member mb = p.table[m];
    const int nx = mb.nx, ny = mb.ny;
    const int local = p.tile_map[2 * tile + 1];
    const int x0 = (local % mb.tiles_x) * kTileW;
    const int y0 = (local / mb.tiles_x) * TH;
    const int64_t off = mb.off;
    const ScalarConsts<T> kc(mb);
    const bool nine = mb.laplacian == PGPE_LAP_NINE;
    const T centre = nine ? kc.c20 : kc.c0;
    T coef[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) coef[s] = p.coef[s * p.members + m];

    c2<T> z[PPT], acc[PPT];
    T env[PPT], n0[PPT][2], ns[PPT][2], accn[PPT][2];
    const int tid = threadIdx.x;
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const int f = tid + 256 * j;
        if (f < NP) {
            const int ly = f / EW;
            const int lx = f - ly * EW;
            const int gy = map_index(y0 + ly - kStepHalo, ny, mb.boundary);
            const int gx = map_index(x0 + lx - kStepHalo, nx, mb.boundary);
            const int g = gy * nx + gx;
            z[j] = p.psi0[off + g];
            env[j] = p.pump[mb.pump_off + g];
#pragma unroll
            for (int q = 0; q < NF; ++q) {
                n0[j][q] = p.res0[q][off + g];
                ns[j][q] = n0[j][q];
            }
            buf0[f] = z[j];
        }
    }
    __syncthreads();

#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const c2<T>* __restrict__ cur = (s & 1) ? buf1 : buf0;
        c2<T>* __restrict__ nxt = (s & 1) ? buf0 : buf1;
        const T a_next = s < 2 ? p.a_half : p.a_full;
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            const int f = tid + 256 * j;
            if (f >= NP) continue;
            // neighbours clamped into the buffer; border points are stale anyway
            const c2<T> ps = cur[f];
            const c2<T> sx = csum<T>(cur[f + 1], cur[f - 1]);
            const c2<T> sy = csum<T>(cur[f + EW], cur[f - EW]);
            c2<T> K;
            if (nine) {
                K = kin9<T>(kc, csum<T>(sy, sx),
                            csum<T>(csum<T>(csum<T>(cur[f - EW - 1], cur[f - EW + 1]), cur[f + EW - 1]), cur[f + EW + 1]));
            } else {
                K = kin5<T>(kc, sx, sy);
            }
            const T P = env[j] * coef[s];
            c2<T> d;
            T dn[2];
            scalar_rhs<T, RES>(mb, kc, ps, K, centre, ns[j], P, false, make_c<T>((T)0, (T)0), d, dn);
            if (s == 0) {
                acc[j] = d;
#pragma unroll
                for (int q = 0; q < NF; ++q) accn[j][q] = dn[q];
            } else if (s < 3) {
                acc[j].x = acc[j].x + (T)2 * d.x;
                acc[j].y = acc[j].y + (T)2 * d.y;
#pragma unroll
                for (int q = 0; q < NF; ++q) accn[j][q] = accn[j][q] + (T)2 * dn[q];
            }
            if (s < 3) {
                c2<T> v;
                v.x = z[j].x + a_next * d.x;
                v.y = z[j].y + a_next * d.y;
                nxt[f] = v;
#pragma unroll
                for (int q = 0; q < NF; ++q) ns[j][q] = n0[j][q] + a_next * dn[q];
            } else {
                const int ly = f / EW;
                const int lx = f - ly * EW;
                const int oy = y0 + ly - kStepHalo, ox = x0 + lx - kStepHalo;
                if (lx < kStepHalo || lx >= kStepHalo + kTileW || ly < kStepHalo || ly >= kStepHalo + TH ||
                    ox >= nx || oy >= ny)
                    continue;
                const int64_t idx = (int64_t)oy * nx + ox;
                c2<T> out;
                out.x = z[j].x + p.dt6 * (acc[j].x + d.x);
                out.y = z[j].y + p.dt6 * (acc[j].y + d.y);
                T n_new[2] = {(T)0, (T)0};
#pragma unroll
                for (int q = 0; q < NF; ++q) {
                    T v = n0[j][q] + p.dt6 * (accn[j][q] + dn[q]);
                    v = v > (T)0 ? v : (T)0;  // set_state clamp (single_reservoir.py:70-73)
                    n_new[q] = v;
                    p.res_out[q][off + idx] = v;
                }
                if (p.noise) {
                    uint64_t step = p.step;
                    if (p.step_base) step += *p.step_base;
                    wigner_site<T>(out.x, out.y, mb.seed, (uint32_t)idx, step, 1u, (T)mb.noise_strength,
                                   (T)mb.noise_gamma, (T)mb.noise_R, NF > 0 ? n_new[0] : (T)0,
                                   (T)mb.noise_dt_over_4v);
                }
                p.psi_out[off + idx] = out;
            }
        }
        if (s < 3) __syncthreads();
    }
}

// ------------------------------------------------------------------ whole step, spinor
// Whole RK4 step of a spinor member tile on a periodic grid: the temporal blocking of
// rk4_step_kernel (64 x TH output points, a 4-point halo recomputed stage by stage, the stage
// inputs of both components double-buffered in LDS, the per-point state in registers) with
// spinor_rhs at every point and the spinor stage kernels' combination order, so that the two
// paths agree bit for bit.  Closed intervals take the stage kernels: their deeper halo would be
// recomputed beyond the wall, and the TE-TM cross term changes sign under a reflection, so the
// mirrored field does not evolve as the mirror image (the reference's spinor step kernel is
// periodic as well, rk4_cuda_solver.py:1835-1853).
template <typename T, int RES, int TH>
__global__ void __launch_bounds__(512) rk4_spinor_step_kernel(const Rk4StepParams<T> p) {
    constexpr int NT = 512;
    constexpr int NF = ResFields<RES>::n;
    constexpr int EW = kTileW + 2 * kStepHalo;
    constexpr int EH = TH + 2 * kStepHalo;
    constexpr int NP = EW * EH;
    constexpr int PPT = (NP + NT - 1) / NT;
    // each component's buffer has a one-row-plus-one pad on both sides (every neighbour
    // f +- EW +- 1 is an LDS address); buffer b of component c starts (2 b + c) PLANE in
    constexpr int PAD = EW + 1;
    constexpr int PLANE = NP + 2 * PAD;
    extern __shared__ unsigned char smem_raw[];
    c2<T>* base = reinterpret_cast<c2<T>*>(smem_raw) + PAD;

    const int tile = xcd_tile(blockIdx.x, p.active_tiles ? *p.active_tiles : p.num_tiles);
    if (tile < 0) return;
    const int m = p.tile_map[2 * tile];
    const pgpe_member mb = p.table[m];
    const int nx = mb.nx, ny = mb.ny;
    const int local = p.tile_map[2 * tile + 1];
    const int x0 = (local % mb.tiles_x) * kTileW;
    const int y0 = (local / mb.tiles_x) * TH;
    const int64_t plane = (int64_t)nx * ny;
    const int64_t poff = mb.off * 2;
    const int res_comps = p.res_comps;
    const bool two = res_comps == 2;
    const int64_t roff = mb.off * res_comps;
    T coef[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) coef[s] = p.coef[s * p.members + m];
    uint64_t step = p.step;
    if (p.noise && p.step_base) step += *p.step_base;

    c2<T> z[PPT][2], acc[PPT][2];
    T env[PPT], n0[PPT][2][2], ns[PPT][2][2], accn[PPT][2][2];
    const int tid = threadIdx.x;
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const int f = tid + NT * j;
#pragma unroll
        for (int rc = 0; rc < 2; ++rc)
#pragma unroll
            for (int q = 0; q < 2; ++q) n0[j][rc][q] = ns[j][rc][q] = accn[j][rc][q] = (T)0;
        if (f < NP) {
            const int ly = f / EW;
            const int lx = f - ly * EW;
            const int gy = map_index(y0 + ly - kStepHalo, ny, mb.boundary);
            const int gx = map_index(x0 + lx - kStepHalo, nx, mb.boundary);
            const int64_t cell = (int64_t)gy * nx + gx;
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                z[j][c] = p.psi0[poff + c * plane + cell];
                base[c * PLANE + f] = z[j][c];
            }
            env[j] = p.pump[mb.pump_off + cell];
#pragma unroll
            for (int rc = 0; rc < 2; ++rc) {
                if (rc >= res_comps) continue;
#pragma unroll
                for (int q = 0; q < NF; ++q) {
                    n0[j][rc][q] = p.res0[q][roff + rc * plane + cell];
                    ns[j][rc][q] = n0[j][rc][q];
                }
            }
        }
    }
    __syncthreads();

#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const c2<T>* cur = base + (s & 1) * 2 * PLANE;
        c2<T>* nxt = base + ((s + 1) & 1) * 2 * PLANE;
        const T a_next = s < 2 ? p.a_half : p.a_full;
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            const int f = tid + NT * j;
            if (f >= NP) continue;
            const T P = env[j] * coef[s];
            c2<T> d[2];
            T dn[2][2] = {{0, 0}, {0, 0}};
            spinor_rhs<T, RES>(mb, cur, cur + PLANE, f, EW, two, ns[j], P, d, dn);
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                if (s == 0) {
                    acc[j][c] = d[c];
                } else if (s < 3) {
                    acc[j][c].x = acc[j][c].x + (T)2 * d[c].x;
                    acc[j][c].y = acc[j][c].y + (T)2 * d[c].y;
                }
            }
#pragma unroll
            for (int rc = 0; rc < 2; ++rc) {
                if (rc >= res_comps) continue;
#pragma unroll
                for (int q = 0; q < NF; ++q) {
                    if (s == 0) accn[j][rc][q] = dn[rc][q];
                    else if (s < 3) accn[j][rc][q] = accn[j][rc][q] + (T)2 * dn[rc][q];
                }
            }
            if (s < 3) {
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    c2<T> v;
                    v.x = z[j][c].x + a_next * d[c].x;
                    v.y = z[j][c].y + a_next * d[c].y;
                    nxt[c * PLANE + f] = v;
                }
#pragma unroll
                for (int rc = 0; rc < 2; ++rc) {
                    if (rc >= res_comps) continue;
#pragma unroll
                    for (int q = 0; q < NF; ++q) ns[j][rc][q] = n0[j][rc][q] + a_next * dn[rc][q];
                }
            } else {
                const int ly = f / EW;
                const int lx = f - ly * EW;
                const int oy = y0 + ly - kStepHalo, ox = x0 + lx - kStepHalo;
                if (lx < kStepHalo || lx >= kStepHalo + kTileW || ly < kStepHalo || ly >= kStepHalo + TH ||
                    ox >= nx || oy >= ny)
                    continue;
                const int64_t cell = (int64_t)oy * nx + ox;
                T n_new[2][2] = {{0, 0}, {0, 0}};
#pragma unroll
                for (int rc = 0; rc < 2; ++rc) {
                    if (rc >= res_comps) continue;
#pragma unroll
                    for (int q = 0; q < NF; ++q) {
                        T v = n0[j][rc][q] + p.dt6 * (accn[j][rc][q] + dn[rc][q]);
                        v = v > (T)0 ? v : (T)0;  // set_state clamp (single_reservoir.py:70-73)
                        n_new[rc][q] = v;
                        p.res_out[q][roff + rc * plane + cell] = v;
                    }
                }
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    c2<T> out;
                    out.x = z[j][c].x + p.dt6 * (acc[j][c].x + d[c].x);
                    out.y = z[j][c].y + p.dt6 * (acc[j][c].y + d[c].y);
                    if (p.noise) {
                        const T act = NF > 0 ? (two ? n_new[c][0] : n_new[0][0]) : (T)0;
                        wigner_site<T>(out.x, out.y, mb.seed, (uint32_t)cell, step, 1u + (uint32_t)c,
                                       (T)mb.noise_strength, (T)mb.noise_gamma, (T)mb.noise_R, act,
                                       (T)mb.noise_dt_over_4v);
                    }
                    p.psi_out[poff + c * plane + cell] = out;
                }
            }
        }
        if (s < 3) __syncthreads();
    }
}

// ------------------------------------------------------------------ whole step, row strips
// The same four-stage temporal block as rk4_step_kernel with a layout that keeps the stencil's
// neighbours in registers.  A strip is 64 columns (one per lane of a wavefront: 56 output
// columns and 4 halo columns each side) by NW R rows (TH = NW R - 8 output rows and 4 halo rows
// each side); wavefront w of the NW-wavefront workgroup owns the R consecutive rows [wR, wR + R)
// of every column, so a point's y neighbours are the same thread's adjacent registers, its x
// neighbours the adjacent lanes' (read with DPP wave shifts, no LDS), and only the first and
// last row of each wavefront cross to the neighbouring wavefronts through LDS, once per stage.
// Per-point arithmetic is rk4_step_kernel's (the neighbour sums in the same order, kin5 / kin9,
// scalar_rhs, the same accumulation), so the two whole-step kernels and the stage kernels agree
// bit for bit on periodic grids.
constexpr int kStripW = kWave - 2 * kStepHalo;   // 56 output columns per strip

// Element e of an array at a uniform base, addressed by a 32-bit byte offset (the base in scalar
// registers, one vector register per access).
template <typename V>
__device__ __forceinline__ V& at_offset(V* base, uint32_t e) {
    return *reinterpret_cast<V*>(reinterpret_cast<char*>(base) + e * (uint32_t)sizeof(V));
}
template <typename V>
__device__ __forceinline__ const V& at_offset(const V* base, uint32_t e) {
    return *reinterpret_cast<const V*>(reinterpret_cast<const char*>(base) + e * (uint32_t)sizeof(V));
}

// Value of the next (lane + 1) or previous (lane - 1) lane of the wavefront (DPP wave_shl:1 /
// wave_shr:1); the last / first lane receives 0, a halo column whose value is never used.
__device__ __forceinline__ int lane_next_i(int v) { return __builtin_amdgcn_mov_dpp(v, 0x130, 0xF, 0xF, true); }
__device__ __forceinline__ int lane_prev_i(int v) { return __builtin_amdgcn_mov_dpp(v, 0x138, 0xF, 0xF, true); }
__device__ __forceinline__ float lane_next(float v) { return __builtin_bit_cast(float, lane_next_i(__builtin_bit_cast(int, v))); }
__device__ __forceinline__ float lane_prev(float v) { return __builtin_bit_cast(float, lane_prev_i(__builtin_bit_cast(int, v))); }
__device__ __forceinline__ double lane_next(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)lane_next_i((int)(uint32_t)u), hi = (uint32_t)lane_next_i((int)(uint32_t)(u >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double lane_prev(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)lane_prev_i((int)(uint32_t)u), hi = (uint32_t)lane_prev_i((int)(uint32_t)(u >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
template <typename T>
__device__ __forceinline__ c2<T> lane_next_c(const c2<T>& v) { return make_c<T>(lane_next(v.x), lane_next(v.y)); }
template <typename T>
__device__ __forceinline__ c2<T> lane_prev_c(const c2<T>& v) { return make_c<T>(lane_prev(v.x), lane_prev(v.y)); }

//
// Rows stale at a stage are skipped: after stage s the outer s + 1 rows of the strip can no
// longer feed an output point, so at stage s the first wavefront leaves out its rows j <= s and
// the last its rows j >= R - 1 - s (20 of the 4 NW R row-stages; the choice is per wavefront, so
// no lane diverges).  A skipped row still passes its stage input on as the next row's south
// neighbour; that value is stale, and so is every point it reaches.  FIVE compiles the 5-point
// Laplacian alone, for tables without 9-point members.
// A read-only array the kernel never writes, addressed through the constant address space so that
// uniform indices load through the scalar unit (member rows, launch maps, coefficients).
template <typename V>
__device__ __forceinline__ const __attribute__((address_space(4))) V* kconst(const V* ptr) {
    return (const __attribute__((address_space(4))) V*)ptr;
}

// A copy of a read-only record (a member row) through the constant address space: dword by dword,
// which the compiler merges into wide scalar loads when the address is uniform.
template <typename V>
__device__ __forceinline__ V load_const(const V* ptr) {
    static_assert(sizeof(V) % 4 == 0, "records are whole dwords");
    V out;
    const __attribute__((address_space(4))) uint32_t* s = (const __attribute__((address_space(4))) uint32_t*)ptr;
    uint32_t* d = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(V) / 4); ++i) d[i] = s[i];
    return out;
}

// Workgroup barrier for the strips' LDS exchange: this wavefront's LDS operations complete
// (lgkmcnt(0)), then s_barrier, with compiler-only fences on both sides so no memory access moves
// across it.  Its outstanding global loads, including LDS copies of a prefetch, stay in flight:
// __syncthreads, or a workgroup fence even restricted to LDS, would wait for those as well.  A
// strip wavefront reads no LDS another wavefront fills by copy, so nothing needs them here.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_s_waitcnt(0xC07F);  // s_waitcnt lgkmcnt(0)
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Where a strip sits: member, member row, first output column and row.
struct StripPos {
    int m, x0, y0;
};

// A strip map entry is (member, (row << 16) | column).
template <typename T, int TH>
__device__ __forceinline__ StripPos strip_pos(const Rk4StepParams<T>& p, int t) {
    StripPos s;
    s.m = kconst(p.tile_map)[2 * t];
    const int local = kconst(p.tile_map)[2 * t + 1];
    s.x0 = (local & 0xFFFF) * kStripW;
    s.y0 = (local >> 16) * TH;
    return s;
}

// Member-local element offsets of a wavefront's R rows in a strip, e = gy nx + gx of the wrapped or
// mirrored point (32-bit: the host keeps a strip member below 2^28 points).  An output point's
// offset is its own index, since output rows and columns never wrap.
template <int R>
__device__ __forceinline__ void strip_offsets(int nx, int ny, int boundary, int x0, int y0, int w, int lane,
                                              uint32_t e[R]) {
    const int gx = map_index(x0 - kStepHalo + lane, nx, boundary);
    const bool periodic = boundary != PGPE_BC_CLOSED;
    int gy = map_index(y0 - kStepHalo + w * R, ny, boundary);
#pragma unroll
    for (int j = 0; j < R; ++j) {
        if (j > 0) {
            // the next row: periodic rows step with a wrap; mirrored rows need the full map
            if (periodic) gy = gy + 1 == ny ? 0 : gy + 1;
            else gy = map_index(y0 - kStepHalo + w * R + j, ny, boundary);
        }
        e[j] = (uint32_t)(gy * nx + gx);
    }
}

// The four stages of one strip from psi0 (z), the pump envelope and n0 of the wavefront's R rows
// held in registers, and the stores of its output points.  Both strip kernels run this.
template <typename T, int RES, int NW, int R, bool ETA, bool NOISE, bool FIVE>
__device__ __forceinline__ void strip_stages(const Rk4StepParams<T>& p, c2<T> (&edge)[2][NW][2][kWave],
                                             const pgpe_member& mb, int m, int x0, int y0, int w, int lane,
                                             const uint32_t (&e)[R], const c2<T> (&z)[R], const T (&env)[R],
                                             const T (&n0)[R][2]) {
#pragma clang fp contract(on)
    constexpr int NF = ResFields<RES>::n;
    constexpr int TH = NW * R - 2 * kStepHalo;
    const int nx = mb.nx, ny = mb.ny;
    const int64_t off = mb.off;
    ScalarConsts<T> kc(mb);
    kc.to_scalar_registers();
    const bool nine = !FIVE && mb.laplacian == PGPE_LAP_NINE;
    const T centre = nine ? kc.c20 : kc.c0;
    T coef[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) coef[s] = kconst(p.coef)[s * p.members + m];
    const bool lane_out = lane >= kStepHalo && lane < kStepHalo + kStripW && x0 - kStepHalo + lane < nx;
    c2<T> cur[R], acc[R];
    T ns[R][2], accn[R][2];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        cur[j] = z[j];
#pragma unroll
        for (int q = 0; q < NF; ++q) ns[j][q] = n0[j][q];
    }

#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const T a_next = s < 2 ? p.a_half : p.a_full;
        edge[s & 1][w][0][lane] = cur[0];
        edge[s & 1][w][1][lane] = cur[R - 1];
        lds_barrier();
        // the first wavefront has no row below its first and the last none above its last: they
        // read their own edge rows instead, which only feed halo points that are stale by the end
        // of the step (a plain clamped index: no select between LDS and register values)
        c2<T> south = edge[s & 1][w > 0 ? w - 1 : 0][1][lane];
        const c2<T> above = edge[s & 1][w < NW - 1 ? w + 1 : NW - 1][0][lane];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const c2<T> ps = cur[j];
            if ((w == 0 && j <= s) || (w == NW - 1 && j >= R - 1 - s)) {
                south = ps;
                continue;
            }
            const c2<T> pn = j + 1 < R ? cur[j + 1] : above;
            const c2<T> pso = south;
            // x neighbours from the adjacent lanes (DPP wave shifts), y neighbours from this lane
            const c2<T> sx = csum<T>(lane_next_c<T>(ps), lane_prev_c<T>(ps));
            const c2<T> sy = csum<T>(pn, pso);
            c2<T> K;
            if (nine) {
                K = kin9<T>(kc, csum<T>(sy, sx),
                            csum<T>(csum<T>(csum<T>(lane_prev_c<T>(pso), lane_next_c<T>(pso)), lane_prev_c<T>(pn)),
                                    lane_next_c<T>(pn)));
            } else {
                K = kin5<T>(kc, sx, sy);
            }
            south = ps;
            const T P = env[j] * coef[s];
            c2<T> d;
            T dn[2];
            scalar_rhs<T, RES, ETA>(mb, kc, ps, K, centre, ns[j], P, false, make_c<T>((T)0, (T)0), d, dn);
            if (s == 0) {
                acc[j] = d;
#pragma unroll
                for (int q = 0; q < NF; ++q) accn[j][q] = dn[q];
            } else if (s < 3) {
                acc[j].x = acc[j].x + (T)2 * d.x;
                acc[j].y = acc[j].y + (T)2 * d.y;
#pragma unroll
                for (int q = 0; q < NF; ++q) accn[j][q] = accn[j][q] + (T)2 * dn[q];
            }
            if (s < 3) {
                cur[j].x = z[j].x + a_next * d.x;
                cur[j].y = z[j].y + a_next * d.y;
#pragma unroll
                for (int q = 0; q < NF; ++q) ns[j][q] = n0[j][q] + a_next * dn[q];
            } else {
                const int ly = w * R + j;  // uniform per wavefront
                if (ly < kStepHalo || ly >= kStepHalo + TH || y0 + ly - kStepHalo >= ny || !lane_out) continue;
                c2<T> out;
                out.x = z[j].x + p.dt6 * (acc[j].x + d.x);
                out.y = z[j].y + p.dt6 * (acc[j].y + d.y);
                T n_new[2] = {(T)0, (T)0};
#pragma unroll
                for (int q = 0; q < NF; ++q) {
                    T v = n0[j][q] + p.dt6 * (accn[j][q] + dn[q]);
                    v = v > (T)0 ? v : (T)0;  // set_state clamp (single_reservoir.py:70-73)
                    n_new[q] = v;
                    at_offset(p.res_out[q] + off, e[j]) = v;
                }
                if (NOISE) {
                    uint64_t step = p.step;
                    if (p.step_base) step += *kconst(p.step_base);
                    wigner_site<T>(out.x, out.y, mb.seed, e[j], step, 1u, (T)mb.noise_strength,
                                   (T)mb.noise_gamma, (T)mb.noise_R, NF > 0 ? n_new[0] : (T)0,
                                   (T)mb.noise_dt_over_4v);
                }
                at_offset(p.psi_out + off, e[j]) = out;
            }
        }
    }
}

// One strip per workgroup, psi0 / n0 / P loaded straight into registers (double precision).
template <typename T, int RES, int NW, int R, bool ETA, bool NOISE, bool FIVE>
__global__ void __launch_bounds__(NW * kWave) rk4_strip_kernel(const Rk4StepParams<T> p) {
    constexpr int NF = ResFields<RES>::n;
    constexpr int TH = NW * R - 2 * kStepHalo;
    static_assert(R >= kStepHalo, "a wavefront must hold at least the halo's rows");
    // first and last row of each wavefront, double buffered over the stages
    __shared__ c2<T> edge[2][NW][2][kWave];

    const int tile = xcd_tile(blockIdx.x, p.active_tiles ? *kconst(p.active_tiles) : p.num_tiles);
    if (tile < 0) return;
    const StripPos sp = strip_pos<T, TH>(p, tile);
    const pgpe_member mb = load_const(p.table + sp.m);
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));  // uniform per wavefront
    uint32_t e[R];
    strip_offsets<R>(mb.nx, mb.ny, mb.boundary, sp.x0, sp.y0, w, lane, e);
    const c2<T>* __restrict__ psi_m = p.psi0 + mb.off;
    const T* __restrict__ env_m = p.pump + mb.pump_off;
    c2<T> z[R];
    T env[R], n0[R][2];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        z[j] = at_offset(psi_m, e[j]);
        env[j] = at_offset(env_m, e[j]);
#pragma unroll
        for (int q = 0; q < NF; ++q) n0[j][q] = at_offset(p.res0[q] + mb.off, e[j]);
    }
    strip_stages<T, RES, NW, R, ETA, NOISE, FIVE>(p, edge, mb, sp.m, sp.x0, sp.y0, w, lane, e, z, env, n0);
}

// Single precision: persistent workgroups over the strips of their XCD's share of the launch map,
// with the next strip's psi0 / P / n0 loaded while the current one is computed.  Workgroup b runs
// on XCD b mod 8 and takes strips x per + b / 8, + G / 8, ... of that XCD's contiguous range
// [x per, (x + 1) per), per = ceil(count / 8), so neighbouring strips still meet in one L2.  Each
// wavefront issues the loads of its R rows of the next strip into a second set of registers
// before the four stages of the current one and moves them over when that strip starts, so the
// loads run under the arithmetic instead of ahead of it (the stage barriers wait for LDS only:
// lds_barrier).  The launch holds as many workgroups per CU as are resident at once (the
// occupancy of the instantiation: its registers decide), so every workgroup starts at once and
// the strips divide evenly.

template <int RES, int NW, int R, bool ETA, bool NOISE, bool FIVE>
__global__ void __launch_bounds__(NW * kWave) rk4_strip_pf_kernel(const Rk4StepParams<float> p) {
    using T = float;
    constexpr int NF = ResFields<RES>::n;
    constexpr int TH = NW * R - 2 * kStepHalo;
    static_assert(R >= kStepHalo, "a wavefront must hold at least the halo's rows");
    __shared__ c2<T> edge[2][NW][2][kWave];

    const int count = p.active_tiles ? *kconst(p.active_tiles) : p.num_tiles;
    const int xcd = blockIdx.x % kXcds;
    const int gpx = gridDim.x / kXcds;
    const int per = (count + kXcds - 1) / kXcds;
    const int t_end = min((xcd + 1) * per, count);
    int t = xcd * per + (int)(blockIdx.x / kXcds);
    if (t >= t_end) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));  // uniform per wavefront

    // this wavefront's rows of strip tt, loaded into the prefetch registers, with the strip's
    // position and row offsets kept for when it is computed
    c2<T> zq[R];
    T envq[R], n0q[R][2];
    StripPos sq;
    uint32_t eq[R];
    auto fetch = [&](int tt) {
        sq = strip_pos<T, TH>(p, tt);
        const auto& mb = kconst(p.table)[sq.m];
        strip_offsets<R>(mb.nx, mb.ny, mb.boundary, sq.x0, sq.y0, w, lane, eq);
        const c2<T>* psi_m = p.psi0 + mb.off;
        const T* env_m = p.pump + mb.pump_off;
#pragma unroll
        for (int j = 0; j < R; ++j) {
            zq[j] = at_offset(psi_m, eq[j]);
            envq[j] = at_offset(env_m, eq[j]);
#pragma unroll
            for (int q = 0; q < NF; ++q) n0q[j][q] = at_offset(p.res0[q] + mb.off, eq[j]);
        }
    };

    fetch(t);
    for (; t < t_end; t += gpx) {
        const StripPos sp = sq;
        const pgpe_member mb = load_const(p.table + sp.m);
        uint32_t e[R];
        c2<T> z[R];
        T env[R], n0[R][2];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            e[j] = eq[j];
            z[j] = zq[j];
            env[j] = envq[j];
#pragma unroll
            for (int q = 0; q < NF; ++q) n0[j][q] = n0q[j][q];
        }
        if (t + gpx < t_end) fetch(t + gpx);
        strip_stages<T, RES, NW, R, ETA, NOISE, FIVE>(p, edge, mb, sp.m, sp.x0, sp.y0, w, lane, e, z, env, n0);
    }
}

// CUs of the current device, asked once per device.
static int device_cus() {
    static int cus[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (cus[dev] == 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cus[dev] = n;
    }
    return cus[dev];
}

template <typename T, int RES, int NW, int R, bool ETA, bool NOISE, bool FIVE>
static void launch_strip_k(const pgpe_rk4_args* a, const Rk4StepParams<T>& p, hipStream_t s) {
    const dim3 block(NW * kWave);
    if constexpr (sizeof(T) == 4) {
        // persistent: the resident workgroups of every CU (at most one per strip), a multiple of
        // the XCD count
        auto kern = rk4_strip_pf_kernel<RES, NW, R, ETA, NOISE, FIVE>;
        static const int per_cu = [&] {
            int n = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, NW * kWave, 0) != hipSuccess || n < 1) n = 1;
            return n;
        }();
        const int grid = std::min(xcd_grid(a->num_tiles), xcd_grid(per_cu * device_cus()));
        hipLaunchKernelGGL(kern, dim3(grid), block, 0, s, p);
    } else {
        hipLaunchKernelGGL((rk4_strip_kernel<T, RES, NW, R, ETA, NOISE, FIVE>), dim3(xcd_grid(a->num_tiles)), block,
                           0, s, p);
    }
}

template <typename T, int RES, int NW, int R, bool FIVE>
static void launch_strip_f(const pgpe_rk4_args* a, const Rk4StepParams<T>& p, hipStream_t s) {
    // the kinetic-relaxation term and the Wigner increment are compiled in only when used: each
    // costs registers (the noise generator ~25 per lane), and fewer lift the occupancy
    if (a->no_eta && !a->noise) launch_strip_k<T, RES, NW, R, false, false, FIVE>(a, p, s);
    else if (a->no_eta) launch_strip_k<T, RES, NW, R, false, true, FIVE>(a, p, s);
    else if (!a->noise) launch_strip_k<T, RES, NW, R, true, false, FIVE>(a, p, s);
    else launch_strip_k<T, RES, NW, R, true, true, FIVE>(a, p, s);
}

template <typename T, int RES, int NW, int R>
static int launch_strip_r(const pgpe_rk4_args* a, const Rk4StepParams<T>& p, hipStream_t s) {
    if (a->five_point) launch_strip_f<T, RES, NW, R, true>(a, p, s);
    else launch_strip_f<T, RES, NW, R, false>(a, p, s);
    PGPE_LAUNCHED("rk4_strip_kernel");
    return PGPE_OK;
}

template <typename T, int RES>
static int launch_strip_res(const pgpe_rk4_args* a, const Rk4StepParams<T>& p, hipStream_t s) {
    // strip height TH = NW R - 8: wavefronts per workgroup x rows per wavefront, less the halo
    switch (a->tile_y) {
        case 8: return launch_strip_r<T, RES, 4, 4>(a, p, s);
        case 16: return launch_strip_r<T, RES, 4, 6>(a, p, s);
        case 24: return launch_strip_r<T, RES, 8, 4>(a, p, s);
        case 32:
        case 40:
            // double precision needs more than the 256 registers a 512-thread workgroup allows
            if constexpr (sizeof(T) == 4) {
                if (a->tile_y == 32) return launch_strip_r<T, RES, 8, 5>(a, p, s);
                return launch_strip_r<T, RES, 8, 6>(a, p, s);
            }
            return fail(PGPE_ERR_ARG, "pgpe_rk4_step: strip heights 32 and 40 are single precision only");
        default: return fail(PGPE_ERR_ARG, "pgpe_rk4_step: strip height (tile_y) must be 8, 16, 24, 32 or 40");
    }
}

template <typename T, int RES, int TH>
static int launch_step_th(const pgpe_rk4_args* a, const Rk4StepParams<T>& p, hipStream_t s) {
    constexpr int NP = (kTileW + 2 * kStepHalo) * (TH + 2 * kStepHalo);
    constexpr int PAD = kTileW + 2 * kStepHalo + 1;
    const size_t shm = 2 * (NP + 2 * PAD) * sizeof(c2<T>);
    if (shm > 160 * 1024) return fail(PGPE_ERR_ARG, "pgpe_rk4_step: tile too tall for LDS at this precision");
    auto kern = rk4_step_kernel<T, RES, TH>;
    if (shm > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute(rk4_step_kernel)");
    }
    hipLaunchKernelGGL(kern, dim3(xcd_grid(a->num_tiles)), dim3(256), shm, s, p);
    PGPE_LAUNCHED("rk4_step_kernel");
    return PGPE_OK;
}

// Spinor whole step, single precision: 64 x 16 tiles on 512 threads (the spinor stage kernels'
// tile map of height 16).  Double precision would hold more than the 256 registers a 512-thread
// workgroup allows, so double-precision spinors take the stage kernels.
template <typename T, int RES>
static int launch_spinor_step(const pgpe_rk4_args* a, const Rk4StepParams<T>& p, hipStream_t s) {
    constexpr int TH = 16;
    if constexpr (sizeof(T) != 4) {
        return fail(PGPE_ERR_ARG, "pgpe_rk4_step: spinor fields are single precision only (the stage kernels take both)");
    } else {
    if (a->strips || a->tile_y != TH) return fail(PGPE_ERR_ARG, "pgpe_rk4_step: spinor fields take tile_y 16 (no strips)");
    constexpr int NP = (kTileW + 2 * kStepHalo) * (TH + 2 * kStepHalo);
    constexpr int PAD = kTileW + 2 * kStepHalo + 1;
    constexpr size_t shm = 4 * (NP + 2 * PAD) * sizeof(c2<T>);
    static_assert(shm <= 160 * 1024, "spinor step tile exceeds the LDS");
    auto kern = rk4_spinor_step_kernel<T, RES, TH>;
    if (shm > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute(rk4_spinor_step_kernel)");
    }
    hipLaunchKernelGGL(kern, dim3(xcd_grid(a->num_tiles)), dim3(512), shm, s, p);
    PGPE_LAUNCHED("rk4_spinor_step_kernel");
    return PGPE_OK;
    }
}

template <typename T, int RES>
static int launch_step_res(const pgpe_rk4_args* a, const Rk4StepParams<T>& p, hipStream_t s) {
    if (a->components == 2) return launch_spinor_step<T, RES>(a, p, s);
    if (a->strips) return launch_strip_res<T, RES>(a, p, s);
    switch (a->tile_y) {
        case 16: return launch_step_th<T, RES, 16>(a, p, s);
        case 32: return launch_step_th<T, RES, 32>(a, p, s);
        case 64:
            if constexpr (sizeof(T) == 4) return launch_step_th<T, RES, 64>(a, p, s);
            return fail(PGPE_ERR_ARG, "pgpe_rk4_step: tile_y 64 is single precision only");
        default: return fail(PGPE_ERR_ARG, "pgpe_rk4_step: tile_y must be 16, 32 or 64");
    }
}

template <typename T>
static int launch_step_typed(const pgpe_rk4_args* a, hipStream_t s) {
    Rk4StepParams<T> p{};
    p.table = a->table;
    p.tile_map = a->tile_map;
    p.num_tiles = a->num_tiles;
    p.noise = a->noise;
    p.members = a->members;
    p.a_half = (T)(0.5 * a->dt);
    p.a_full = (T)(1.0 * a->dt);
    p.dt6 = (T)(a->dt / 6.0);
    p.psi0 = (const c2<T>*)a->psi0;
    p.psi_out = (c2<T>*)a->psi_out;
    for (int f = 0; f < 2; ++f) {
        p.res0[f] = (const T*)a->res0[f];
        p.res_out[f] = (T*)a->res_out[f];
    }
    p.pump = (const T*)a->pump;
    p.coef = (const T*)a->coef;
    p.step = a->step;
    p.step_base = a->step_base;
    p.active_tiles = a->active_tiles;
    p.res_comps = a->components == 2 ? a->res_components : 1;
    switch (a->reservoir) {
        case PGPE_RES_NONE: return launch_step_res<T, PGPE_RES_NONE>(a, p, s);
        case PGPE_RES_SINGLE: return launch_step_res<T, PGPE_RES_SINGLE>(a, p, s);
        case PGPE_RES_DOUBLE: return launch_step_res<T, PGPE_RES_DOUBLE>(a, p, s);
        case PGPE_RES_QDOUBLE: return launch_step_res<T, PGPE_RES_QDOUBLE>(a, p, s);
        default: return fail(PGPE_ERR_ARG, "pgpe_rk4_step: unknown reservoir kind");
    }
}

template <typename T, int RES, int TH>
static int launch_kind(const pgpe_rk4_args* a, const Rk4Params<T>& p, hipStream_t s) {
    const dim3 grid(xcd_grid(a->num_tiles)), block(256);
    if (a->components == 2) {
        const int rc = a->res_components;
        if (a->stage == 0)
            hipLaunchKernelGGL((rk4_spinor_stage_kernel<T, RES, TH, 0>), grid, block, 0, s, p, rc);
        else if (a->stage < 3)
            hipLaunchKernelGGL((rk4_spinor_stage_kernel<T, RES, TH, 1>), grid, block, 0, s, p, rc);
        else
            hipLaunchKernelGGL((rk4_spinor_stage_kernel<T, RES, TH, 2>), grid, block, 0, s, p, rc);
        PGPE_LAUNCHED("rk4_spinor_stage_kernel");
        return PGPE_OK;
    }
    if (a->stage == 0)
        hipLaunchKernelGGL((rk4_stage_kernel<T, RES, TH, 0>), grid, block, 0, s, p);
    else if (a->stage < 3)
        hipLaunchKernelGGL((rk4_stage_kernel<T, RES, TH, 1>), grid, block, 0, s, p);
    else
        hipLaunchKernelGGL((rk4_stage_kernel<T, RES, TH, 2>), grid, block, 0, s, p);
    PGPE_LAUNCHED("rk4_stage_kernel");
    return PGPE_OK;
}

template <typename T, int RES>
static int launch_th(const pgpe_rk4_args* a, const Rk4Params<T>& p, hipStream_t s) {
    switch (a->tile_y) {
        case 8: return launch_kind<T, RES, 8>(a, p, s);
        case 16: return launch_kind<T, RES, 16>(a, p, s);
        case 32: return launch_kind<T, RES, 32>(a, p, s);
        default: return fail(PGPE_ERR_ARG, "pgpe_rk4_stage: tile_y must be 8, 16 or 32");
    }
}

template <typename T>
static int launch_typed(const pgpe_rk4_args* a, hipStream_t s) {
    Rk4Params<T> p{};
    p.table = a->table;
    p.tile_map = a->tile_map;
    p.num_tiles = a->num_tiles;
    p.noise = a->noise;
    static const double scale[4] = {0.0, 0.5, 0.5, 1.0};
    p.a = (T)(scale[a->stage] * a->dt);
    p.dt6 = (T)(a->dt / 6.0);
    p.psi0 = (const c2<T>*)a->psi0;
    p.kprev = (const c2<T>*)a->kprev;
    p.k1 = (const c2<T>*)a->k1;
    p.k2 = (const c2<T>*)a->k2;
    p.kout = (c2<T>*)a->kout;
    p.psi_out = (c2<T>*)a->psi_out;
    for (int f = 0; f < 2; ++f) {
        p.res0[f] = (const T*)a->res0[f];
        p.kres_prev[f] = (const T*)a->kres_prev[f];
        p.kres1[f] = (const T*)a->kres1[f];
        p.kres2[f] = (const T*)a->kres2[f];
        p.kres_out[f] = (T*)a->kres_out[f];
        p.res_out[f] = (T*)a->res_out[f];
    }
    p.pump = (const T*)a->pump;
    p.coef = (const T*)a->coef;
    p.step = a->step;
    p.step_base = a->step_base;
    p.potential = (const c2<T>*)a->potential;
    p.active_tiles = a->active_tiles;
    switch (a->reservoir) {
        case PGPE_RES_NONE: return launch_th<T, PGPE_RES_NONE>(a, p, s);
        case PGPE_RES_SINGLE: return launch_th<T, PGPE_RES_SINGLE>(a, p, s);
        case PGPE_RES_DOUBLE: return launch_th<T, PGPE_RES_DOUBLE>(a, p, s);
        case PGPE_RES_QDOUBLE: return launch_th<T, PGPE_RES_QDOUBLE>(a, p, s);
        default: return fail(PGPE_ERR_ARG, "pgpe_rk4_stage: unknown reservoir kind");
    }
}

// Host-side validation of the pointers a stage dereferences.
static int check_args(const pgpe_rk4_args* a) {
    if (!a) return fail(PGPE_ERR_ARG, "pgpe_rk4_stage: null args");
    if (a->stage < 0 || a->stage > 3) return fail(PGPE_ERR_ARG, "pgpe_rk4_stage: stage must be 0..3");
    if (a->num_tiles <= 0 || a->members <= 0) return fail(PGPE_ERR_ARG, "pgpe_rk4_stage: empty member table");
    if (a->components != 1 && a->components != 2) return fail(PGPE_ERR_ARG, "pgpe_rk4_stage: components must be 1 or 2");
    if (a->components == 2 && a->res_components != 1 && a->res_components != 2)
        return fail(PGPE_ERR_ARG, "pgpe_rk4_stage: res_components must be 1 or 2");
    if (a->components == 2 && a->potential)
        return fail(PGPE_ERR_ARG, "pgpe_rk4_stage: spinor fields take no potential here");
    if (!a->table || !a->tile_map || !a->psi0 || !a->pump || !a->coef)
        return fail(PGPE_ERR_ARG, "pgpe_rk4_stage: null table, field, pump or coefficient");
    const int nf = a->reservoir == PGPE_RES_NONE ? 0 : (a->reservoir == PGPE_RES_SINGLE ? 1 : 2);
    for (int f = 0; f < nf; ++f)
        if (!a->res0[f]) return fail(PGPE_ERR_ARG, "pgpe_rk4_stage: null reservoir field");
    if (a->stage > 0) {
        if (!a->kprev) return fail(PGPE_ERR_ARG, "pgpe_rk4_stage: null kprev");
        for (int f = 0; f < nf; ++f)
            if (!a->kres_prev[f]) return fail(PGPE_ERR_ARG, "pgpe_rk4_stage: null kres_prev");
    }
    if (a->stage < 3) {
        if (!a->kout) return fail(PGPE_ERR_ARG, "pgpe_rk4_stage: null kout");
        for (int f = 0; f < nf; ++f)
            if (!a->kres_out[f]) return fail(PGPE_ERR_ARG, "pgpe_rk4_stage: null kres_out");
    } else {
        if (!a->k1 || !a->k2 || !a->psi_out) return fail(PGPE_ERR_ARG, "pgpe_rk4_stage: null k1/k2/psi_out");
        if (a->psi_out == a->psi0) return fail(PGPE_ERR_ARG, "pgpe_rk4_stage: psi_out must not alias psi0");
        for (int f = 0; f < nf; ++f)
            if (!a->kres1[f] || !a->kres2[f] || !a->res_out[f])
                return fail(PGPE_ERR_ARG, "pgpe_rk4_stage: null reservoir stage buffers");
    }
    return PGPE_OK;
}

}  // namespace pgpe

using namespace pgpe;

extern "C" int pgpe_rk4_stage(const pgpe_rk4_args* a, void* stream) {
    int rc = check_args(a);
    if (rc) return rc;
    if (a->precision == 0) return launch_typed<float>(a, (hipStream_t)stream);
    if (a->precision == 1) return launch_typed<double>(a, (hipStream_t)stream);
    return fail(PGPE_ERR_ARG, "pgpe_rk4_stage: precision must be 0 or 1");
}

extern "C" int pgpe_rk4_step(const pgpe_rk4_args* a, void* stream) {
    if (!a) return fail(PGPE_ERR_ARG, "pgpe_rk4_step: null args");
    if (a->num_tiles <= 0 || a->members <= 0) return fail(PGPE_ERR_ARG, "pgpe_rk4_step: empty member table");
    if (a->components != 1 && a->components != 2) return fail(PGPE_ERR_ARG, "pgpe_rk4_step: components must be 1 or 2");
    if (a->components == 2 && a->res_components != 1 && a->res_components != 2)
        return fail(PGPE_ERR_ARG, "pgpe_rk4_step: res_components must be 1 or 2");
    if (!a->table || !a->tile_map || !a->psi0 || !a->psi_out || !a->pump || !a->coef)
        return fail(PGPE_ERR_ARG, "pgpe_rk4_step: null table, field, pump or coefficient");
    if (a->psi_out == a->psi0) return fail(PGPE_ERR_ARG, "pgpe_rk4_step: psi_out must not alias psi0");
    if (a->potential) return fail(PGPE_ERR_ARG, "pgpe_rk4_step: a potential needs the stage kernels (pgpe_rk4_stage)");
    const int nf = a->reservoir == PGPE_RES_NONE ? 0 : (a->reservoir == PGPE_RES_SINGLE ? 1 : 2);
    for (int f = 0; f < nf; ++f) {
        if (!a->res0[f] || !a->res_out[f]) return fail(PGPE_ERR_ARG, "pgpe_rk4_step: null reservoir field");
        // the halo reads neighbouring tiles' reservoir, so the update cannot be in place
        if (a->res_out[f] == a->res0[f]) return fail(PGPE_ERR_ARG, "pgpe_rk4_step: res_out must not alias res0");
    }
    if (a->precision == 0) return launch_step_typed<float>(a, (hipStream_t)stream);
    if (a->precision == 1) return launch_step_typed<double>(a, (hipStream_t)stream);
    return fail(PGPE_ERR_ARG, "pgpe_rk4_step: precision must be 0 or 1");
}
