// Batched 2-D FFT (row + column LDS Stockham kernels) and the fused
// split-step FFT step of the driven-dissipative GPE.
//
// A 2-D transform of a batch of (ny, nx) planes is two launches:
//   rows    : every row (contiguous along x) transformed in LDS,
//   columns : a workgroup loads C whole columns (C consecutive x, read as
//             C-element row segments so HBM access stays coalesced),
//             transforms them along y in LDS and writes them back.
// The split-step step fuses the pointwise physics into the transforms so each
// pass reads and writes the field once:
//   CPU order  (split_step_fft_solver.py:97-159):
//     rows  : reservoir half-step (P at t+dt/4), forward FFT along x
//     cols  : forward FFT along y, x exp(-i hbar k^2 dt/(4m)), inverse along y
//     rows  : inverse along x, exact nonlinear step, Wigner noise, forward along x
//     cols  : as above
//     rows  : inverse along x, reservoir half-step (P at t+3dt/4)
//   Chained steps of the CPU order (inside a graph chunk) replace the last row pass of a step
//   and the first of the next by one: inverse along x, this step's reservoir half-step (P at
//   t+3dt/4), the next step's (P at t+dt+dt/4), forward along x.  The operations and their
//   order are those of the two passes, so the chain is bit-identical to separate steps.
//   K-outer (split_step_fft_cuda_solver.py:106-184 order):
//     rows  : forward;  cols : propagator;
//     rows  : inverse, R(dt/2) N(dt) R(dt/2), noise, forward;
//     cols  : propagator;  rows : inverse.
#include <cmath>
#include <new>
#include <vector>

#include "split_common.hpp"

namespace pgpe {


// ---------------------------------------------------------------- rows
template <typename T, int V, int MODE, int RES, int NT>
__global__ void __launch_bounds__(NT) fft_rows_kernel(const SplitParams<T> p) {
    extern __shared__ unsigned char smem_raw[];
    c2<T>* buf = reinterpret_cast<c2<T>*>(smem_raw);
    constexpr int E = NT * V;
    const int n = p.nx;
    const int rows_per = E / n;
    const int ld = seq_stride<T>(n);
    c2<T>* s_tw = buf + rows_per * ld;  // twiddles staged once per group: every pass reads them from LDS
    const FastDiv by_n((uint32_t)n);
    auto slot = [&](int i) {  // LDS slot of the group's i-th element (row i / n, column i % n)
        const int r = by_n.div(i);
        return r * ld + pad_at(i - r * n);
    };
    const int total_rows = p.planes * p.ny;
    const int row0 = blockIdx.x * rows_per;
    const int count = min(rows_per, total_rows - row0);
    if (count <= 0) return;
    stage_twiddles<T, NT>(s_tw, p.tw, n);
    const int elems = count * n;
    const int64_t g0 = (int64_t)row0 * n;
    const bool res_here = RES != 0;
    uint64_t step = p.step;
    if (p.step_base) step += *p.step_base;

    // ---- load (+ prologue) ----
    constexpr int CH = stage_chunk<T, V>();
    if constexpr (MODE == ROW_SS_FIRST) {
        struct Pt { c2<T> v; T n[2]; T P; };
        staged<Pt, CH, NT>(
            elems,
            [&](int i) {
                Pt q;
                q.v = p.in[g0 + i];
                const int r = by_n.div(i);
                const int row = row0 + r;
                const int m = row / p.ny;
                const int64_t cell = (int64_t)(row - m * p.ny) * n + (i - r * n);
                q.n[0] = q.n[1] = (T)0;
                load_res<T, RES>(p, g0 + i, q.n);
                q.P = p.pump[p.table[m].pump_off + cell] * p.coef[m];
                return q;
            },
            [&](int i, Pt q) {
                const int m = (row0 + by_n.div(i)) / p.ny;
                res_half_step<T, RES>(p.table[m], q.v.x * q.v.x + q.v.y * q.v.y, q.P, p.dt / (T)2, q.n);
                if (res_here) store_res<T, RES>(p, g0 + i, q.n);
                buf[slot(i)] = q.v;
            });
    } else {
        staged<c2<T>, CH, NT>(elems, [&](int i) { return p.in[g0 + i]; }, [&](int i, c2<T> v) { buf[slot(i)] = v; });
    }
    __syncthreads();
    const bool first_inverse = (MODE == ROW_INV || MODE == ROW_SS_MID || MODE == ROW_SS_LAST || MODE == ROW_KO_MID ||
                                MODE == ROW_SS_JOIN);
    lds_fft<T, V, NT>(buf, p.dims, ld, count, s_tw, first_inverse);

    if constexpr (MODE == ROW_SS_JOIN) {
        // inverse done: scale, this step's last reservoir half-step (coefficient row 1), the next
        // step's first (row 2), then forward again in place
        struct Pt { c2<T> v; T n[2]; T P1, P2; };
        staged<Pt, CH, NT>(
            elems,
            [&](int i) {
                Pt q;
                q.v = buf[slot(i)];
                q.n[0] = q.n[1] = (T)0;
                load_res<T, RES>(p, g0 + i, q.n);
                const int r = by_n.div(i);
                const int row = row0 + r;
                const int m = row / p.ny;
                const int64_t cell = (int64_t)(row - m * p.ny) * n + (i - r * n);
                const T env = p.pump[p.table[m].pump_off + cell];
                q.P1 = env * p.coef[p.members + m];
                q.P2 = env * p.coef[2 * p.members + m];
                return q;
            },
            [&](int i, Pt q) {
                c2<T> v = q.v;
                v.x *= p.scale;
                v.y *= p.scale;
                const int m = (row0 + by_n.div(i)) / p.ny;
                res_half_step<T, RES>(p.table[m], v.x * v.x + v.y * v.y, q.P1, p.dt / (T)2, q.n);
                res_half_step<T, RES>(p.table[m], v.x * v.x + v.y * v.y, q.P2, p.dt / (T)2, q.n);
                if (res_here) store_res<T, RES>(p, g0 + i, q.n);
                buf[slot(i)] = v;
            });
        __syncthreads();
        lds_fft<T, V, NT>(buf, p.dims, ld, count, s_tw, false);
    }

    if constexpr (MODE == ROW_SS_MID || MODE == ROW_KO_MID) {
        // inverse done: scale, local physics, then forward again in place
        struct Pt { c2<T> v; T n[2]; T P1, P2; };
        staged<Pt, CH, NT>(
            elems,
            [&](int i) {
                Pt q;
                q.v = buf[slot(i)];
                q.n[0] = q.n[1] = (T)0;
                load_res<T, RES>(p, g0 + i, q.n);
                if constexpr (MODE == ROW_KO_MID) {
                    const int r = by_n.div(i);
                    const int row = row0 + r;
                    const int m = row / p.ny;
                    const int64_t cell = (int64_t)(row - m * p.ny) * n + (i - r * n);
                    const T env = p.pump[p.table[m].pump_off + cell];
                    q.P1 = env * p.coef[m];
                    q.P2 = env * p.coef[p.members + m];
                }
                return q;
            },
            [&](int i, Pt q) {
                c2<T> v = q.v;
                v.x *= p.scale;
                v.y *= p.scale;
                const int r = by_n.div(i);
                const int row = row0 + r;
                const int m = row / p.ny;
                const int64_t cell = (int64_t)(row - m * p.ny) * n + (i - r * n);
                const pgpe_member& mb = p.table[m];
                T* nn = q.n;
                if constexpr (MODE == ROW_KO_MID) {
                    res_half_step<T, RES>(mb, v.x * v.x + v.y * v.y, q.P1, p.dt / (T)2, nn);
                    nonlinear_step<T>(mb, RES ? nn[0] : (T)0, RES >= 2 ? nn[1] : (T)0, p.dt, v);
                    res_half_step<T, RES>(mb, v.x * v.x + v.y * v.y, q.P2, p.dt / (T)2, nn);
                    if (res_here) store_res<T, RES>(p, g0 + i, nn);
                } else {
                    nonlinear_step<T>(mb, RES ? nn[0] : (T)0, RES >= 2 ? nn[1] : (T)0, p.dt, v);
                }
                if (p.noise) {
                    wigner_site<T>(v.x, v.y, mb.seed, (uint32_t)cell, step, 1u, (T)mb.noise_strength,
                                   (T)mb.noise_gamma, (T)mb.noise_R, RES ? nn[0] : (T)0, (T)mb.noise_dt_over_4v);
                }
                buf[slot(i)] = v;
            });
        __syncthreads();
        lds_fft<T, V, NT>(buf, p.dims, ld, count, s_tw, false);
    }

    // ---- store (+ epilogue) ----
    if constexpr (MODE == ROW_SS_LAST) {
        struct Pt { c2<T> v; T n[2]; T P; };
        staged<Pt, CH, NT>(
            elems,
            [&](int i) {
                Pt q;
                q.v = buf[slot(i)];
                const int r = by_n.div(i);
                const int row = row0 + r;
                const int m = row / p.ny;
                const int64_t cell = (int64_t)(row - m * p.ny) * n + (i - r * n);
                q.n[0] = q.n[1] = (T)0;
                load_res<T, RES>(p, g0 + i, q.n);
                q.P = p.pump[p.table[m].pump_off + cell] * p.coef[p.members + m];
                return q;
            },
            [&](int i, Pt q) {
                c2<T> v = q.v;
                v.x *= p.scale;
                v.y *= p.scale;
                const int m = (row0 + by_n.div(i)) / p.ny;
                res_half_step<T, RES>(p.table[m], v.x * v.x + v.y * v.y, q.P, p.dt / (T)2, q.n);
                if (res_here) store_res<T, RES>(p, g0 + i, q.n);
                p.out[g0 + i] = v;
            });
    } else {
        staged<c2<T>, CH, NT>(
            elems, [&](int i) { return buf[slot(i)]; },
            [&](int i, c2<T> v) {
                if constexpr (MODE == ROW_INV) {
                    v.x *= p.scale;
                    v.y *= p.scale;
                }
                p.out[g0 + i] = v;
            });
    }
}

// ---------------------------------------------------------------- columns
template <typename T, int V, int MODE, int NT>
__global__ void __launch_bounds__(NT) fft_cols_kernel(const SplitParams<T> p) {
    extern __shared__ unsigned char smem_raw[];
    c2<T>* buf = reinterpret_cast<c2<T>*>(smem_raw);
    constexpr int E = NT * V;
    const int n = p.ny;
    const int cols_per = E / n;
    const int ld = seq_stride<T>(n);
    c2<T>* s_tw = buf + cols_per * ld;  // twiddles staged once per group
    const int groups_per_plane = (p.nx + cols_per - 1) / cols_per;
    // XCD-aware: the groups one XCD runs side by side are neighbours along x, so the row
    // segments they share come through the same L2
    const int g = xcd_tile(blockIdx.x, groups_per_plane * p.planes);
    if (g < 0) return;
    const int plane = g / groups_per_plane;
    const int x0 = (g - plane * groups_per_plane) * cols_per;
    const int count = min(cols_per, p.nx - x0);
    const int64_t base = (int64_t)plane * n * p.nx;
    const FastDiv by_count((uint32_t)count), by_n((uint32_t)n);
    stage_twiddles<T, NT>(s_tw, p.tw, n);
    // load column segments: element (y, x0 + c) -> buf[c * ld + y]
    constexpr int CH = stage_chunk<T, V>();
    struct Pt { c2<T> v; int at; };
    staged<Pt, CH, NT>(
        count * n,
        [&](int i) {
            const int y = by_count.div(i);
            const int c = i - y * count;
            return Pt{p.in[base + (int64_t)y * p.nx + x0 + c], c * ld + pad_at(y)};
        },
        [&](int, Pt q) { buf[q.at] = q.v; });
    __syncthreads();
    lds_fft<T, V, NT>(buf, p.dims, ld, count, s_tw, MODE == COL_INV);
    if constexpr (MODE == COL_PROP) {
        // x exp(-i hbar k^2 h / (2 m)), evaluated in double as the reference does
        const pgpe_member& mb = p.table[plane];
        const double hbar = mb.hbar;
        const double inv2m = 1.0 / (2 * mb.m_eff);
        struct K { double kx, ky; };
        staged<K, CH, NT>(
            count * n, [&](int i) { const int c = by_n.div(i); return K{p.kx[x0 + c], p.ky[i - c * n]}; },
            [&](int i, K q) {
                const int c = by_n.div(i);
                const int y = i - c * n;
                const double k2 = q.kx * q.kx + q.ky * q.ky;
                const double theta = ((-(hbar * k2)) * p.prop_h) * inv2m;
                double s, co;
                sincos(theta, &s, &co);
                const c2<T> v = buf[c * ld + pad_at(y)];
                const double re = (double)v.x * co - (double)v.y * s;
                const double im = (double)v.x * s + (double)v.y * co;
                buf[c * ld + pad_at(y)] = make_c<T>((T)re, (T)im);
            });
        __syncthreads();
        lds_fft<T, V, NT>(buf, p.dims, ld, count, s_tw, true);
    }
    staged<Pt, CH, NT>(
        count * n,
        [&](int i) {
            const int y = by_count.div(i);
            const int c = i - y * count;
            return Pt{buf[c * ld + pad_at(y)], y * p.nx + x0 + c};
        },
        [&](int, Pt q) { p.out[base + q.at] = q.v; });
}

// ---------------------------------------------------------------- tall columns, split in two
// A forward or inverse column transform of length n = 2h (a power of two) as one radix-2
// decimation-in-frequency step in registers and two h-point LDS transforms:
//   a[y] = x[y] + x[y + h],  b[y] = (x[y] - x[y + h]) w^y,  w = exp(-+2 pi i / n),
//   X[2k] = DFT_h(a)[k],     X[2k + 1] = DFT_h(b)[k].
// Every thread loads its V pairs (all 2V loads in flight), keeps its b values in registers while
// the group transforms and stores the even outputs, then transforms b into the odd outputs.
// The LDS holds C half columns (C = 8 complex64 or 4 complex128: 64-byte row segments, as the
// one-pass kernel), half of what the one-pass kernel needs, so two groups share a CU and one
// group's loads and stores overlap the other's passes.  FP32 roots carry tails (the slot
// y + h of `root`), as the plan's early-pass twiddles do.
template <typename T, int V, int MODE, int NT, int CLS>
__global__ void __launch_bounds__(NT, 2 * NT / 256) fft_cols_half_kernel(const SplitParams<T> p,
                                                                        const c2<T>* __restrict__ root) {
    extern __shared__ unsigned char smem_raw[];
    c2<T>* buf = reinterpret_cast<c2<T>*>(smem_raw);
    constexpr int E = NT * V;
    const int h = p.ny >> 1;
    const int cols_per = E / h;
    const int ld = seq_stride<T>(h);
    c2<T>* s_tw = buf + cols_per * ld;
    const int groups_per_plane = (p.nx + cols_per - 1) / cols_per;
    const int g = xcd_tile(blockIdx.x, groups_per_plane * p.planes);
    if (g < 0) return;
    const int plane = g / groups_per_plane;
    const int x0 = (g - plane * groups_per_plane) * cols_per;
    const int count = min(cols_per, p.nx - x0);
    const int64_t base = (int64_t)plane * p.ny * p.nx + x0;
    const int64_t half = (int64_t)h * p.nx;
    const int elems = count * h;
    const FastDiv by_count((uint32_t)count);
    constexpr bool inverse = MODE == COL_INV;
    stage_twiddles<T, NT>(s_tw, p.tw, h);

    // the pairs in batches of PB: 2 PB loads in flight per thread, within the register budget
    constexpr int PB = V < 8 ? V : 8;
    c2<T> b[V];
#pragma unroll
    for (int k0 = 0; k0 < V; k0 += PB) {
        c2<T> u[PB], v[PB];
#pragma unroll
        for (int k = 0; k < PB; ++k) {
            const int i = (int)threadIdx.x + (k0 + k) * NT;
            if (i < elems) {
                const int y = by_count.div(i);
                const int64_t at = base + (int64_t)y * p.nx + (i - y * count);
                u[k] = p.in[at];
                v[k] = p.in[at + half];
            }
        }
#pragma unroll
        for (int k = 0; k < PB; ++k) {
            const int i = (int)threadIdx.x + (k0 + k) * NT;
            if (i < elems) {
                const int y = by_count.div(i);
                const int c = i - y * count;
                const c2<T> a = make_c<T>(u[k].x + v[k].x, u[k].y + v[k].y);
                const c2<T> d = make_c<T>(u[k].x - v[k].x, u[k].y - v[k].y);
                c2<T> w = root[y];
                if (inverse) w.y = -w.y;
                if constexpr (sizeof(T) == 4) {
                    c2<T> lo = root[h + y];
                    if (inverse) lo.y = -lo.y;
                    b[k0 + k] = cmul_split(d, w, lo);
                } else {
                    b[k0 + k] = cmul<T>(d, w);
                }
                buf[c * ld + pad_at(y)] = a;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int odd = 0; odd < 2; ++odd) {
        if (odd) {
#pragma unroll
            for (int k = 0; k < V; ++k) {
                const int i = (int)threadIdx.x + k * NT;
                if (i < elems) {
                    const int y = by_count.div(i);
                    buf[(i - y * count) * ld + pad_at(y)] = b[k];
                }
            }
            __syncthreads();
        }
        lds_fft<T, V, NT, CLS>(buf, p.dims, ld, count, s_tw, inverse);
        // outputs k of this half go to rows 2k + odd (PB LDS reads in flight, then PB stores)
        staged<c2<T>, PB, NT>(
            elems,
            [&](int i) {
                const int k = by_count.div(i);
                return buf[(i - k * count) * ld + pad_at(k)];
            },
            [&](int i, c2<T> v) {
                const int k = by_count.div(i);
                p.out[base + (int64_t)(2 * k + odd) * p.nx + (i - k * count)] = v;
            });
        if (!odd) __syncthreads();
    }
}


template <typename T>
static void twiddles(int n, std::vector<c2<T>>& out) {
    out.resize(n);
    for (int m = 0; m < n; ++m) {
        const long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
        out[m] = make_c<T>((T)(double)cosl(a), (T)(double)sinl(a));
    }
}

// FP32 tails of the n-th roots: lo = root - (float)root, rounded.
static void twiddle_tails(int n, std::vector<float2>& out) {
    out.resize(n);
    for (int m = 0; m < n; ++m) {
        const long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
        const long double c = cosl(a), s = sinl(a);
        const float ch = (float)(double)c, sh = (float)(double)s;
        out[m] = make_c<float>((float)(double)(c - (long double)ch), (float)(double)(s - (long double)sh));
    }
}

// k = 2 pi fftfreq(n, d): integer index times 1/(n d), then times 2 pi (periodic_grid.py:66-67).
static void wavenumbers(int n, double d, std::vector<double>& k) {
    k.resize(n);
    const double val = 1.0 / (n * d);
    const int N = (n - 1) / 2 + 1;
    for (int i = 0; i < n; ++i) {
        const int f = i < N ? i : i - n;
        k[i] = 2.0 * 3.141592653589793 * ((double)f * val);
    }
}

template <typename T, int V, int MODE, int RES, int NT>
static int launch_rows_v(const pgpe_fft_plan* plan, const SplitParams<T>& p0, hipStream_t s) {
    constexpr int E = NT * V;
    SplitParams<T> p = p0;
    p.dims = plan->dims_x[radix_class<V>()];
    p.tw = (const c2<T>*)plan->tw_x[radix_class<V>()];
    const int rows_per = E / plan->nx;
    const int total_rows = p.planes * p.ny;
    const dim3 grid((total_rows + rows_per - 1) / rows_per);
    const size_t shm = (size_t)(rows_per * seq_stride<T>(plan->nx) + tw_slots<T>(plan->nx)) * sizeof(c2<T>);
    auto kern = fft_rows_kernel<T, V, MODE, RES, NT>;
    if (shm > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute(fft_rows_kernel)");
    }
    hipLaunchKernelGGL(kern, grid, dim3(NT), shm, s, p);
    PGPE_LAUNCHED("fft_rows_kernel");
    return PGPE_OK;
}

template <typename T, int MODE, int RES>
static int launch_rows(const pgpe_fft_plan* plan, const SplitParams<T>& p, hipStream_t s) {
    switch (plan->row_shape) {
        case ROWS_64x4: return launch_rows_v<T, 4, MODE, RES, 64>(plan, p, s);
        case ROWS_128x4: return launch_rows_v<T, 4, MODE, RES, 128>(plan, p, s);
        case ROWS_256x4: return launch_rows_v<T, 4, MODE, RES, 256>(plan, p, s);
        case ROWS_256x8: return launch_rows_v<T, 8, MODE, RES, 256>(plan, p, s);
        case ROWS_256x12: return launch_rows_v<T, 12, MODE, RES, 256>(plan, p, s);
        default: return launch_rows_v<T, 16, MODE, RES, 256>(plan, p, s);
    }
}

template <typename T, int V, int MODE, int NT>
static int launch_cols_v(const pgpe_fft_plan* plan, const SplitParams<T>& p0, hipStream_t s) {
    constexpr int E = NT * V;
    SplitParams<T> p = p0;
    p.dims = plan->dims_y[radix_class<V>()];
    p.tw = (const c2<T>*)plan->tw_y[radix_class<V>()];
    const int cols_per = E / plan->ny;
    const int groups = (plan->nx + cols_per - 1) / cols_per;
    const size_t shm = (size_t)(cols_per * seq_stride<T>(plan->ny) + tw_slots<T>(plan->ny)) * sizeof(c2<T>);
    auto kern = fft_cols_kernel<T, V, MODE, NT>;
    if (shm > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute(fft_cols_kernel)");
    }
    hipLaunchKernelGGL(kern, dim3(xcd_grid(groups * p.planes)), dim3(NT), shm, s, p);
    PGPE_LAUNCHED("fft_cols_kernel");
    return PGPE_OK;
}


// Column groups: as many columns per group as make 64-byte row segments (8 complex64 or
// 4 complex128) when the batch has enough columns for >= 512 groups (two per CU), fewer on
// small batches (more groups instead); the first shape that holds them, with the padded
// columns and the staged twiddles within the 160 KB of LDS.  Lengths with a radix 3 take
// 256 x 12.  Returns an index into kColShapes, 7 for 256 x 12, or -1.
template <typename T>
static int column_shape(const pgpe_fft_plan* plan, int planes) {
    const int n = plan->ny;
    auto lds = [&](int cols) { return (size_t)(cols * seq_stride<T>(n) + tw_slots<T>(n)) * sizeof(c2<T>); };
    if (n % 3 == 0) return 256 * 12 >= n && lds(256 * 12 / n) <= 163840 ? 7 : -1;
    const int cols_total = plan->nx * planes;
    int want = sizeof(T) == 4 ? 8 : 4;
    while (want > 1 && cols_total / want < 512) want /= 2;
    int best = -1;
    for (int k = 0; k < 7; ++k) {
        const int cols = kColShapes[k][0] * kColShapes[k][1] / n;
        if (cols < 1) continue;
        if (lds(cols) > 163840) break;
        best = k;
        if (cols >= want) break;
    }
    return best;
}

template <typename T, int V, int MODE, int NT, int CLS>
static int launch_cols_half_v(const pgpe_fft_plan* plan, const SplitParams<T>& p0, hipStream_t s) {
    constexpr int E = NT * V;
    SplitParams<T> p = p0;
    const int h = plan->ny / 2;
    p.dims = plan->dims_yh[CLS];
    p.tw = (const c2<T>*)plan->tw_yh[CLS];
    const int cols_per = E / h;
    const int groups = (plan->nx + cols_per - 1) / cols_per;
    const size_t shm = (size_t)(cols_per * seq_stride<T>(h) + tw_slots<T>(h)) * sizeof(c2<T>);
    auto kern = fft_cols_half_kernel<T, V, MODE, NT, CLS>;
    if (shm > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute(fft_cols_half_kernel)");
    }
    hipLaunchKernelGGL(kern, dim3(xcd_grid(groups * p.planes)), dim3(NT), shm, s, p, (const c2<T>*)plan->root_y);
    PGPE_LAUNCHED("fft_cols_half_kernel");
    return PGPE_OK;
}

// Columns split in two (fft_cols_half_kernel) where one-pass groups of 64-byte row segments would
// leave one group per CU: complex64 columns of 2048 (8 per group) or 4096 (4 per group) points
// and complex128 columns of 1024 (4 per group), on batches of >= 512 groups.  The complex64
// groups hold 16 b values per thread under the 128 registers of four waves per SIMD, so their
// half transforms take the radix-8 plan (radix-16 passes would spill).  Sets *done = false when
// the one-pass kernel is to be used.
template <typename T, int MODE>
static int launch_cols_half(const pgpe_fft_plan* plan, const SplitParams<T>& p, hipStream_t s, bool* done) {
    *done = false;
    if constexpr (MODE != COL_FWD && MODE != COL_INV) {
        return PGPE_OK;
    } else {
        if (!plan->root_y) return PGPE_OK;
        const int n = plan->ny;
        const int64_t cols = (int64_t)plan->nx * p.planes;
        if constexpr (sizeof(T) == 4) {
            if ((n == 2048 && cols / 8 >= 512) || (n == 4096 && cols / 4 >= 512)) {
                *done = true;
                return launch_cols_half_v<T, 16, MODE, 512, 1>(plan, p, s);
            }
        } else {
            if (n == 1024 && cols / 4 >= 512) {
                *done = true;
                return launch_cols_half_v<T, 8, MODE, 256, 1>(plan, p, s);
            }
        }
        return PGPE_OK;
    }
}

template <typename T, int MODE>
static int launch_cols(const pgpe_fft_plan* plan, const SplitParams<T>& p, hipStream_t s) {
    bool done;
    const int rc = launch_cols_half<T, MODE>(plan, p, s, &done);
    if (done || rc) return rc;
    switch (column_shape<T>(plan, p.planes)) {
        case 0: return launch_cols_v<T, 4, MODE, 64>(plan, p, s);
        case 1: return launch_cols_v<T, 4, MODE, 128>(plan, p, s);
        case 2: return launch_cols_v<T, 4, MODE, 256>(plan, p, s);
        case 3: return launch_cols_v<T, 8, MODE, 256>(plan, p, s);
        case 4: return launch_cols_v<T, 16, MODE, 256>(plan, p, s);
        case 5: return launch_cols_v<T, 32, MODE, 256>(plan, p, s);
        case 6: return launch_cols_v<T, 32, MODE, 512>(plan, p, s);
        case 7: return launch_cols_v<T, 12, MODE, 256>(plan, p, s);
        default: return fail(PGPE_ERR_UNSUPPORTED, "fft: column too tall for one LDS group");
    }
}

template <typename T>
static int fft2_typed(const pgpe_fft_plan* plan, const void* in, void* out, int planes, int inverse, hipStream_t s) {
    SplitParams<T> p{};
    p.nx = plan->nx;
    p.ny = plan->ny;
    p.planes = planes;
    p.kx = plan->kx;
    p.ky = plan->ky;
    p.scale = (T)(1.0 / ((double)plan->nx * (double)plan->ny));
    int rc;
    if (!inverse) {
        p.in = (c2<T>*)in;
        p.out = (c2<T>*)out;
        if ((rc = launch_rows<T, ROW_FWD, 0>(plan, p, s))) return rc;
        p.in = (c2<T>*)out;
        return launch_cols<T, COL_FWD>(plan, p, s);
    }
    p.in = (c2<T>*)in;
    p.out = (c2<T>*)out;
    if ((rc = launch_cols<T, COL_INV>(plan, p, s))) return rc;
    p.in = (c2<T>*)out;
    return launch_rows<T, ROW_INV, 0>(plan, p, s);
}


template <typename T, int RES>
static int split_typed(const pgpe_split_args* a, hipStream_t s) {
    const pgpe_fft_plan* plan = (const pgpe_fft_plan*)a->plan;
    SplitParams<T> p = split_params<T>(a);
    c2<T>* psi = (c2<T>*)a->psi;
    c2<T>* tmp = (c2<T>*)a->tmp;
    int rc;
    // pass 1 (skipped when the field enters in tmp from a chained step)
    if (!(a->chain & PGPE_CHAIN_IN)) {
        p.in = psi;
        p.out = tmp;
        if (a->ordering == 0) rc = launch_rows<T, ROW_SS_FIRST, RES>(plan, p, s);
        else rc = launch_rows<T, ROW_FWD, RES>(plan, p, s);
        if (rc) return rc;
    }
    // pass 2: propagator along y
    p.in = tmp;
    p.out = tmp;
    if ((rc = launch_cols<T, COL_PROP>(plan, p, s))) return rc;
    // pass 3: local physics between inverse and forward row transforms
    if (a->ordering == 0) rc = launch_rows<T, ROW_SS_MID, RES>(plan, p, s);
    else rc = launch_rows<T, ROW_KO_MID, RES>(plan, p, s);
    if (rc) return rc;
    // pass 4
    if ((rc = launch_cols<T, COL_PROP>(plan, p, s))) return rc;
    // pass 5: inverse rows back into psi (+ second reservoir half-step for the CPU order), or, for
    // a chained step, the join with the next step's first pass, in place in tmp
    if (a->chain & PGPE_CHAIN_OUT) {
        p.in = tmp;
        p.out = tmp;
        return launch_rows<T, ROW_SS_JOIN, RES>(plan, p, s);
    }
    p.in = tmp;
    p.out = psi;
    if (a->ordering == 0) return launch_rows<T, ROW_SS_LAST, RES>(plan, p, s);
    return launch_rows<T, ROW_INV, RES>(plan, p, s);
}

}  // namespace pgpe

using namespace pgpe;

extern "C" {

int pgpe_fft_plan_create(int nx, int ny, double dx, double dy, int precision, pgpe_fft_plan** out) {
    if (!out) return fail(PGPE_ERR_ARG, "pgpe_fft_plan_create: null output");
    *out = nullptr;
    if (precision != 0 && precision != 1) return fail(PGPE_ERR_ARG, "pgpe_fft_plan_create: precision must be 0 or 1");
    int vx = 0, vy = 0;
    if (!check_length(nx, vx) || !check_length(ny, vy))
        return fail(PGPE_ERR_UNSUPPORTED, "pgpe_fft_plan_create: sizes must be 2^a or 3*2^a with 2 <= n <= 3072");
    pgpe_fft_plan* plan = new (std::nothrow) pgpe_fft_plan{};
    if (!plan) return fail(PGPE_ERR_ARG, "pgpe_fft_plan_create: out of host memory");
    plan->nx = nx;
    plan->ny = ny;
    plan->dx = dx;
    plan->dy = dy;
    plan->precision = precision;
    plan->vx = vx;
    plan->vy = vy;
    plan->row_shape = row_shape(nx);
    {
        // a row group holds threads x per_thread / nx padded rows and the twiddles in LDS: very
        // short complex128 rows (6 points) pad past the 160 KB
        int threads, per_thread;
        row_shape_size(plan->row_shape, threads, per_thread);
        const int rows = threads * per_thread / nx;
        const size_t shm = precision ? (size_t)(rows * seq_stride<double>(nx) + tw_slots<double>(nx)) * sizeof(double2)
                                     : (size_t)(rows * seq_stride<float>(nx) + tw_slots<float>(nx)) * sizeof(float2);
        if (shm > 163840) {
            delete plan;
            return fail(PGPE_ERR_UNSUPPORTED, "pgpe_fft_plan_create: rows this short do not fit a row group in this precision");
        }
    }
    // pass lists: x classes 0..2, y classes 0..2, then (tall columns) ny / 2 classes 0..2, in
    // one device block
    const int nh = (ny >= 4 && (ny & (ny - 1)) == 0) ? ny / 2 : 0;
    const int nlists = nh ? 9 : 6;
    std::vector<FftPass> lists[9], all;
    for (int c = 0; c < 3; ++c) {
        build_passes(nx, c, lists[c]);
        build_passes(ny, c, lists[3 + c]);
        if (nh) build_passes(nh, c, lists[6 + c]);
    }
    int offs[9];
    for (int k = 0; k < nlists; ++k) {
        offs[k] = (int)all.size();
        all.insert(all.end(), lists[k].begin(), lists[k].end());
    }
    std::vector<double> kx, ky;
    wavenumbers(nx, dx, kx);
    wavenumbers(ny, dy, ky);
    const size_t csz = precision ? sizeof(double2) : sizeof(float2);
    hipError_t e = hipMalloc((void**)&plan->passes, sizeof(FftPass) * all.size());
    if (e == hipSuccess) e = hipMemcpy(plan->passes, all.data(), sizeof(FftPass) * all.size(), hipMemcpyHostToDevice);
    for (int k = 0; k < nlists && e == hipSuccess; ++k) {
        const int n = k < 3 ? nx : (k < 6 ? ny : nh);
        FftDims& d = k < 3 ? plan->dims_x[k] : (k < 6 ? plan->dims_y[k - 3] : plan->dims_yh[k - 6]);
        d.n = n;
        d.npass = (int)lists[k].size();
        d.passes = plan->passes + offs[k];
        void** tw = k < 3 ? &plan->tw_x[k] : (k < 6 ? &plan->tw_y[k - 3] : &plan->tw_yh[k - 6]);
        const size_t slots = precision ? (size_t)tw_slots<double>(n) : (size_t)tw_slots<float>(n);
        e = hipMalloc(tw, csz * slots);
        if (e != hipSuccess) break;
        if (precision) {
            std::vector<double2> root, tab;
            twiddles<double>(n, root);
            pass_twiddles(n, lists[k], root, tab);
            e = hipMemcpy(*tw, tab.data(), csz * slots, hipMemcpyHostToDevice);
        } else {
            // heads in slots [0, n - 1), a zero pad, then the tails of slots [0, kTwLo - 1)
            std::vector<float2> root, tails, tab, tail_tab;
            twiddles<float>(n, root);
            twiddle_tails(n, tails);
            pass_twiddles(n, lists[k], root, tab);
            pass_twiddles(n, lists[k], tails, tail_tab);
            std::vector<float2> all_slots(slots, make_c<float>(0.0f, 0.0f));
            for (int j = 0; j < n - 1; ++j) all_slots[j] = tab[j];
            for (int j = 0; j < n - 1 && j < kTwLo - 1; ++j) all_slots[n + j] = tail_tab[j];
            e = hipMemcpy(*tw, all_slots.data(), csz * slots, hipMemcpyHostToDevice);
        }
    }
    if (nh && e == hipSuccess) {
        // roots of the splitting step: ny / 2 heads (FP32: then ny / 2 tails)
        if (precision) {
            std::vector<double2> root;
            twiddles<double>(ny, root);
            e = hipMalloc(&plan->root_y, sizeof(double2) * nh);
            if (e == hipSuccess) e = hipMemcpy(plan->root_y, root.data(), sizeof(double2) * nh, hipMemcpyHostToDevice);
        } else {
            std::vector<float2> root, tails;
            twiddles<float>(ny, root);
            twiddle_tails(ny, tails);
            std::vector<float2> both(root.begin(), root.begin() + nh);
            both.insert(both.end(), tails.begin(), tails.begin() + nh);
            e = hipMalloc(&plan->root_y, sizeof(float2) * 2 * nh);
            if (e == hipSuccess) e = hipMemcpy(plan->root_y, both.data(), sizeof(float2) * 2 * nh, hipMemcpyHostToDevice);
        }
    }
    if (e == hipSuccess) e = hipMalloc((void**)&plan->kx, sizeof(double) * nx);
    if (e == hipSuccess) e = hipMalloc((void**)&plan->ky, sizeof(double) * ny);
    if (e == hipSuccess) e = hipMemcpy(plan->kx, kx.data(), sizeof(double) * nx, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(plan->ky, ky.data(), sizeof(double) * ny, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        pgpe_fft_plan_destroy(plan);  // the first error is the one reported
        return hip_fail(e, "pgpe_fft_plan_create");
    }
    *out = plan;
    return PGPE_OK;
}

int pgpe_fft_plan_destroy(pgpe_fft_plan* plan) {
    if (!plan) return PGPE_OK;
    hipError_t e = hipSuccess, f;
    void* ptrs[] = {plan->tw_x[0],  plan->tw_x[1],  plan->tw_x[2],  plan->tw_y[0], plan->tw_y[1],
                    plan->tw_y[2],  plan->tw_yh[0], plan->tw_yh[1], plan->tw_yh[2], plan->root_y,
                    (void*)plan->passes, (void*)plan->kx, (void*)plan->ky};
    for (void* ptr : ptrs)
        if (ptr && (f = hipFree(ptr)) != hipSuccess && e == hipSuccess) e = f;
    delete plan;
    return e == hipSuccess ? PGPE_OK : hip_fail(e, "pgpe_fft_plan_destroy");
}

int pgpe_fft2(const pgpe_fft_plan* plan, const void* in, void* out, int planes, int inverse, void* stream) {
    if (!plan || !in || !out) return fail(PGPE_ERR_ARG, "pgpe_fft2: null argument");
    if (planes <= 0) return fail(PGPE_ERR_ARG, "pgpe_fft2: planes must be positive");
    if (plan->precision == 0) return fft2_typed<float>(plan, in, out, planes, inverse, (hipStream_t)stream);
    return fft2_typed<double>(plan, in, out, planes, inverse, (hipStream_t)stream);
}

int pgpe_fft_columns(const pgpe_fft_plan* plan, const void* in, void* out, int width, int planes, void* stream) {
    if (!plan || !in || !out) return fail(PGPE_ERR_ARG, "pgpe_fft_columns: null argument");
    if (planes <= 0 || width <= 0) return fail(PGPE_ERR_ARG, "pgpe_fft_columns: width and planes must be positive");
    pgpe_fft_plan q = *plan;  // the column kernels read only the y pass list and the width
    q.nx = width;
    if (plan->precision == 0) {
        SplitParams<float> p{};
        p.nx = width;
        p.ny = plan->ny;
        p.planes = planes;
        p.in = (float2*)in;
        p.out = (float2*)out;
        return launch_cols<float, COL_FWD>(&q, p, (hipStream_t)stream);
    }
    SplitParams<double> p{};
    p.nx = width;
    p.ny = plan->ny;
    p.planes = planes;
    p.in = (double2*)in;
    p.out = (double2*)out;
    return launch_cols<double, COL_FWD>(&q, p, (hipStream_t)stream);
}

int pgpe_split_step(const pgpe_split_args* a, void* stream) {
    if (!a || !a->plan || !a->table || !a->psi || !a->tmp || !a->pump || !a->coef)
        return fail(PGPE_ERR_ARG, "pgpe_split_step: null argument");
    if (a->members <= 0) return fail(PGPE_ERR_ARG, "pgpe_split_step: members must be positive");
    if (a->ordering != 0 && a->ordering != 1) return fail(PGPE_ERR_ARG, "pgpe_split_step: ordering must be 0 or 1");
    if (a->components < 0 || a->components > 2 || a->res_components < 0 || a->res_components > 2)
        return fail(PGPE_ERR_ARG, "pgpe_split_step: components and res_components must be 1 or 2");
    if (a->chain & ~(PGPE_CHAIN_IN | PGPE_CHAIN_OUT))
        return fail(PGPE_ERR_ARG, "pgpe_split_step: unknown chain bits");
    if (a->chain && (a->ordering != 0 || a->components == 2))
        return fail(PGPE_ERR_UNSUPPORTED, "pgpe_split_step: chained steps are the scalar CPU operator order only");
    const pgpe_fft_plan* plan = (const pgpe_fft_plan*)a->plan;
    if (plan->precision != a->precision) return fail(PGPE_ERR_ARG, "pgpe_split_step: plan precision mismatch");
    const int nf = a->reservoir == PGPE_RES_NONE ? 0 : (a->reservoir == PGPE_RES_SINGLE ? 1 : 2);
    for (int f = 0; f < nf; ++f)
        if (!a->res[f]) return fail(PGPE_ERR_ARG, "pgpe_split_step: null reservoir field");
    hipStream_t s = (hipStream_t)stream;
    if (a->components == 2) return split_spinor_step(a, s);
    const bool dbl = a->precision == 1;
    switch (a->reservoir) {
        case PGPE_RES_NONE: return dbl ? split_typed<double, 0>(a, s) : split_typed<float, 0>(a, s);
        case PGPE_RES_SINGLE: return dbl ? split_typed<double, 1>(a, s) : split_typed<float, 1>(a, s);
        case PGPE_RES_DOUBLE: return dbl ? split_typed<double, 2>(a, s) : split_typed<float, 2>(a, s);
        case PGPE_RES_QDOUBLE: return dbl ? split_typed<double, 3>(a, s) : split_typed<float, 3>(a, s);
        default: return fail(PGPE_ERR_ARG, "pgpe_split_step: unknown reservoir kind");
    }
}

}  // extern "C"
