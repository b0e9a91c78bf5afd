// In-LDS mixed-radix (2, 3, 4, 8, 16) Stockham FFT used by the row and column kernels.
//
// A workgroup of NT threads (256 unless a launcher picks 64, 128 or 512) transforms
// `count` sequences of length n that sit in LDS.  Each pass is the Stockham autosort iteration: butterfly j of
// radix r reads r inputs n/r apart, multiplies input q by w^(q (j mod Ns)) with
// w = exp(-+2 pi i / (Ns r)), runs an r-point DFT in registers and writes the
// outputs Ns apart at (j / Ns) Ns r + (j mod Ns).  The pass list comes from the
// plan (built on the host for three radix classes): the largest radices the
// per-thread element count V allows (16 when V is a multiple of 16, else 8 or
// 4), balanced so that n = 2^a takes ceil(a / 4) passes at V = 16 or 32 (a
// 2048-point transform: 16, 16, 8), then one radix-3 pass when 3 divides n.
// Twiddles come from host-computed n-th roots of unity (long double cos/sin
// rounded once), so the FP64 transform keeps ~1e-16 relative accuracy; the
// constants inside the radix-3/8/16 DFTs are sqrt(3)/2, cos(pi/8), sin(pi/8) and
// 1/sqrt(2), which FP32 takes as a head and a tail (see kmul).
//
// Layout: element k of a sequence sits at k + k/16 (one pad slot per 16
// elements: a stride-16 scatter of the first radix-16 pass then spreads over
// the LDS banks) and sequences are seq_stride(n) apart, chosen so that
// consecutive sequences start 64 bytes apart modulo 256 (the column kernels'
// transposing loads hit distinct banks).  The buffer is single: every thread
// gathers its butterflies' inputs into registers, a barrier, then all scatter.
#pragma once

#include "common.hpp"

namespace pgpe {

constexpr int kFftThreads = 256;
constexpr int kFftMaxPasses = 8;

// One Stockham pass of a plan: its radix, the product ns of the earlier radices and the
// division magics of n / radix and ns (FastDiv; 0 when the divisor is 1).
struct FftPass {
    int radix;
    int ns;
    uint32_t nb_m;
    uint32_t ns_m;
};

// A transform length with its pass list (device memory, read through the scalar cache).
struct FftDims {
    int n;
    int npass;
    const FftPass* passes;
};

// Loads then stores `elems` items strided over the group's threads, CH at a time: a thread
// issues its CH loads before it waits on the first, so the group keeps CH x 256 loads in
// flight (one at a time per thread when load and store alternate, which leaves the
// few-workgroup column passes bound by memory latency).  load(i) -> P, store(i, P).
template <typename P, int CH, int NT = kFftThreads, typename Load, typename Store>
__device__ __forceinline__ void staged(int elems, Load load, Store store) {
    for (int i0 = 0; i0 < elems; i0 += NT * CH) {
        P r[CH];
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            const int i = i0 + k * NT + (int)threadIdx.x;
            if (i < elems) r[k] = load(i);
        }
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            const int i = i0 + k * NT + (int)threadIdx.x;
            if (i < elems) store(i, r[k]);
        }
    }
}

// Loads in flight per thread for V elements per thread: all of them up to 64 VGPRs of data.
template <typename T, int V>
constexpr int stage_chunk() {
    constexpr int cap = sizeof(T) == 4 ? 32 : 16;
    return V < cap ? V : cap;
}

// An irrational constant K as head + tail: hi = K rounded to T, lo = (K - hi) rounded (0 in
// FP64).  FP32 multiplies by both, so a product by K is unbiased.  With the head alone,
// |1/sqrt(2) (1 - i)|^2 = 2 hi^2 = 1 - 3.4e-8 and cos^2 + sin^2 of pi/8 = 1 - 5.7e-8: every
// radix-8 pass would shrink the energy by ~1e-8, a loss that compounds over repeated
// transforms (240 2-D transforms of a 512^2 field lost 1.1e-5 of the norm).
template <typename T>
struct KSplit {
    T hi, lo;
};

template <typename T>
__device__ __forceinline__ constexpr KSplit<T> ksplit(double k) {
    return KSplit<T>{(T)k, (T)(k - (double)(T)k)};
}

// x K
template <typename T>
__device__ __forceinline__ T kmul(T x, KSplit<T> k) {
    if constexpr (sizeof(T) == 4) return __builtin_fmaf(x, k.hi, x * k.lo);
    else return x * k.hi;
}

// a x + b y
template <typename T>
__device__ __forceinline__ T kdot(KSplit<T> a, T x, KSplit<T> b, T y) {
    if constexpr (sizeof(T) == 4) return __builtin_fmaf(a.hi, x, __builtin_fmaf(b.hi, y, a.lo * x + b.lo * y));
    else return a.hi * x + b.hi * y;
}

template <typename T>
__device__ __forceinline__ KSplit<T> kneg(KSplit<T> k) {
    return KSplit<T>{-k.hi, -k.lo};
}

template <typename T>
__device__ __forceinline__ c2<T> cmul(c2<T> a, c2<T> b) {
    return make_c<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Multiply by -i (forward) or +i (inverse).
template <typename T>
__device__ __forceinline__ c2<T> rot90(c2<T> a, bool inverse) {
    return inverse ? make_c<T>(-a.y, a.x) : make_c<T>(a.y, -a.x);
}

template <typename T, int R>
__device__ __forceinline__ void dft_small(c2<T>* v, bool inverse) {
    if constexpr (R == 2) {
        const c2<T> a = v[0], b = v[1];
        v[0] = make_c<T>(a.x + b.x, a.y + b.y);
        v[1] = make_c<T>(a.x - b.x, a.y - b.y);
    } else if constexpr (R == 4) {
        const c2<T> a = v[0], b = v[1], c = v[2], d = v[3];
        const c2<T> s0 = make_c<T>(a.x + c.x, a.y + c.y);
        const c2<T> s1 = make_c<T>(a.x - c.x, a.y - c.y);
        const c2<T> s2 = make_c<T>(b.x + d.x, b.y + d.y);
        const c2<T> s3 = rot90<T>(make_c<T>(b.x - d.x, b.y - d.y), inverse);  // -+i (b - d)
        v[0] = make_c<T>(s0.x + s2.x, s0.y + s2.y);
        v[1] = make_c<T>(s1.x + s3.x, s1.y + s3.y);
        v[2] = make_c<T>(s0.x - s2.x, s0.y - s2.y);
        v[3] = make_c<T>(s1.x - s3.x, s1.y - s3.y);
    } else {  // R == 3
        constexpr KSplit<T> h = ksplit<T>(0.86602540378443864676372317075294);  // sqrt(3)/2
        const c2<T> a = v[0], b = v[1], c = v[2];
        const c2<T> t1 = make_c<T>(b.x + c.x, b.y + c.y);
        const c2<T> t2 = make_c<T>(a.x - (T)0.5 * t1.x, a.y - (T)0.5 * t1.y);
        const c2<T> d = make_c<T>(kmul<T>(b.x - c.x, h), kmul<T>(b.y - c.y, h));
        const c2<T> t3 = rot90<T>(d, inverse);  // -+i sqrt(3)/2 (b - c)
        v[0] = make_c<T>(a.x + t1.x, a.y + t1.y);
        v[1] = make_c<T>(t2.x + t3.x, t2.y + t3.y);
        v[2] = make_c<T>(t2.x - t3.x, t2.y - t3.y);
    }
}

// Unsigned division by a divisor fixed for a pass: q = floor(x m / 2^32) with
// m = floor((2^32 - 1) / d) + 1 (one 32-bit division to set up, one mul_hi per use), exact
// whenever x * d < 2^32 (here x < 2^16, d <= 4096); m = 0 marks d = 1.
__host__ __device__ __forceinline__ uint32_t div_magic(uint32_t d) { return d == 1u ? 0u : 0xFFFFFFFFu / d + 1u; }

struct FastDiv {
    uint32_t m;
    __device__ __forceinline__ explicit FastDiv(uint32_t d) : m(div_magic(d)) {}
    __device__ __forceinline__ static FastDiv from_magic(uint32_t magic) {
        FastDiv f(1u);
        f.m = magic;
        return f;
    }
    __device__ __forceinline__ int div(int x) const { return m ? (int)__umulhi((uint32_t)x, m) : x; }
};

// Padded position of element k inside its sequence.
__host__ __device__ __forceinline__ constexpr int pad_at(int k) { return k + (k >> 4); }

// Distance between consecutive sequences: holds pad_at(n - 1) + 1 slots and starts the next
// sequence 64 bytes further modulo 256.
template <typename T>
__host__ __device__ __forceinline__ constexpr int seq_stride(int n) {
    const int p = n + (n >> 4);
    const int unit = 256 / (int)sizeof(c2<T>);
    const int off = 64 / (int)sizeof(c2<T>);
    return ((p - off + unit - 1) / unit) * unit + off;
}

// x * exp(-+2 pi i / 8): (1 -+ i) / sqrt(2).
template <typename T>
__device__ __forceinline__ c2<T> mul_w8(c2<T> a, bool inverse) {
    constexpr KSplit<T> h = ksplit<T>(0.70710678118654752440084436210485);
    return inverse ? make_c<T>(kmul<T>(a.x - a.y, h), kmul<T>(a.y + a.x, h))
                   : make_c<T>(kmul<T>(a.x + a.y, h), kmul<T>(a.y - a.x, h));
}

// x * exp(-+2 pi i m / 16) for the products m = n2 k1 of a 4 x 4 split.
template <typename T>
__device__ __forceinline__ c2<T> mul_w16(c2<T> a, int m, bool inverse) {
    constexpr KSplit<T> c = ksplit<T>(0.92387953251128675612818318939679);  // cos(pi / 8)
    constexpr KSplit<T> s = ksplit<T>(0.38268343236508977172845998403040);  // sin(pi / 8)
    const KSplit<T> si = inverse ? s : kneg<T>(s);
    const KSplit<T> ci = inverse ? c : kneg<T>(c);
    switch (m) {
        case 0: return a;
        case 1: return make_c<T>(kdot<T>(c, a.x, kneg<T>(si), a.y), kdot<T>(c, a.y, si, a.x));  // c -+ i s
        case 2: return mul_w8<T>(a, inverse);
        case 3: return make_c<T>(kdot<T>(s, a.x, kneg<T>(ci), a.y), kdot<T>(s, a.y, ci, a.x));  // s -+ i c
        case 4: return rot90<T>(a, inverse);
        case 6: return rot90<T>(mul_w8<T>(a, inverse), inverse);
        default:  // m = 9: -(c -+ i s)
            return make_c<T>(-kdot<T>(c, a.x, kneg<T>(si), a.y), -kdot<T>(c, a.y, si, a.x));
    }
}

template <typename T>
__device__ __forceinline__ void dft4_of(c2<T>& a, c2<T>& b, c2<T>& c, c2<T>& d, bool inverse) {
    c2<T> v[4] = {a, b, c, d};
    dft_small<T, 4>(v, inverse);
    a = v[0];
    b = v[1];
    c = v[2];
    d = v[3];
}

// 8-point DFT: one radix-2 split, the odd half rotated by w8^k, two 4-point DFTs.
template <typename T>
__device__ __forceinline__ void dft8(c2<T>* v, bool inverse) {
    c2<T> u[4], w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        u[k] = make_c<T>(v[k].x + v[k + 4].x, v[k].y + v[k + 4].y);
        w[k] = make_c<T>(v[k].x - v[k + 4].x, v[k].y - v[k + 4].y);
    }
    w[1] = mul_w8<T>(w[1], inverse);
    w[2] = rot90<T>(w[2], inverse);
    w[3] = rot90<T>(mul_w8<T>(w[3], inverse), inverse);
    dft4_of<T>(u[0], u[1], u[2], u[3], inverse);
    dft4_of<T>(w[0], w[1], w[2], w[3], inverse);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        v[2 * m] = u[m];
        v[2 * m + 1] = w[m];
    }
}

// 16-point DFT as 4 x 4: 4-point DFTs over n1 of x[4 n1 + n2], twiddles w16^(n2 k1),
// 4-point DFTs over n2; output k1 + 4 k2.
template <typename T>
__device__ __forceinline__ void dft16(c2<T>* v, bool inverse) {
    c2<T> y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        y[n2][0] = v[n2];
        y[n2][1] = v[4 + n2];
        y[n2][2] = v[8 + n2];
        y[n2][3] = v[12 + n2];
        dft4_of<T>(y[n2][0], y[n2][1], y[n2][2], y[n2][3], inverse);
#pragma unroll
        for (int k1 = 1; k1 < 4; ++k1) y[n2][k1] = mul_w16<T>(y[n2][k1], n2 * k1, inverse);
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        dft4_of<T>(y[0][k1], y[1][k1], y[2][k1], y[3][k1], inverse);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = y[k2][k1];
    }
}

template <typename T, int R>
__device__ __forceinline__ void dft_radix(c2<T>* v, bool inverse) {
    if constexpr (R == 16) dft16<T>(v, inverse);
    else if constexpr (R == 8) dft8<T>(v, inverse);
    else dft_small<T, R>(v, inverse);
}

// FP32 twiddle tails.  A twiddle rounded to FP32 has |w|^2 - 1 of up to 1e-7, and over the
// few distinct roots of the early passes these errors do not average out (the 16th roots lose
// 2.3e-8 of the energy per element and pass).  That makes every FP32 2-D transform shrink the
// energy by ~8e-8, which compounds over repeated transforms.  FP32 tables therefore carry a
// tail lo = w - hi for the slots below kTwLo - 1 (the passes with ns R <= kTwLo), and
// those passes multiply by hi + lo.  The later passes use 512th or finer roots, whose rounding
// errors average to ~1e-9 per pass.
constexpr int kTwLo = 256;

// Table slots of an n-point plan (device and LDS alike): FP64, the n - 1 twiddles; FP32, the
// n - 1 heads, one pad slot, then kTwLo tails (tail of slot j at n + j, zero past n - 2).
template <typename T>
__host__ __device__ __forceinline__ constexpr int tw_slots(int n) {
    return sizeof(T) == 4 ? n + kTwLo : (n > 1 ? n - 1 : 1);
}

// Stages a plan's twiddles into LDS (the per-pass table, see the plan builder): pass p
// (radix R, ns = product of the earlier radices) finds w^(q jm n / (ns R)) at
// ns - 1 + jm (R - 1) + q - 1 for jm < ns, 1 <= q < R, so a butterfly reads its R - 1
// factors from consecutive slots and neighbouring butterflies from neighbouring runs.
// Callers put a barrier between this and lds_fft.
template <typename T, int NT = kFftThreads>
__device__ __forceinline__ void stage_twiddles(c2<T>* s_tw, const c2<T>* __restrict__ tw, int n) {
    staged<c2<T>, (sizeof(T) == 4 ? 16 : 8), NT>(tw_slots<T>(n), [&](int i) { return tw[i]; },
                                                 [&](int i, c2<T> w) { s_tw[i] = w; });
}

// x (hi + lo) for an FP32 twiddle with its tail.
__device__ __forceinline__ c2<float> cmul_split(c2<float> x, c2<float> hi, c2<float> lo) {
    const float re = __builtin_fmaf(x.x, hi.x, __builtin_fmaf(-x.y, hi.y, __builtin_fmaf(x.x, lo.x, -(x.y * lo.y))));
    const float im = __builtin_fmaf(x.x, hi.y, __builtin_fmaf(x.y, hi.x, __builtin_fmaf(x.x, lo.y, x.y * lo.x)));
    return make_c<float>(re, im);
}

// One Stockham pass over `count` sequences ld apart by a group of NT threads; V = elements
// per thread; tw = this pass's staged twiddles.
template <typename T, int R, int V, int NT = kFftThreads>
__device__ __forceinline__ void stockham_pass(c2<T>* buf, int n, int ld, int count, const FftPass& ps,
                                              const c2<T>* __restrict__ tw, bool inverse) {
    constexpr int B = V / R;
    const int nb = n / R;
    const int Ns = ps.ns;
    const int total = nb * count;
    const bool tails = sizeof(T) == 4 && Ns * R <= kTwLo;  // this pass's FP32 slots carry tails
    const FastDiv by_nb = FastDiv::from_magic(ps.nb_m), by_ns = FastDiv::from_magic(ps.ns_m);
    c2<T> v[B][R];
    int base[B], k0[B];
#pragma unroll
    for (int i = 0; i < B; ++i) {
        const int b = threadIdx.x + i * NT;
        base[i] = -1;
        if (b < total) {
            const int s = by_nb.div(b);
            const int j = b - s * nb;
            const int jq = by_ns.div(j);
            const int jm = j - jq * Ns;
            base[i] = s * ld;
#pragma unroll
            for (int q = 0; q < R; ++q) {
                c2<T> x = buf[base[i] + pad_at(j + q * nb)];
                if (q > 0 && jm > 0) {
                    const int at = jm * (R - 1) + q - 1;
                    c2<T> w = tw[at];
                    if (inverse) w.y = -w.y;
                    if constexpr (sizeof(T) == 4) {
                        if (tails) {
                            c2<T> lo = tw[n + at];
                            if (inverse) lo.y = -lo.y;
                            x = cmul_split(x, w, lo);
                        } else {
                            x = cmul<T>(x, w);
                        }
                    } else {
                        x = cmul<T>(x, w);
                    }
                }
                v[i][q] = x;
            }
            dft_radix<T, R>(v[i], inverse);
            k0[i] = jq * Ns * R + jm;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < B; ++i) {
        if (base[i] >= 0) {
#pragma unroll
            for (int q = 0; q < R; ++q) buf[base[i] + pad_at(k0[i] + q * Ns)] = v[i][q];
        }
    }
    __syncthreads();
}

// Sequence s occupies buf[s * ld + pad_at(k)], k < n, ld >= pad_at(n - 1) + 1; tw holds the
// staged twiddles of the plan d, whose radices divide V (the launchers pick the plan by V).
// CLS is the plan's radix class (radix_class below by default); passes of larger radices are not
// compiled in, so a kernel short of registers can run a class-1 plan at V = 16.
template <typename T, int V, int NT = kFftThreads, int CLS = ((V % 16 == 0 && V < 64) ? 2 : (V % 8 == 0 ? 1 : 0))>
__device__ __forceinline__ void lds_fft(c2<T>* buf, const FftDims& d, int ld, int count,
                                        const c2<T>* __restrict__ tw, bool inverse) {
    for (int p = 0; p < d.npass; ++p) {
        const FftPass ps = d.passes[p];
        const c2<T>* t = tw + (ps.ns - 1);
        if (ps.radix == 16) {
            if constexpr (CLS >= 2 && V % 16 == 0 && V < 64) stockham_pass<T, 16, V, NT>(buf, d.n, ld, count, ps, t, inverse);
        } else if (ps.radix == 8) {
            if constexpr (CLS >= 1 && V % 8 == 0) stockham_pass<T, 8, V, NT>(buf, d.n, ld, count, ps, t, inverse);
        } else if (ps.radix == 4) {
            if constexpr (V % 4 == 0) stockham_pass<T, 4, V, NT>(buf, d.n, ld, count, ps, t, inverse);
        } else if (ps.radix == 2) {
            // V = 64 serves only columns of 1024 or more (the launchers guarantee it), whose
            // plans hold radices 8 and 4: 32 radix-2 butterflies per thread would spill registers
            if constexpr (V < 64) stockham_pass<T, 2, V, NT>(buf, d.n, ld, count, ps, t, inverse);
        } else {
            if constexpr (V % 3 == 0) stockham_pass<T, 3, V, NT>(buf, d.n, ld, count, ps, t, inverse);
        }
    }
}

// Plan class of a kernel with V elements per thread: 0 = radices up to 4, 1 = up to 8,
// 2 = up to 16 (radix 16 at V = 64 would hold 4 x 16 inputs and their twiddles at once,
// past the 512 VGPRs, so V = 64 takes class 1).
template <int V>
constexpr int radix_class() {
    return (V % 16 == 0 && V < 64) ? 2 : (V % 8 == 0 ? 1 : 0);
}

}  // namespace pgpe
