// In-LDS mixed-radix (2, 3, 4) Stockham FFT used by the row and column kernels.
//
// A workgroup of 256 threads transforms `count` sequences of length n that
// sit contiguously in LDS (buf[s * n + k]).  Each pass is the Stockham
// autosort iteration: butterfly j of radix r reads r inputs n/r apart,
// multiplies input q by w^(q (j mod Ns)) with w = exp(-+2 pi i / (Ns r)), runs an
// r-point DFT and writes the outputs Ns apart at (j / Ns) Ns r + (j mod Ns).
// Twiddles come from a host-computed table of the n-th roots of unity
// (float64 cos/sin rounded once), so the FP64 transform keeps ~1e-16
// relative accuracy.  The LDS buffer is single: every thread gathers its
// butterflies' inputs into registers, a barrier, then all scatter.
#pragma once

#include "common.hpp"

namespace pgpe {

constexpr int kFftThreads = 256;
constexpr int kFftMaxPasses = 8;

struct FftDims {
    int n;                          // transform length
    int npass;
    int radix[kFftMaxPasses];       // 2, 3 or 4
};

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
        const T h = (T)0.86602540378443864676372317075294;  // sqrt(3)/2
        const c2<T> a = v[0], b = v[1], c = v[2];
        const c2<T> t1 = make_c<T>(b.x + c.x, b.y + c.y);
        const c2<T> t2 = make_c<T>(a.x - (T)0.5 * t1.x, a.y - (T)0.5 * t1.y);
        const c2<T> d = make_c<T>(h * (b.x - c.x), h * (b.y - c.y));
        const c2<T> t3 = rot90<T>(d, inverse);  // -+i sqrt(3)/2 (b - c)
        v[0] = make_c<T>(a.x + t1.x, a.y + t1.y);
        v[1] = make_c<T>(t2.x + t3.x, t2.y + t3.y);
        v[2] = make_c<T>(t2.x - t3.x, t2.y - t3.y);
    }
}

// One Stockham pass over `count` sequences; V = elements per thread.
template <typename T, int R, int V>
__device__ __forceinline__ void stockham_pass(c2<T>* buf, int n, int ld, int count, int Ns,
                                              const c2<T>* __restrict__ tw, bool inverse) {
    constexpr int B = V / R;
    const int nb = n / R;
    const int total = nb * count;
    const int tstride = n / (Ns * R);  // twiddle index step in the n-th root table
    c2<T> v[B][R];
    int dst[B];
#pragma unroll
    for (int i = 0; i < B; ++i) {
        const int b = threadIdx.x + i * kFftThreads;
        dst[i] = -1;
        if (b < total) {
            const int s = b / nb;
            const int j = b - s * nb;
            const int base = s * ld;
            const int jm = j % Ns;
#pragma unroll
            for (int q = 0; q < R; ++q) {
                c2<T> x = buf[base + j + q * nb];
                if (q > 0 && jm > 0) {
                    c2<T> w = tw[q * jm * tstride];
                    if (inverse) w.y = -w.y;
                    x = cmul<T>(x, w);
                }
                v[i][q] = x;
            }
            dft_small<T, R>(v[i], inverse);
            dst[i] = base + (j / Ns) * Ns * R + jm;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < B; ++i) {
        if (dst[i] >= 0) {
#pragma unroll
            for (int q = 0; q < R; ++q) buf[dst[i] + q * Ns] = v[i][q];
        }
    }
    __syncthreads();
}

// Sequence s occupies buf[s * ld + k], k < n (ld >= n; padding avoids LDS bank conflicts).
template <typename T, int V>
__device__ void lds_fft(c2<T>* buf, const FftDims& d, int ld, int count, const c2<T>* __restrict__ tw,
                        bool inverse) {
    int Ns = 1;
    for (int p = 0; p < d.npass; ++p) {
        const int r = d.radix[p];
        if (r == 4) {
            if constexpr (V % 4 == 0) stockham_pass<T, 4, V>(buf, d.n, ld, count, Ns, tw, inverse);
        } else if (r == 2) {
            stockham_pass<T, 2, V>(buf, d.n, ld, count, Ns, tw, inverse);
        } else {
            if constexpr (V % 3 == 0) stockham_pass<T, 3, V>(buf, d.n, ld, count, Ns, tw, inverse);
        }
        Ns *= r;
    }
}

}  // namespace pgpe
