// Plan object shared by the FFT, split-step and Lawson kernels.
#pragma once

#include <vector>

#include "fft_core.hpp"

struct pgpe_fft_plan {
    int nx, ny;
    int precision;
    double dx, dy;
    // per radix class (radix_class<V>(): radices up to 4, 8, 16): the pass list and the
    // per-pass twiddle table (n - 1 entries, precision-typed c2) of each dimension
    pgpe::FftDims dims_x[3], dims_y[3];
    void* tw_x[3];
    void* tw_y[3];
    pgpe::FftPass* passes;  // device block holding the pass lists
    double* kx;             // 2 pi fftfreq(nx, dx)
    double* ky;
    // Tall columns split in two (ny a power of two, ny >= 4): the pass lists and per-pass
    // twiddles of the ny / 2-point transform per radix class, and the ny-th roots
    // exp(-2 pi i y / ny), y < ny / 2, of the splitting step (FP32: the heads, then their tails).
    pgpe::FftDims dims_yh[3];
    void* tw_yh[3];
    void* root_y;
    int vx, vy;   // elements per thread of the Lawson row / column kernels (16, or 12 with a radix 3)
    int row_shape;  // threads x elements per thread of the FFT / split-step row kernels
                    // (pgpe::RowShape); column groups are sized per launch
};

namespace pgpe {

// n = 2^a or 3 * 2^a with 2 <= n <= 256 v, v = elements per thread of the Lawson kernels
// (16, or 12 when 3 divides n).
inline bool check_length(int n, int& v) {
    int m = n, threes = 0;
    while (m % 3 == 0) { m /= 3; ++threes; }
    if (threes > 1 || n < 2) return false;
    while (m % 2 == 0) m /= 2;
    if (m != 1) return false;
    v = threes ? 12 : 16;
    return n <= 256 * v;
}

// Pass list of an n-point transform in radix class cls: the powers of two with exponents
// up to cls + 2, balanced over ceil(a / (cls + 2)) passes, then radix 3 when 3 divides n.
inline void build_passes(int n, int cls, std::vector<FftPass>& out) {
    out.clear();
    const int emax = cls + 2;
    int a = 0;
    while (((n >> a) & 1) == 0) ++a;
    int left = (a + emax - 1) / emax;
    int ns = 1;
    while (a > 0) {
        const int e = (a + left - 1) / left;
        const int r = 1 << e;
        out.push_back(FftPass{r, ns, div_magic((uint32_t)(n / r)), div_magic((uint32_t)ns)});
        ns *= r;
        a -= e;
        --left;
    }
    if (n % 3 == 0) out.push_back(FftPass{3, ns, div_magic((uint32_t)(n / 3)), div_magic((uint32_t)ns)});
}

// Per-pass twiddle table from the n-th roots of unity root[k] = exp(-2 pi i k / n):
// slot ns - 1 + jm (r - 1) + q - 1 of pass (r, ns) holds root[q jm n / (ns r)].
template <typename C>
inline void pass_twiddles(int n, const std::vector<FftPass>& passes, const std::vector<C>& root, std::vector<C>& out) {
    out.assign(n > 1 ? n - 1 : 1, root[0]);
    for (const FftPass& ps : passes) {
        const int stride = n / (ps.ns * ps.radix);
        for (int jm = 0; jm < ps.ns; ++jm)
            for (int q = 1; q < ps.radix; ++q) out[ps.ns - 1 + jm * (ps.radix - 1) + q - 1] = root[q * jm * stride];
    }
}

// Row group shapes (threads x elements per thread).
enum RowShape { ROWS_64x4, ROWS_128x4, ROWS_256x4, ROWS_256x8, ROWS_256x12, ROWS_256x16 };

// One row per group at 4 elements per thread on 64, 128 or 256 threads (n <= 1024: more,
// smaller workgroups on small grids, several rows per group below 256), then 16 elements per
// thread (two 2048-point rows per group: the staged twiddles serve both, and the radix-16 plan
// takes three passes); lengths with a radix 3 take 256 x 12.
inline int row_shape(int n) {
    if (n % 3 == 0) return ROWS_256x12;
    if (n <= 256) return ROWS_64x4;
    if (n <= 512) return ROWS_128x4;
    if (n <= 1024) return ROWS_256x4;
    return ROWS_256x16;
}

// Threads x elements per thread of a row shape.
inline void row_shape_size(int shape, int& threads, int& per_thread) {
    static const int sizes[][2] = {{64, 4}, {128, 4}, {256, 4}, {256, 8}, {256, 12}, {256, 16}};
    threads = sizes[shape][0];
    per_thread = sizes[shape][1];
}

}  // namespace pgpe
