// Plan object shared by the FFT, split-step and Lawson kernels.
#pragma once

#include "fft_core.hpp"

struct pgpe_fft_plan {
    int nx, ny;
    int precision;
    double dx, dy;
    pgpe::FftDims dx_dims, dy_dims;
    void* tw_x;   // n-th roots of unity for nx (precision-typed c2)
    void* tw_y;
    double* kx;   // 2 pi fftfreq(nx, dx)
    double* ky;
    int vx, vy;   // elements per thread of the Lawson row / column kernels (16, or 12 with a radix 3)
    int rv;       // elements per thread of the FFT / split-step row kernels: the fewest that hold
                  // one row (more workgroups on small grids); column groups are sized per launch
};

namespace pgpe {

// Radix plan of n = 2^a or 3 * 2^a: radix-4 passes, one radix-2 if a is odd,
// one radix-3; v = elements per thread (16, or 12 when 3 divides n).
inline bool factor(int n, FftDims& d, int& v) {
    d.n = n;
    d.npass = 0;
    int m = n, threes = 0;
    while (m % 3 == 0) { m /= 3; ++threes; }
    if (threes > 1) return false;
    int twos = 0;
    while (m % 2 == 0) { m /= 2; ++twos; }
    if (m != 1 || n < 2) return false;
    for (int i = 0; i < twos / 2; ++i) d.radix[d.npass++] = 4;
    if (twos % 2) d.radix[d.npass++] = 2;
    if (threes) d.radix[d.npass++] = 3;
    v = threes ? 12 : 16;
    return n <= 256 * v && d.npass <= kFftMaxPasses;
}

// Elements per thread of a 256-thread group holding `need` elements: 4, 8, 16 (power-of-two
// lengths; every pass needs V divisible by its radix) or 12 (lengths with a radix 3).
inline int small_v(int n, int need) {
    if (n % 3 == 0) return 12;
    for (int v = 4; v <= 16; v *= 2)
        if (need <= 256 * v) return v;
    return 16;
}


}  // namespace pgpe
