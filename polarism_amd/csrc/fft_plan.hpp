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
    int vx, vy;   // elements per thread of the row / column kernels
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


}  // namespace pgpe
