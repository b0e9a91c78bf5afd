// Spinor dispersion of the TE-TM / Zeeman model in k-space (dispersion.py:34-83):
// H(k) = [[E + Z/2, beta (kx - i ky)^2], [beta (kx + i ky)^2, E - Z/2]], E = hbar^2 k^2 / 2m,
// and exp(-i H tau / hbar) by the traceless closed form with sinc.  The Hamiltonian and the
// two angles (E tau / hbar, omega tau / hbar) are formed in double; complex128 fields take
// sin/cos and the matrix in double, complex64 fields reduce the angles to [-pi, pi] in double
// and take sin/cos and the matrix in float (the angles reach ~10^2 rad at the grid's largest k,
// so the reduction keeps the float evaluation at a few ulp).
#pragma once

#include "common.hpp"

namespace pgpe {

// x - 2 pi round(x / 2 pi) in double with 2 pi split into a double and its remainder
// (2 pi = 6.283185307179586 + 2.4492935982947064e-16), returned as float in [-pi, pi].
__device__ __forceinline__ float reduce_two_pi(double x) {
    const double k = rint(x * 0.15915494309189535);  // 1 / (2 pi)
    double r = fma(-k, 6.283185307179586, x);
    r = fma(-k, 2.4492935982947064e-16, r);
    return (float)r;
}

// out = exp(-i H(k) tau / hbar) v for the two components v[0] (psi+), v[1] (psi-).
template <typename T>
__device__ __forceinline__ void spinor_propagate(const pgpe_member& mb, double kxv, double kyv, double tau,
                                                 const c2<T>* v, c2<T>* out) {
    const double hbar = mb.hbar;
    const double k2 = kxv * kxv + kyv * kyv;
    const double E = hbar * hbar * k2 / (2 * mb.m_eff);
    const double z = mb.zeeman / 2;
    // coupling beta (kx - i ky)^2
    const double ur = mb.te_tm * (kxv * kxv - kyv * kyv);
    const double ui = mb.te_tm * (-2.0 * kxv * kyv);
    const double omega = sqrt(z * z + (ur * ur + ui * ui));
    const double t = tau / hbar;
    const double theta = omega * t;
    if constexpr (sizeof(T) == 4) {
        float ps, pc, st, ct;
        sincosf(reduce_two_pi(-E * t), &ps, &pc);  // exp(-i eps t)
        sincosf(reduce_two_pi(theta), &st, &ct);
        const float thf = (float)theta;
        const float so = thf != 0.0f ? (float)t * (st / thf) : (float)t;  // t sinc(theta / pi)
        const float zf = (float)z, urf = (float)ur, uif = (float)ui;
        const float m00r = pc * ct + ps * so * zf, m00i = ps * ct - pc * so * zf;
        const float m11r = pc * ct - ps * so * zf, m11i = ps * ct + pc * so * zf;
        const float m01r = so * (ps * urf + pc * uif), m01i = so * (ps * uif - pc * urf);
        const float m10r = so * (ps * urf - pc * uif), m10i = so * (-ps * uif - pc * urf);
        const float ar = v[0].x, ai = v[0].y, br = v[1].x, bi = v[1].y;
        out[0] = make_c<T>(m00r * ar - m00i * ai + m01r * br - m01i * bi, m00r * ai + m00i * ar + m01r * bi + m01i * br);
        out[1] = make_c<T>(m10r * ar - m10i * ai + m11r * br - m11i * bi, m10r * ai + m10i * ar + m11r * bi + m11i * br);
        return;
    } else {
        double ps, pc;
        sincos(-E * t, &ps, &pc);                      // exp(-i eps t)
        double st, ct;
        sincos(theta, &st, &ct);
        const double so = theta != 0.0 ? t * (st / theta) : t;  // t sinc(theta / pi)
        // M00 = ph (cos - i so z), M11 = ph (cos + i so z), M01 = -i ph so u, M10 = -i ph so conj(u)
        const double m00r = pc * ct + ps * so * z, m00i = ps * ct - pc * so * z;
        const double m11r = pc * ct - ps * so * z, m11i = ps * ct + pc * so * z;
        // -i ph so u = so * (-i)(pc + i ps)(ur + i ui) = so * [(ps ur + pc ui) + i (ps ui - pc ur)]
        const double m01r = so * (ps * ur + pc * ui), m01i = so * (ps * ui - pc * ur);
        const double m10r = so * (ps * ur - pc * ui), m10i = so * (-ps * ui - pc * ur);
        const double ar = v[0].x, ai = v[0].y, br = v[1].x, bi = v[1].y;
        out[0] = make_c<T>((T)(m00r * ar - m00i * ai + m01r * br - m01i * bi),
                           (T)(m00r * ai + m00i * ar + m01r * bi + m01i * br));
        out[1] = make_c<T>((T)(m10r * ar - m10i * ai + m11r * br - m11i * bi),
                           (T)(m10r * ai + m10i * ar + m11r * bi + m11i * br));
    }
}

}  // namespace pgpe
