// Spinor dispersion of the TE-TM / Zeeman model in k-space (dispersion.py:34-83):
// H(k) = [[E + Z/2, beta (kx - i ky)^2], [beta (kx + i ky)^2, E - Z/2]], E = hbar^2 k^2 / 2m,
// and exp(-i H tau / hbar) by the traceless closed form with sinc, evaluated in double.
#pragma once

#include "common.hpp"

namespace pgpe {

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

}  // namespace pgpe
