// Test infrastructure only: real spherical harmonics in the tournier basis for the CPU oracle --
// the associated Legendre table, the series' derivatives on the sphere and the Newton climb to a
// peak -- shared by the peak search (sd_oracle.cpp) and streamline tracking (track_oracle.cpp).
//
// Restated from the reference (paths relative to the reference repository):
//   cpp/src/sh/harmonics.cpp   coefficient_index, power_int, RecurrenceConstants, fill_legendre,
//                              Legendre::d_theta / d2_theta, derivatives
//   cpp/src/sh/peaks.cpp       climb
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>

namespace oracle_sh {

using Index = std::int64_t;

constexpr int SH_MAX_LMAX = 20;
using LegendreTable = std::array<std::array<double, SH_MAX_LMAX + 1>, SH_MAX_LMAX + 1>;

// Position of (degree, order) in a series (harmonics.hpp coefficient_index).
inline Index sh_index(Index degree, Index order, bool full) {
    return full ? degree * degree + degree + order : degree * (degree + 1) / 2 + order;
}

// base^exponent by binary exponentiation (harmonics.cpp power_int).
inline double power_int(double base, Index exponent) {
    double result = 1.0;
    while (true) {
        if (exponent % 2 == 1) {
            result = result * base;
        }
        exponent /= 2;
        if (exponent == 0) {
            break;
        }
        base = base * base;
    }
    return result;
}

inline double ladder_up(Index degree, Index order) {
    return std::sqrt(static_cast<double>((degree - order) * (degree + order + 1)));
}
inline double ladder_down(Index degree, Index order) {
    return std::sqrt(static_cast<double>((degree + order) * (degree - order + 1)));
}

// The recurrence constants (harmonics.cpp RecurrenceConstants): the diagonal
// P(m, m) / sin^m = (-1)^m sqrt((2m + 1) / (4 pi) prod_{k<=m} (2k - 1) / (2k)) and the factors
// sqrt((4 l^2 - 1) / (l^2 - m^2)).
struct ShConstants {
    double diagonal[SH_MAX_LMAX + 1];
    double factor[SH_MAX_LMAX + 1][SH_MAX_LMAX + 1];
    ShConstants() {
        double ratio = 1.0;
        for (Index m = 0; m <= SH_MAX_LMAX; ++m) {
            const double mf = static_cast<double>(m);
            if (m > 0) {
                ratio = ratio * ((2.0 * mf - 1.0) / (2.0 * mf));
            }
            const double sign = m % 2 == 1 ? -1.0 : 1.0;
            diagonal[m] = sign * std::sqrt((2.0 * mf + 1.0) / (4.0 * M_PI) * ratio);
            for (Index l = 0; l <= SH_MAX_LMAX; ++l) {
                const double lf = static_cast<double>(l);
                factor[l][m] = l > m ? std::sqrt((4.0 * lf * lf - 1.0) / (lf * lf - mf * mf)) : 0.0;
            }
        }
    }
};
inline const ShConstants& sh_constants() {
    static const ShConstants c;
    return c;
}

// Normalised associated Legendre values P(l, m)(cos theta), 0 <= m <= l <= lmax
// (harmonics.cpp fill_legendre): the diagonal, the first off-diagonal, then the three-term
// recurrence in the degree.
inline void fill_legendre(Index lmax, double cos_theta, double sin_theta, LegendreTable& v) {
    const ShConstants& c = sh_constants();
    for (Index m = 0; m <= lmax; ++m) {
        v[m][m] = c.diagonal[m] * power_int(sin_theta, m);
        if (m < lmax) {
            v[m + 1][m] = std::sqrt(2.0 * static_cast<double>(m) + 3.0) * cos_theta * v[m][m];
        }
        for (Index l = m + 2; l <= lmax; ++l) {
            v[l][m] = c.factor[l][m] * (cos_theta * v[l - 1][m] - v[l - 2][m] / c.factor[l - 1][m]);
        }
    }
}

// dP(l, m) / dtheta and d2P(l, m) / dtheta2 from the ladder relations (harmonics.cpp
// Legendre::d_theta, d2_theta).
inline double legendre_d_theta(const LegendreTable& v, Index l, Index m) {
    if (l == 0) {
        return 0.0;
    }
    if (m == 0) {
        return ladder_up(l, 0) * v[l][1];
    }
    const double above = m < l ? ladder_up(l, m) * v[l][m + 1] : 0.0;
    return (above - ladder_down(l, m) * v[l][m - 1]) / 2.0;
}

inline double legendre_d2_theta(const LegendreTable& v, Index l, Index m) {
    if (l == 0) {
        return 0.0;
    }
    if (m == 0) {
        const double u0 = ladder_up(l, 0);
        return (u0 * ladder_up(l, 1) * v[l][2] - u0 * u0 * v[l][0]) / 2.0;
    }
    const double uh = ladder_up(l, m);
    const double dh = ladder_down(l, m);
    double total = -(uh * uh + dh * dh) * v[l][m];
    if (m == 1) {
        total = total - ladder_down(l, 1) * ladder_down(l, 0) * v[l][1];
    } else {
        total = total + ladder_down(l, m) * ladder_down(l, m - 1) * v[l][m - 2];
    }
    if (m + 2 <= l) {
        total = total + ladder_up(l, m) * ladder_up(l, m + 1) * v[l][m + 2];
    }
    return total / 4.0;
}

struct ShDerivatives {
    double value = 0.0, d_theta = 0.0, d_phi = 0.0, d_theta_theta = 0.0, d_theta_phi = 0.0,
           d_phi_phi = 0.0;
};

// The series and its derivatives on the sphere at (theta, phi) (harmonics.cpp derivatives):
// accumulated degree by degree, order 0 then orders 1..l; phi-derivatives divided by sin theta
// (per radian on the sphere), with the pole limit where sin theta < 1e-4.
inline ShDerivatives sh_derivatives(const double* c, Index lmax, bool full, double theta, double phi) {
    LegendreTable v{};
    fill_legendre(lmax, std::cos(theta), std::sin(theta), v);
    const double sin_theta = std::sin(theta);
    const bool at_pole = sin_theta < 1e-4;
    ShDerivatives r;
    const Index step = full ? 1 : 2;
    for (Index l = 0; l <= lmax; l += step) {
        const double zonal = c[sh_index(l, 0, full)];
        r.value = r.value + zonal * v[l][0];
        r.d_theta = r.d_theta + zonal * legendre_d_theta(v, l, 0);
        r.d_theta_theta = r.d_theta_theta + zonal * legendre_d2_theta(v, l, 0);
        for (Index m = 1; m <= l; ++m) {
            const double mf = static_cast<double>(m);
            const double cc = c[sh_index(l, m, full)];
            const double sc = c[sh_index(l, -m, full)];
            const double s_m = std::sin(mf * phi);
            const double c_m = std::cos(mf * phi);
            const double angular = M_SQRT2 * (cc * c_m + sc * s_m);
            const double angular_d_phi = M_SQRT2 * (sc * c_m - cc * s_m);
            const double lv = v[l][m];
            const double ldt = legendre_d_theta(v, l, m);
            r.value = r.value + angular * lv;
            r.d_theta = r.d_theta + angular * ldt;
            r.d_theta_theta = r.d_theta_theta + angular * legendre_d2_theta(v, l, m);
            if (at_pole) {
                r.d_phi = r.d_phi + angular_d_phi * ldt;
            } else {
                r.d_phi = r.d_phi + mf * angular_d_phi * lv;
                r.d_theta_phi = r.d_theta_phi + mf * angular_d_phi * ldt;
                r.d_phi_phi = r.d_phi_phi - mf * mf * angular * lv;
            }
        }
    }
    if (!at_pole) {
        r.d_phi = r.d_phi / sin_theta;
        r.d_theta_phi = r.d_theta_phi / sin_theta;
        r.d_phi_phi = r.d_phi_phi / (sin_theta * sin_theta);
    }
    return r;
}

struct ShPeak {
    double dir[3];
    double amplitude;
};

// Newton climb on the sphere from a unit start (peaks.cpp climb): the step along the unit
// gradient is |slope / curvature| capped at max_step; the climb returns the direction after the
// first step below tol with the amplitude before it, or nothing after max_iters steps.
inline bool sh_climb(const double* c, Index lmax, bool full, const double start[3], Index max_iters,
              double tol, double max_step, ShPeak& out) {
    double d[3] = {start[0], start[1], start[2]};
    for (Index it = 0; it < max_iters; ++it) {
        const double theta = std::acos(std::clamp(d[2], -1.0, 1.0));
        const double phi = std::atan2(d[1], d[0]);
        const ShDerivatives h = sh_derivatives(c, lmax, full, theta, phi);
        const double gn = std::sqrt(h.d_theta * h.d_theta + h.d_phi * h.d_phi);
        double at = 0.0;
        double ap = 0.0;
        if (gn != 0.0) {
            at = h.d_theta / gn;
            ap = h.d_phi / gn;
        }
        const double slope = at * h.d_theta + ap * h.d_phi;
        const double curvature =
            at * at * h.d_theta_theta + 2.0 * at * ap * h.d_theta_phi + ap * ap * h.d_phi_phi;
        double step = curvature != 0.0 ? std::fabs(slope / curvature) : 0.0;
        step = std::fmin(step, max_step);
        const double st = std::sin(theta);
        const double ct = std::cos(theta);
        const double sp = std::sin(phi);
        const double cp = std::cos(phi);
        const double tt[3] = {ct * cp, ct * sp, -st};
        const double tp[3] = {-sp, cp, 0.0};
        for (int a = 0; a < 3; ++a) {
            d[a] = d[a] + step * (at * tt[a] + ap * tp[a]);
        }
        const double norm = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        for (int a = 0; a < 3; ++a) {
            d[a] = d[a] / norm;
        }
        if (step < tol) {
            out = ShPeak{{d[0], d[1], d[2]}, h.value};
            return true;
        }
    }
    return false;
}

} // namespace oracle_sh
