// Peaks of SH series on the sphere: one voxel per wavefront, one Newton climb per lane.
//
// Restates larmor's peak search (cpp/src/sh/peaks.cpp:33-133, with the derivatives of
// cpp/src/sh/harmonics.cpp:73-185):
//  - the series' amplitudes on the search sphere (given, or computed from the basis with the sum
//    over coefficients in order) go to LDS; a vertex whose positive amplitude no neighbour
//    exceeds is a local maximum;
//  - each local maximum starts a climb on its own lane: at most 50 Newton steps along the unit
//    gradient in the (theta, phi) tangent frame, the step |slope / curvature| capped at 0.2,
//    ending at the first step below 1e-4 with the amplitude before that step;
//  - the climbs' peaks are merged in vertex order, uniformly across the wave: a peak within 0.99
//    in cosine (|cos| for a symmetric series) of a kept one replaces it when larger, otherwise
//    it is appended; then the largest peaks, ties in the order kept, are cut at the absolute and
//    relative thresholds and symmetric peaks are turned to z >= 0.
//
// A climb's Legendre values live in registers, one degree at a time: LMAX and FULL are template
// parameters and every index is a compile-time constant.  The kernels instantiated for lmax 20
// take any lower lmax at run time (from the constants), so every lmax up to 20 is served, the
// common ones (12 and below symmetric, 8 and below full) by exact instantiations.  The
// transcendental functions (acos, atan2, sin, cos) are the device library's, so the results match
// the CPU oracle to rounding, not to the bit.
#pragma once

#include "wave.hpp"

namespace lsd {

constexpr int SH_TABLE = 21;       // recurrence constants up to lmax 20
constexpr int PEAK_VERTEX_SLOTS = 48; // search spheres of at most 3072 vertices
constexpr int PEAK_FOUND_SLOTS = 2;   // at most 128 distinct peaks per voxel

// Constant tables, made on the host with the same expressions as the CPU oracle.
struct ShConstantsDev {
    const double* diagonal; // (21): P(m, m) / sin^m theta
    const double* factor;   // (21, 21): sqrt((4 l^2 - 1) / (l^2 - m^2)), l > m
    const double* up;       // (21, 21): sqrt((l - m)(l + m + 1)), m <= l
    const double* down;     // (21, 21): sqrt((l + m)(l - m + 1)), m <= l
    const double* offdiag;  // (21): sqrt(2 m + 3)
    int lmax;               // the series' lmax (read by the lmax-20 instantiations)
};

// The lmax of the instantiation that serves every lmax above the exact ones.
constexpr int SH_WIDE_LMAX = 20;

// The series' lmax: LMAX, or for the wide instantiation the run-time value.
template <int LMAX>
__device__ __forceinline__ int series_lmax(const ShConstantsDev& k) {
    return LMAX == SH_WIDE_LMAX ? k.lmax : LMAX;
}

constexpr int sh_index(int l, int m, bool full) {
    return full ? l * l + l + m : l * (l + 1) / 2 + m;
}

__device__ __forceinline__ double power_int_dev(double base, int exponent) {
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

// Degree l of the normalised associated Legendre values, row[m] = P(l, m) for m <= l, from the
// two degrees before it: fill_legendre's recurrences (P(m, m) from the diagonal constant, P(m + 1,
// m) from P(m, m), then in l), the same expression for every value, so the same values as a table
// filled order by order.
template <int LMAX>
__device__ __forceinline__ void legendre_row(int l, double cos_t, double sin_t,
                                             const double (&prev1)[LMAX + 1],
                                             const double (&prev2)[LMAX + 1],
                                             double (&row)[LMAX + 1], const ShConstantsDev& c) {
#pragma unroll
    for (int m = 0; m <= LMAX; ++m) {
        if (m > l) {
            break;
        }
        if (m == l) {
            row[m] = c.diagonal[m] * power_int_dev(sin_t, m);
        } else if (m == l - 1) {
            row[m] = c.offdiag[m] * cos_t * prev1[m];
        } else {
            row[m] = c.factor[l * SH_TABLE + m] *
                     (cos_t * prev1[m] - prev2[m] / c.factor[(l - 1) * SH_TABLE + m]);
        }
    }
}

__device__ __forceinline__ double ladder_up_dev(const ShConstantsDev& k, int l, int m) {
    return k.up[l * SH_TABLE + m];
}
__device__ __forceinline__ double ladder_down_dev(const ShConstantsDev& k, int l, int m) {
    return k.down[l * SH_TABLE + m];
}

// d/dtheta of P(l, m) from the degree's row (legendre_d_theta).
template <int LMAX>
__device__ __forceinline__ double legendre_d_theta_row(const ShConstantsDev& k, int l, int m,
                                                       const double (&row)[LMAX + 1]) {
    if (l == 0) return 0.0;
    if (m == 0) return ladder_up_dev(k, l, 0) * row[1];
    const double above = m < l ? ladder_up_dev(k, l, m) * row[m < LMAX ? m + 1 : m] : 0.0;
    return (above - ladder_down_dev(k, l, m) * row[m - 1]) / 2.0;
}

// d2/dtheta2 of P(l, m) from the degree's row (legendre_d2_theta).  At l = 1 the m = 0 form's
// P(l, 2) term has a zero ladder factor; it is taken as zero.
template <int LMAX>
__device__ __forceinline__ double legendre_d2_theta_row(const ShConstantsDev& k, int l, int m,
                                                        const double (&row)[LMAX + 1]) {
    if (l == 0) return 0.0;
    if (m == 0) {
        const double u0 = ladder_up_dev(k, l, 0);
        const double p2 = l >= 2 ? row[2 <= LMAX ? 2 : 0] : 0.0;
        return (u0 * ladder_up_dev(k, l, 1) * p2 - u0 * u0 * row[0]) / 2.0;
    }
    const double uh = ladder_up_dev(k, l, m);
    const double dh = ladder_down_dev(k, l, m);
    double total = -(uh * uh + dh * dh) * row[m];
    if (m == 1) {
        total = total - ladder_down_dev(k, l, 1) * ladder_down_dev(k, l, 0) * row[1];
    } else {
        total = total + ladder_down_dev(k, l, m) * ladder_down_dev(k, l, m - 1) * row[m - 2];
    }
    if (m + 2 <= l) {
        total = total + ladder_up_dev(k, l, m) * ladder_up_dev(k, l, m + 1) *
                            row[m + 2 <= LMAX ? m + 2 : m];
    }
    return total / 4.0;
}

struct ShDerivs {
    double value, d_theta, d_phi, d_theta_theta, d_theta_phi, d_phi_phi;
};

// derivatives (harmonics.cpp:139-185) of the series c at (theta, phi), per lane, given
// sin theta, cos theta, sin phi and cos phi (the values of sin(m phi), cos(m phi) at m = 1).
// The sums run over the degrees in increasing l and within a degree over m = 0..l; the Legendre
// values come one degree at a time (legendre_row), so two degrees are live at once.
template <int LMAX, bool FULL>
__device__ __forceinline__ ShDerivs sh_derivs(const double* c, double sin_theta,
                                              double cos_theta, double phi, double sin_phi,
                                              double cos_phi, const ShConstantsDev& k) {
    const int lmax = series_lmax<LMAX>(k);
    const bool at_pole = sin_theta < 1e-4;
    double s_m[LMAX + 1];
    double c_m[LMAX + 1];
#pragma unroll
    for (int m = 1; m <= LMAX; ++m) {
        if (m > lmax) {
            break;
        }
        if (m == 1) {
            s_m[m] = sin_phi;
            c_m[m] = cos_phi;
        } else {
            sincos(static_cast<double>(m) * phi, &s_m[m], &c_m[m]);
        }
    }
    ShDerivs r{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double prev2[LMAX + 1];
    double prev1[LMAX + 1];
#pragma unroll
    for (int l = 0; l <= LMAX; ++l) {
        if (l > lmax) {
            break;
        }
        double row[LMAX + 1];
        legendre_row<LMAX>(l, cos_theta, sin_theta, prev1, prev2, row, k);
        if (FULL || l % 2 == 0) {
            const double zonal = c[sh_index(l, 0, FULL)];
            r.value = r.value + zonal * row[0];
            r.d_theta = r.d_theta + zonal * legendre_d_theta_row<LMAX>(k, l, 0, row);
            r.d_theta_theta = r.d_theta_theta + zonal * legendre_d2_theta_row<LMAX>(k, l, 0, row);
#pragma unroll
            for (int m = 1; m <= l; ++m) {
                const double mf = static_cast<double>(m);
                const double cc = c[sh_index(l, m, FULL)];
                const double sc = c[sh_index(l, -m, FULL)];
                const double angular = M_SQRT2 * (cc * c_m[m] + sc * s_m[m]);
                const double angular_d_phi = M_SQRT2 * (sc * c_m[m] - cc * s_m[m]);
                const double lv = row[m];
                const double ldt = legendre_d_theta_row<LMAX>(k, l, m, row);
                r.value = r.value + angular * lv;
                r.d_theta = r.d_theta + angular * ldt;
                r.d_theta_theta =
                    r.d_theta_theta + angular * legendre_d2_theta_row<LMAX>(k, l, m, row);
                if (at_pole) {
                    r.d_phi = r.d_phi + angular_d_phi * ldt;
                } else {
                    r.d_phi = r.d_phi + mf * angular_d_phi * lv;
                    r.d_theta_phi = r.d_theta_phi + mf * angular_d_phi * ldt;
                    r.d_phi_phi = r.d_phi_phi - mf * mf * angular * lv;
                }
            }
        }
#pragma unroll
        for (int m = 0; m <= l; ++m) {
            if (m < l) {
                prev2[m] = prev1[m];
            }
            prev1[m] = row[m];
        }
    }
    if (!at_pole) {
        r.d_phi = r.d_phi / sin_theta;
        r.d_theta_phi = r.d_theta_phi / sin_theta;
        r.d_phi_phi = r.d_phi_phi / (sin_theta * sin_theta);
    }
    return r;
}

// climb (peaks.cpp:33-77) from d (unit), per lane; active lanes only.  Returns whether it
// stopped; d and amp then hold the peak.
template <int LMAX, bool FULL>
__device__ __forceinline__ bool sh_climb_lane(const double* c, double (&d)[3], double& amp,
                                              int max_iters, double tol, double max_step,
                                              const ShConstantsDev& k) {
    for (int it = 0; it < max_iters; ++it) {
        const double z = d[2] < -1.0 ? -1.0 : (d[2] > 1.0 ? 1.0 : d[2]);
        const double theta = acos(z);
        const double phi = atan2(d[1], d[0]);
        double st, ct, sp, cp;
        sincos(theta, &st, &ct);
        sincos(phi, &sp, &cp);
        const ShDerivs h = sh_derivs<LMAX, FULL>(c, st, ct, phi, sp, cp, k);
        const double gn = sqrt(h.d_theta * h.d_theta + h.d_phi * h.d_phi);
        double at = 0.0;
        double ap = 0.0;
        if (gn != 0.0) {
            at = h.d_theta / gn;
            ap = h.d_phi / gn;
        }
        const double slope = at * h.d_theta + ap * h.d_phi;
        const double curvature =
            at * at * h.d_theta_theta + 2.0 * at * ap * h.d_theta_phi + ap * ap * h.d_phi_phi;
        double step = curvature != 0.0 ? fabs(slope / curvature) : 0.0;
        step = fmin(step, max_step);
        d[0] = d[0] + step * (at * (ct * cp) + ap * (-sp));
        d[1] = d[1] + step * (at * (ct * sp) + ap * cp);
        d[2] = d[2] + step * (at * (-st) + ap * 0.0);
        const double norm = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        d[0] = d[0] / norm;
        d[1] = d[1] / norm;
        d[2] = d[2] / norm;
        if (step < tol) {
            amp = h.value;
            return true;
        }
    }
    return false;
}

struct PeaksArgs {
    const double* coeffs;    // (n_vox, n_coeffs), tournier basis
    long n_vox;
    int n_coeffs;
    const double* on_sphere; // (n_vox, nv), or null: computed from basis_t
    const double* basis_t;   // (n_coeffs, nv)
    const double* vertices;  // (nv, 3)
    const int* neighbours;   // (nv, 6)
    int nv;
    int n_peaks;             // <= 64
    double relative_threshold;
    double threshold;
    ShConstantsDev k;
    double* directions;      // (n_vox, n_peaks, 3)
    double* amplitudes;      // (n_vox, n_peaks)
    int* overflow;           // voxels with more distinct peaks than the wave keeps
};

// LDS per wave for a sphere of nv vertices: the amplitudes (doubles) and the ranked
// local-maximum list (ints), one vertex slot of 64 per lane row.
inline int peak_vertex_slots(int nv) { return nv > 0 ? (nv + WAVE - 1) / WAVE : 1; }
inline int peaks_lds_bytes_per_wave(int nv) {
    return peak_vertex_slots(nv) * WAVE * (8 + 4);
}

template <int LMAX, bool FULL>
__global__ void __launch_bounds__(WAVE * 4) peaks_kernel(PeaksArgs A) {
    const int nvs = A.nv > 0 ? (A.nv + WAVE - 1) / WAVE : 1;
    constexpr int NF = PEAK_FOUND_SLOTS;
    extern __shared__ double smem[];
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / WAVE);
    const long gw = static_cast<long>(blockIdx.x) * (blockDim.x / WAVE) + w;
    const long n_waves = static_cast<long>(gridDim.x) * (blockDim.x / WAVE);
    char* base = reinterpret_cast<char*>(smem) + static_cast<long>(w) * nvs * WAVE * (8 + 4);
    double* amp_lds = reinterpret_cast<double*>(base);
    int* order_lds = reinterpret_cast<int*>(base + nvs * WAVE * 8);
    const int l = lane_id();
    const double nan = __builtin_nan("");

    for (long v = gw; v < A.n_vox; v += n_waves) {
        const double* c = A.coeffs + v * A.n_coeffs;
        // Amplitudes on the sphere into LDS (vertex d at amp_lds[d]).
#pragma unroll 1
        for (int d = l; d < A.nv; d += WAVE) {
            double x = 0.0;
            if (A.on_sphere) {
                x = A.on_sphere[v * A.nv + d];
            } else {
                for (int q = 0; q < A.n_coeffs; ++q) {
                    x = x + A.basis_t[static_cast<long>(q) * A.nv + d] * c[q];
                }
            }
            amp_lds[d] = x;
        }
        wave_lds_sync();
        const bool finite_first = A.nv > 0 && isfinite(amp_lds[0]);
        // Local maxima, ranked in vertex order.
        int n_max = 0;
        if (finite_first) {
#pragma unroll 1
            for (int t = 0; t < nvs; ++t) {
                const int d = l + WAVE * t;
                const double here = d < A.nv ? amp_lds[d] : 0.0;
                bool is_max = d < A.nv && here > 0.0;
                if (is_max) {
                    for (int q = 0; q < 6; ++q) {
                        is_max = is_max && amp_lds[A.neighbours[d * 6 + q]] <= here;
                    }
                }
                const unsigned long long bits = ballot(is_max);
                if (is_max) {
                    const unsigned long long below = (1ull << l) - 1ull;
                    order_lds[n_max + __popcll(bits & below)] = d;
                }
                n_max += __popcll(bits);
            }
        }
        wave_lds_sync();

        // Kept peaks: peak p in lane p % 64, slot p / 64.
        double px[NF], py[NF], pz[NF], pa[NF];
#pragma unroll
        for (int s = 0; s < NF; ++s) {
            px[s] = py[s] = pz[s] = pa[s] = 0.0;
        }
        int n_found = 0;
        bool overflow = false;
        for (int first = 0; first < n_max; first += WAVE) {
            const int count = n_max - first < WAVE ? n_max - first : WAVE;
            double d[3] = {0.0, 0.0, 0.0};
            double amp = 0.0;
            bool ok = false;
            if (l < count) {
                const int vert = order_lds[first + l];
                d[0] = A.vertices[vert * 3 + 0];
                d[1] = A.vertices[vert * 3 + 1];
                d[2] = A.vertices[vert * 3 + 2];
                ok = sh_climb_lane<LMAX, FULL>(c, d, amp, 50, 1e-4, 0.2, A.k);
            }
            // Merge in vertex order.
            for (int j = 0; j < count; ++j) {
                const bool okj = readlane(static_cast<int>(ok), j) != 0;
                const double aj = readlane(amp, j);
                if (!okj || !(aj > 0.0)) {
                    continue;
                }
                const double xj = readlane(d[0], j);
                const double yj = readlane(d[1], j);
                const double zj = readlane(d[2], j);
                int match = -1;
#pragma unroll
                for (int s = 0; s < NF; ++s) {
                    const int p = l + WAVE * s;
                    const double cosine = xj * px[s] + yj * py[s] + zj * pz[s];
                    const bool same = p < n_found && (FULL ? cosine : fabs(cosine)) > 0.99;
                    const unsigned long long bits = ballot(same);
                    if (match < 0 && bits) {
                        match = WAVE * s + __ffsll(static_cast<long long>(bits)) - 1;
                    }
                }
                int target = -1;
                if (match < 0) {
                    if (n_found < WAVE * NF) {
                        target = n_found;
                        n_found += 1;
                    } else {
                        overflow = true;
                    }
                } else {
                    double am = 0.0;
#pragma unroll
                    for (int s = 0; s < NF; ++s) {
                        am = (match / WAVE) == s ? readlane(pa[s], match % WAVE) : am;
                    }
                    if (aj > am) {
                        target = match;
                    }
                }
                if (target >= 0) {
#pragma unroll
                    for (int s = 0; s < NF; ++s) {
                        if (l + WAVE * s == target) {
                            px[s] = xj;
                            py[s] = yj;
                            pz[s] = zj;
                            pa[s] = aj;
                        }
                    }
                }
            }
        }
        if (overflow && l == 0) {
            atomicAdd(A.overflow, 1);
        }

        // The largest peaks, ties in the order kept (a stable sort), cut at the thresholds.
        bool taken[NF];
#pragma unroll
        for (int s = 0; s < NF; ++s) {
            taken[s] = false;
        }
        const int n_rank = n_found < A.n_peaks ? n_found : A.n_peaks;
        int my_pick = -1;
        double largest = 0.0;
        for (int r = 0; r < n_rank; ++r) {
            double bv = 0.0;
            int bi = -1;
#pragma unroll
            for (int s = 0; s < NF; ++s) {
                const int p = l + WAVE * s;
                if (p < n_found && !taken[s] && (bi < 0 || pa[s] > bv)) {
                    bv = pa[s];
                    bi = p;
                }
            }
            wave_argmax_first(bv, bi);
            if (r == 0) {
                largest = bv;
            }
            if (bv < A.threshold || bv < A.relative_threshold * largest) {
                break;
            }
#pragma unroll
            for (int s = 0; s < NF; ++s) {
                taken[s] = taken[s] || (l + WAVE * s == bi);
            }
            my_pick = l == r ? bi : my_pick;
        }
        // Lane r writes rank r: the picked peak's direction and amplitude, or NaN.
        double ox = nan, oy = nan, oz = nan, oa = nan;
#pragma unroll
        for (int s = 0; s < NF; ++s) {
            const int src = my_pick < 0 ? 0 : my_pick % WAVE;
            const double gx = gather(px[s], src);
            const double gy = gather(py[s], src);
            const double gz = gather(pz[s], src);
            const double ga = gather(pa[s], src);
            if (my_pick >= 0 && my_pick / WAVE == s) {
                ox = gx;
                oy = gy;
                oz = gz;
                oa = ga;
            }
        }
        if (my_pick >= 0 && !FULL && oz < 0.0) {
            ox = -ox;
            oy = -oy;
            oz = -oz;
        }
        if (l < A.n_peaks) {
            const long o = v * A.n_peaks + l;
            A.amplitudes[o] = oa;
            A.directions[3 * o + 0] = ox;
            A.directions[3 * o + 1] = oy;
            A.directions[3 * o + 2] = oz;
        }
        wave_lds_sync();
    }
}

// Batched derivatives: series i at (theta[i], phi[i]) -> out (n, 6).
template <int LMAX, bool FULL>
__global__ void __launch_bounds__(WAVE * 4) sh_derivatives_kernel(const double* coeffs, long n,
                                                                  int n_coeffs, const double* theta,
                                                                  const double* phi, double* out,
                                                                  ShConstantsDev k) {
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double st, ct, sp, cp;
    sincos(theta[i], &st, &ct);
    sincos(phi[i], &sp, &cp);
    const ShDerivs r = sh_derivs<LMAX, FULL>(coeffs + i * n_coeffs, st, ct, phi[i], sp, cp, k);
    out[6 * i + 0] = r.value;
    out[6 * i + 1] = r.d_theta;
    out[6 * i + 2] = r.d_phi;
    out[6 * i + 3] = r.d_theta_theta;
    out[6 * i + 4] = r.d_theta_phi;
    out[6 * i + 5] = r.d_phi_phi;
}

// Batched climbs: series i from starts[i] -> directions (n, 3), amplitudes (n); NaN when a climb
// does not stop.
template <int LMAX, bool FULL>
__global__ void __launch_bounds__(WAVE * 4) sh_climb_kernel(const double* coeffs, long n,
                                                            int n_coeffs, const double* starts,
                                                            int max_iters, double tol,
                                                            double max_step, double* directions,
                                                            double* amplitudes, ShConstantsDev k) {
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d[3] = {starts[3 * i], starts[3 * i + 1], starts[3 * i + 2]};
    double amp = 0.0;
    const bool ok = sh_climb_lane<LMAX, FULL>(coeffs + i * n_coeffs, d, amp, max_iters, tol,
                                              max_step, k);
    const double nan = __builtin_nan("");
    directions[3 * i + 0] = ok ? d[0] : nan;
    directions[3 * i + 1] = ok ? d[1] : nan;
    directions[3 * i + 2] = ok ? d[2] : nan;
    amplitudes[i] = ok ? amp : nan;
}

} // namespace lsd
