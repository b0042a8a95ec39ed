// The (weighted) linear least-squares fit of one voxel's log signal on one wavefront.
//
// Follows larmor's fit_log_signal_block and reweighted_fit (cpp/src/dwi/log_fit.cpp:28-62,
// 159-185) for a design without constraints, the weighted linear least squares of Veraart et al.
// (NeuroImage 81, 2013, 335-346): the ordinary fit p0 = X+ y, then, for a weighted fit, refits
// with each measurement weighted by the square of the signal the current fit predicts, until no
// predicted log signal moves by more than a tolerance or after a number of refits.  A refit whose
// normal matrix is singular or whose result is not finite leaves the voxel at the ordinary fit.
//
// The design X (n x p) has unit-norm columns; the parameters of the original design are the
// scaled ones divided by the column norms.  Measurements are spread over lanes (row r in lane
// r % 64, slot r / 64); parameters are uniform.  Sums over measurements are wave sums (each
// lane adds its rows in slot order, then the XOR butterfly); a prediction x_r . p is summed in
// parameter order.  The weights S^2 are exp(2 (x_r . p - max_r x_r . p)): relative to the
// largest predicted signal, so the exponential cannot overflow (a common factor does not change
// the fit).
//
// The normal equations (X^T W X) q = X^T W y are solved by LU with partial pivoting, the
// augmented matrix [A | b] held one column per lane (column j in lane j, its rows in registers):
// a pivot search is a lane's own scan, a row swap and an elimination step are per-lane register
// updates with the multipliers read from the pivot lane.
#pragma once

#include "qr_wave.hpp"

namespace lsd {

// e^x for x <= 0 (and NaN), by Cody-Waite reduction x = k ln2 + r, |r| <= ln2 / 2, with ln2
// split into a 32-bit head (k ln2_hi exact) and its remainder, and the Taylor series of e^r to
// degree 13 (truncation below 1e-17 relative) by Horner's rule, scaled by 2^k.  The same
// operations on the host give the same bits.
__device__ __forceinline__ double exp_nonpositive(double x) {
    if (!(x > -746.0)) {
        return x != x ? x : 0.0;
    }
    const double k = rint(x * 0x1.71547652b82fep+0);
    const double r = (x - k * 0x1.62e42fee00000p-1) - k * 0x1.a39ef35793c76p-33;
    double e = 0x1.6124613a86d09p-33;   // 1/13!
    e = e * r + 0x1.1eed8eff8d898p-29;  // 1/12!
    e = e * r + 0x1.ae64567f544e4p-26;  // 1/11!
    e = e * r + 0x1.27e4fb7789f5cp-22;  // 1/10!
    e = e * r + 0x1.71de3a556c734p-19;  // 1/9!
    e = e * r + 0x1.a01a01a01a01ap-16;  // 1/8!
    e = e * r + 0x1.a01a01a01a01ap-13;  // 1/7!
    e = e * r + 0x1.6c16c16c16c17p-10;  // 1/6!
    e = e * r + 0x1.1111111111111p-7;   // 1/5!
    e = e * r + 0x1.5555555555555p-5;   // 1/4!
    e = e * r + 0x1.5555555555555p-3;   // 1/3!
    e = e * r + 0.5;
    e = e * r + 1.0;
    e = e * r + 1.0;
    return ldexp(e, static_cast<int>(k));
}

struct LogFitDev {
    const double* xt;     // (p, n): the unit-column design, transposed
    const double* pinv;   // (p, n): its pseudo-inverse
    const double* scales; // (p): the column norms of the original design
    int n;
    int p;
    int weighted;
    int max_reweightings;
    double tol;
};

// Solves A q = b (A p x p, uniform) by LU with partial pivoting; false when a pivot is exactly 0.
// a[i] of lane j is A(i, j) for j < p and b_i for lane p.
template <int PMAX>
__device__ __forceinline__ bool lu_solve_columns(double (&a)[PMAX], int p, double (&q)[PMAX]) {
    const int l = lane_id();
    auto at = [&](int i) {
        double v = 0.0;
#pragma unroll
        for (int u = 0; u < PMAX; ++u) {
            v = u == i ? a[u] : v;
        }
        return v;
    };
#pragma unroll
    for (int k = 0; k < PMAX; ++k) {
        if (k >= p) {
            break;
        }
        // The pivot: the first row i >= k of largest |A(i, k)| (lane k's scan).
        int best = k;
        double best_abs = fabs(a[k]);
#pragma unroll
        for (int i = k + 1; i < PMAX; ++i) {
            if (i < p && fabs(a[i]) > best_abs) {
                best = i;
                best_abs = fabs(a[i]);
            }
        }
        const int piv = readlane(best, k);
        const double pivot = readlane(at(piv), k);
        if (pivot == 0.0) {
            return false;
        }
        if (piv != k) {
            const double row_k = a[k];
            const double row_p = at(piv);
#pragma unroll
            for (int u = 0; u < PMAX; ++u) {
                a[u] = u == k ? row_p : (u == piv ? row_k : a[u]);
            }
        }
        const double ukj = a[k];
#pragma unroll
        for (int i = k + 1; i < PMAX; ++i) {
            if (i < p) {
                const double lik = readlane(a[i], k) / pivot;
                const double updated = a[i] - lik * ukj;
                a[i] = (l > k && l <= p) ? updated : a[i];
            }
        }
    }
    // Back substitution on U q = b', the terms subtracted in increasing column order.
#pragma unroll
    for (int k = PMAX - 1; k >= 0; --k) {
        q[k] = 0.0;
        if (k < p) {
            double s = readlane(a[k], p);
#pragma unroll
            for (int j = k + 1; j < PMAX; ++j) {
                if (j < p) {
                    s = s - readlane(a[k], j) * q[j];
                }
            }
            q[k] = s / readlane(a[k], k);
        }
    }
    return true;
}

// The fit of one voxel: params (scaled back to the original design) and the number of weighted
// refits kept (-1 when a refit failed and the ordinary fit was kept).
template <int RPL, int PMAX>
__device__ __forceinline__ int log_fit_voxel(const LogFitDev& F, const RowVec<RPL>& y,
                                             double (&params)[PMAX]) {
    const int l = lane_id();
    double x[PMAX][RPL];
#pragma unroll
    for (int a = 0; a < PMAX; ++a) {
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            const int r = l + WAVE * t;
            x[a][t] = (a < F.p && r < F.n) ? F.xt[static_cast<long>(a) * F.n + r] : 0.0;
        }
    }
    double ordinary[PMAX];
    bool finite = true;
#pragma unroll
    for (int a = 0; a < PMAX; ++a) {
        ordinary[a] = 0.0;
        if (a < F.p) {
            const double* row = F.pinv + static_cast<long>(a) * F.n;
            ordinary[a] = wave_sum_f<RPL>(0, F.n, [&](int t) {
                const int r = l + WAVE * t;
                return (r < F.n ? row[r] : 0.0) * y.v[t];
            });
            finite = finite && isfinite(ordinary[a]);
        }
    }
    double cur[PMAX];
#pragma unroll
    for (int a = 0; a < PMAX; ++a) {
        cur[a] = ordinary[a];
    }
    int refits = 0;
    if (F.weighted && finite) {
        auto predict = [&](const double (&q)[PMAX], double (&out)[RPL]) {
#pragma unroll
            for (int t = 0; t < RPL; ++t) {
                double s = 0.0;
#pragma unroll
                for (int a = 0; a < PMAX; ++a) {
                    if (a < F.p) {
                        s = s + x[a][t] * q[a];
                    }
                }
                out[t] = s;
            }
        };
        double pred[RPL];
        predict(cur, pred);
        for (int it = 0; it < F.max_reweightings; ++it) {
            double lane_max = -__builtin_inf();
#pragma unroll
            for (int t = 0; t < RPL; ++t) {
                const int r = l + WAVE * t;
                lane_max = (r < F.n && pred[t] > lane_max) ? pred[t] : lane_max;
            }
            const double top = wave_max(lane_max);
            double w[RPL];
#pragma unroll
            for (int t = 0; t < RPL; ++t) {
                w[t] = exp_nonpositive(2.0 * (pred[t] - top));
            }
            // The upper triangle of X^T W X (mirrored) and X^T W y, into the lanes' columns.
            double col[PMAX];
#pragma unroll
            for (int u = 0; u < PMAX; ++u) {
                col[u] = 0.0;
            }
#pragma unroll
            for (int a = 0; a < PMAX; ++a) {
                if (a < F.p) {
                    double wx[RPL];
#pragma unroll
                    for (int t = 0; t < RPL; ++t) {
                        wx[t] = w[t] * x[a][t];
                    }
                    const double rhs = wave_sum_f<RPL>(0, F.n, [&](int t) { return wx[t] * y.v[t]; });
                    col[a] = l == F.p ? rhs : col[a];
#pragma unroll
                    for (int b = a; b < PMAX; ++b) {
                        if (b < F.p) {
                            const double s = wave_sum_f<RPL>(0, F.n, [&](int t) { return wx[t] * x[b][t]; });
                            col[a] = l == b ? s : col[a];
                            col[b] = l == a ? s : col[b];
                        }
                    }
                }
            }
            double next[PMAX];
            if (!lu_solve_columns<PMAX>(col, F.p, next)) {
                refits = -1;
                break;
            }
            bool next_finite = true;
#pragma unroll
            for (int a = 0; a < PMAX; ++a) {
                next_finite = next_finite && (a >= F.p || isfinite(next[a]));
            }
            if (!next_finite) {
                refits = -1;
                break;
            }
            double next_pred[RPL];
            predict(next, next_pred);
            double lane_change = 0.0;
#pragma unroll
            for (int t = 0; t < RPL; ++t) {
                const int r = l + WAVE * t;
                const double d = fabs(next_pred[t] - pred[t]);
                lane_change = (r < F.n && d > lane_change) ? d : lane_change;
            }
            const double change = wave_max(lane_change);
#pragma unroll
            for (int a = 0; a < PMAX; ++a) {
                cur[a] = next[a];
            }
#pragma unroll
            for (int t = 0; t < RPL; ++t) {
                pred[t] = next_pred[t];
            }
            refits = it + 1;
            if (!(change > F.tol)) {
                break;
            }
        }
        if (refits < 0) {
#pragma unroll
            for (int a = 0; a < PMAX; ++a) {
                cur[a] = ordinary[a];
            }
        }
    }
#pragma unroll
    for (int a = 0; a < PMAX; ++a) {
        params[a] = a < F.p ? cur[a] / F.scales[a] : 0.0;
    }
    return refits;
}

} // namespace lsd
