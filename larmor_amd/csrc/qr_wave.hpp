// Column-pivoted Householder least squares and the generic Lawson-Hanson NNLS on one wavefront.
//
// Restates, step for step, the least squares of larmor's NNLS (cpp/src/sd/nnls.cpp:least_squares,
// lines 51-71: Eigen's ColPivHouseholderQR with threshold max(rows, cols) * eps, Q^T d over the
// first `rank` reflectors, then R11 back substitution scattered through the column permutation)
// and solve_nnls_with_limit (nnls.cpp:258-278, lawson_hanson 120-153).  Eigen's QR is
// Eigen 5.0.1's published algorithm: LAPACK xGEQP3 column pivoting with the LAWN 176 norm
// downdate and Householder vectors beta = -sign(c0) ||x||, tau = (beta - c0) / beta.
//
// Rows are spread over lanes (RowVec layout); every row sum is a wave sum (wave.hpp), the order
// the CPU oracle uses.  Columns come from a store: RegCols keeps up to NQ columns in registers
// (the fraction fits, a handful of columns), MemCols keeps any number in a per-wave global
// scratch slab (the rare QR fallback of the ridge NNLS).
#pragma once

#include <cfloat>
#include <utility>

#include "wave.hpp"

namespace lsd {

constexpr double DBL_EPS = 2.220446049250313e-16;

// Up to NQ columns of RPL row slots each, in registers.  Column indices are uniform; a column
// is picked by a compile-time-unrolled select over the NQ column vectors, so the columns stay
// separate register vectors.
template <int NQ, int RPL>
struct RegCols {
    static constexpr int kRpl = RPL;
    dvec<RPL> col[NQ];

    template <int... Q>
    __device__ __forceinline__ double get_(int c, int t, std::integer_sequence<int, Q...>) const {
        double x = 0.0;
        ((x = c == Q ? static_cast<double>(col[Q][t]) : x), ...);
        return x;
    }
    template <int... Q>
    __device__ __forceinline__ void set_(int c, int t, double x, std::integer_sequence<int, Q...>) {
        ((col[Q][t] = c == Q ? x : static_cast<double>(col[Q][t])), ...);
    }
    __device__ __forceinline__ double get(int c, int t) const {
        return get_(c, t, std::make_integer_sequence<int, NQ>{});
    }
    __device__ __forceinline__ void set(int c, int t, double x) {
        set_(c, t, x, std::make_integer_sequence<int, NQ>{});
    }
    // Element (row r, column c), both uniform.
    __device__ __forceinline__ double elem(int c, int r) const {
        return readlane(get(c, r / WAVE), r % WAVE);
    }
    __device__ __forceinline__ void set_elem(int c, int r, double x) {
        if (lane_id() == r % WAVE) {
            set(c, r / WAVE, x);
        }
    }
    __device__ __forceinline__ void swap_cols(int a, int b) {
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            const double va = get(a, t);
            const double vb = get(b, t);
            set(a, t, vb);
            set(b, t, va);
        }
    }
};

// Columns in a per-wave slab (LDS or global memory), column-major with leading dimension
// RPL * 64.
template <int RPL>
struct MemCols {
    static constexpr int kRpl = RPL;
    static constexpr int kLd = RPL * WAVE;
    double* base;

    __device__ __forceinline__ double get(int c, int t) const {
        return base[static_cast<long>(c) * kLd + t * WAVE + lane_id()];
    }
    __device__ __forceinline__ void set(int c, int t, double x) {
        base[static_cast<long>(c) * kLd + t * WAVE + lane_id()] = x;
    }
    // A uniform element, read as a vector load of its 64-row segment plus readlane (never
    // through the scalar cache, as the slab is written in the same kernel).
    __device__ __forceinline__ double elem(int c, int r) const {
        const double x = base[static_cast<long>(c) * kLd + (r / WAVE) * WAVE + lane_id()];
        return readlane(x, r % WAVE);
    }
    __device__ __forceinline__ void set_elem(int c, int r, double x) {
        if (lane_id() == r % WAVE) {
            base[static_cast<long>(c) * kLd + r] = x;
        }
    }
    __device__ __forceinline__ void swap_cols(int a, int b) {
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            const double va = get(a, t);
            const double vb = get(b, t);
            set(a, t, vb);
            set(b, t, va);
        }
    }
};

// The wave sum over rows lo <= r < hi of f(t), f evaluated on the lane's slot t.  f is
// evaluated on every slot (its reads stay inside the lane's slots) and the terms outside the
// range are left out by a select, so the lanes do not diverge.
template <int RPL, class F>
__device__ __forceinline__ double wave_sum_f(int lo, int hi, F f) {
    const int l = lane_id();
    double s = 0.0;
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
        const int r = l + WAVE * t;
        const double term = f(t);
        const double sum = s + term;
        s = (r >= lo && r < hi) ? sum : s;
    }
    return butterfly_sum(s);
}

// Least squares of d on the `cols` columns of A (rows x cols), A overwritten by its QR.
// coef receives the coefficient of each column 0..cols-1 (zero past the numerical rank).
// d is consumed.  NSC slots cover the column count.
template <int NSC, class Cols>
__device__ __forceinline__ void least_squares_wave(Cols& A, int rows, int cols, RowVec<Cols::kRpl>& d,
                                   LaneVec<NSC>& coef) {
    constexpr int RPL = Cols::kRpl;
    const int l = lane_id();
    const int size = rows < cols ? rows : cols;
    LaneVec<NSC> nu;
    LaneVec<NSC> nd;
    LaneVec<NSC> hc;
    LaneVec<NSC, int> trans;
    nu.fill(0.0);
    nd.fill(0.0);
    hc.fill(0.0);
    trans.fill(0);
    for (int c = 0; c < cols; ++c) {
        const double s = wave_sum_f<RPL>(0, rows, [&](int t) {
            const double x = A.get(c, t);
            return x * x;
        });
        const double nrm = sqrt(s);
        nd.set(c, nrm);
        nu.set(c, nrm);
    }
    double max_norm = 0.0;
    for (int c = 0; c < cols; ++c) {
        const double v = nu.get(c);
        if (c == 0 || v > max_norm) {
            max_norm = v;
        }
    }
    const double scaled = max_norm * DBL_EPS;
    const double threshold_helper = (scaled * scaled) / static_cast<double>(rows);
    const double norm_downdate_threshold = sqrt(DBL_EPS);
    int nonzero_pivots = size;
    double maxpivot = 0.0;

    for (int k = 0; k < size; ++k) {
        int biggest = k;
        double bv = nu.get(k);
        for (int c = k + 1; c < cols; ++c) {
            const double v = nu.get(c);
            if (v > bv) {
                bv = v;
                biggest = c;
            }
        }
        const double biggest_sq = bv * bv;
        if (nonzero_pivots == size &&
            biggest_sq < threshold_helper * static_cast<double>(rows - k)) {
            nonzero_pivots = k;
        }
        trans.set(k, biggest);
        if (k != biggest) {
            A.swap_cols(k, biggest);
            const double u1 = nu.get(k);
            const double u2 = nu.get(biggest);
            nu.set(k, u2);
            nu.set(biggest, u1);
            const double d1 = nd.get(k);
            const double d2 = nd.get(biggest);
            nd.set(k, d2);
            nd.set(biggest, d1);
        }

        // Householder vector of column k.
        const double c0 = A.elem(k, k);
        const double tail_sq = wave_sum_f<RPL>(k + 1, rows, [&](int t) {
            const double x = A.get(k, t);
            return x * x;
        });
        double tau;
        double beta;
        if (tail_sq <= DBL_MIN) {
            tau = 0.0;
            beta = c0;
#pragma unroll
            for (int t = 0; t < RPL; ++t) {
                const int r = l + WAVE * t;
                if (r > k && r < rows) {
                    A.set(k, t, 0.0);
                }
            }
        } else {
            beta = sqrt(c0 * c0 + tail_sq);
            if (c0 >= 0.0) {
                beta = -beta;
            }
            const double denom = c0 - beta;
#pragma unroll
            for (int t = 0; t < RPL; ++t) {
                const int r = l + WAVE * t;
                if (r > k && r < rows) {
                    A.set(k, t, A.get(k, t) / denom);
                }
            }
            tau = (beta - c0) / beta;
        }
        hc.set(k, tau);
        A.set_elem(k, k, beta);
        if (fabs(beta) > maxpivot) {
            maxpivot = fabs(beta);
        }

        // Apply H_k to the columns right of k.
        if (rows - k == 1) {
            for (int c = k + 1; c < cols; ++c) {
                A.set_elem(c, k, A.elem(c, k) * (1.0 - tau));
            }
        } else if (tau != 0.0) {
            for (int c = k + 1; c < cols; ++c) {
                double tmp = wave_sum_f<RPL>(k + 1, rows,
                                             [&](int t) { return A.get(k, t) * A.get(c, t); });
                const double top = A.elem(c, k);
                tmp = tmp + top;
                A.set_elem(c, k, top - tau * tmp);
#pragma unroll
                for (int t = 0; t < RPL; ++t) {
                    const int r = l + WAVE * t;
                    if (r > k && r < rows) {
                        A.set(c, t, A.get(c, t) - (tau * A.get(k, t)) * tmp);
                    }
                }
            }
        }

        // Norm downdate (LAWN 176).
        for (int c = k + 1; c < cols; ++c) {
            const double nuc = nu.get(c);
            if (nuc != 0.0) {
                double temp = fabs(A.elem(c, k)) / nuc;
                temp = (1.0 + temp) * (1.0 - temp);
                temp = temp < 0.0 ? 0.0 : temp;
                const double ratio = nuc / nd.get(c);
                const double temp2 = temp * (ratio * ratio);
                if (temp2 <= norm_downdate_threshold) {
                    const double s = wave_sum_f<RPL>(k + 1, rows, [&](int t) {
                        const double x = A.get(c, t);
                        return x * x;
                    });
                    const double nrm = sqrt(s);
                    nd.set(c, nrm);
                    nu.set(c, nrm);
                } else {
                    nu.set(c, nuc * sqrt(temp));
                }
            }
        }
    }

    // Column permutation from the transpositions.
    LaneVec<NSC, int> perm;
#pragma unroll
    for (int t = 0; t < NSC; ++t) {
        perm.v[t] = l + WAVE * t;
    }
    for (int k = 0; k < size; ++k) {
        const int other = trans.get(k);
        const int a = perm.get(k);
        const int b = perm.get(other);
        perm.set(k, b);
        perm.set(other, a);
    }

    const double threshold = static_cast<double>(rows > cols ? rows : cols) * DBL_EPS;
    const double premultiplied = fabs(maxpivot) * threshold;
    int rank = 0;
    for (int i = 0; i < nonzero_pivots; ++i) {
        rank += fabs(A.elem(i, i)) > premultiplied ? 1 : 0;
    }

    // Q^T d over the first `rank` reflectors, H_0 first.
    for (int k = 0; k < rank; ++k) {
        const double tau = hc.get(k);
        if (rows - k == 1) {
            d.set(k, d.get(k) * (1.0 - tau));
        } else if (tau != 0.0) {
            double tmp = wave_sum_f<RPL>(k + 1, rows, [&](int t) { return A.get(k, t) * d.v[t]; });
            const double top = d.get(k);
            tmp = tmp + top;
            d.set(k, top - tau * tmp);
#pragma unroll
            for (int t = 0; t < RPL; ++t) {
                const int r = l + WAVE * t;
                if (r > k && r < rows) {
                    d.v[t] = d.v[t] - (tau * A.get(k, t)) * tmp;
                }
            }
        }
    }
    // R11 b = (Q^T d)[0:rank], column-oriented back substitution.
    for (int i = rank - 1; i >= 0; --i) {
        double ri = d.get(i);
        if (ri != 0.0) {
            ri = ri / A.elem(i, i);
        }
        d.set(i, ri);
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            const int r = l + WAVE * t;
            if (r < i) {
                d.v[t] = d.v[t] - A.get(i, t) * ri;
            }
        }
    }
    coef.fill(0.0);
    for (int p = 0; p < rank; ++p) {
        coef.set(perm.get(p), d.get(p));
    }
}

// Default NNLS gradient tolerance of a design (nnls.cpp:default_gradient_tolerance):
// 10 eps max_c sum_r |C(r, c)| max(rows, cols), the column sums being wave sums.  P provides
// the columns: P.get(c, t) is column c at the lane's row slot t (c uniform).
template <int NQ, int RPL, class P>
__device__ __forceinline__ double default_tolerance_wave(const P& C, int rows, int cols) {
    double norm1 = 0.0;
    if (rows * cols > 0) {
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            if (c < cols) {
                const double s = wave_sum_f<RPL>(0, rows, [&](int t) { return fabs(C.get(c, t)); });
                if (c == 0 || s > norm1) {
                    norm1 = s;
                }
            }
        }
    }
    return 10.0 * DBL_EPS * norm1 * static_cast<double>(rows > cols ? rows : cols);
}

// The generic Lawson-Hanson NNLS on up to NQ columns (solve_nnls_with_limit).  C provides the
// columns (C.get(c, t), c uniform); each least-squares fit copies the active columns into the
// wave's LDS slab `work` (RPL * 64 doubles per column, ncols columns) and runs the pivoted QR
// there.  x receives the solution (uniform).  Returns false when the inner-step limit stopped it.
template <int NQ, int RPL, class P>
__device__ __forceinline__ bool nnls_small(const P& C, int rows, int ncols, const RowVec<RPL>& d,
                                           double tol, int max_inner, double (&x)[NQ],
                                           double* work) {
    bool active[NQ];
    double g[NQ];
    double cand[NQ];
#pragma unroll
    for (int c = 0; c < NQ; ++c) {
        active[c] = false;
        x[c] = 0.0;
    }
    auto gradient_at = [&](const double (&sol)[NQ], double (&out)[NQ]) {
        double r[RPL];
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < NQ; ++c) {
                if (c < ncols) {
                    s = s + C.get(c, t) * sol[c];
                }
            }
            r[t] = d.v[t] - s;
        }
        // Only an inactive column's gradient is ever read (the entering choice), so the
        // active ones are left at zero.
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            out[c] = 0.0;
            if (c < ncols && !active[c]) {
                out[c] = wave_sum_f<RPL>(0, rows, [&](int t) { return C.get(c, t) * r[t]; });
            }
        }
    };
    auto fit_on = [&](double (&out)[NQ]) {
        MemCols<RPL> A{work};
        int ka = 0;
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            out[c] = 0.0;
            if (c < ncols && active[c]) {
#pragma unroll
                for (int t = 0; t < RPL; ++t) {
                    A.set(ka, t, C.get(c, t));
                }
                ++ka;
            }
        }
        if (ka == 0) {
            return;
        }
        RowVec<RPL> dd = d;
        LaneVec<1> coef;
        least_squares_wave<1>(A, rows, ka, dd, coef);
        int p = 0;
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            if (c < ncols && active[c]) {
                out[c] = coef.get(p);
                ++p;
            }
        }
    };

    gradient_at(x, g);
    int n_inner = 0;
    int n_outer = 0;
    while (true) {
        // Each entering step adds a column; a solve still running after 8 n + 64 of them can
        // only be cycling on rounding, and stops (reported as not converged).
        if (++n_outer > 8 * ncols + 64) {
            return false;
        }
        int entering = -1;
        double best = 0.0;
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            if (c < ncols && !active[c] && (entering < 0 || g[c] > best)) {
                entering = c;
                best = g[c];
            }
        }
        if (entering < 0 || best <= tol) {
            break;
        }
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            active[c] = active[c] || c == entering;
        }
        fit_on(cand);
        while (true) {
            bool any = false;
#pragma unroll
            for (int c = 0; c < NQ; ++c) {
                any = any || (c < ncols && active[c] && cand[c] <= 0.0);
            }
            if (!any) {
                break;
            }
            if (++n_inner > max_inner) {
#pragma unroll
                for (int c = 0; c < NQ; ++c) {
                    x[c] = cand[c];
                }
                return false;
            }
            double step = __builtin_inf();
            int blocking = 0;
#pragma unroll
            for (int c = 0; c < NQ; ++c) {
                if (c < ncols && active[c] && cand[c] <= 0.0) {
                    const double distance = x[c] - cand[c];
                    const double s = distance > 0.0 ? x[c] / distance : 0.0;
                    if (s < step) {
                        step = s;
                        blocking = c;
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < NQ; ++c) {
                if (c < ncols) {
                    x[c] = x[c] + step * (cand[c] - x[c]);
                }
            }
#pragma unroll
            for (int c = 0; c < NQ; ++c) {
                if (c == blocking) {
                    x[c] = 0.0;
                    active[c] = false;
                }
                active[c] = active[c] && fabs(x[c]) >= tol;
            }
            fit_on(cand);
        }
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            x[c] = cand[c];
        }
        gradient_at(x, g);
    }
    return true;
}

} // namespace lsd
