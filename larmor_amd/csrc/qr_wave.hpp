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

namespace LSD_NS {

constexpr double DBL_EPS = 2.220446049250313e-16;
// The machine epsilon and the smallest normal number of the arithmetic type.
#ifdef LSD_FP32
constexpr float REAL_EPS = FLT_EPSILON;
constexpr float REAL_MIN = FLT_MIN;
#else
constexpr double REAL_EPS = DBL_EPS;
constexpr double REAL_MIN = DBL_MIN;
#endif

// Up to NQ columns of RPL row slots each, in registers.  Column indices are uniform; a column
// is picked by a compile-time-unrolled select over the NQ column vectors, so the columns stay
// separate register vectors.
template <int NQ, int RPL>
struct RegCols {
    static constexpr int kRpl = RPL;
    rvec<RPL> col[NQ];

    template <int... Q>
    __device__ __forceinline__ real get_(int c, int t, std::integer_sequence<int, Q...>) const {
        real x = real(0.0);
        ((x = c == Q ? static_cast<real>(col[Q][t]) : x), ...);
        return x;
    }
    template <int... Q>
    __device__ __forceinline__ void set_(int c, int t, real x, std::integer_sequence<int, Q...>) {
        ((col[Q][t] = c == Q ? x : static_cast<real>(col[Q][t])), ...);
    }
    __device__ __forceinline__ real get(int c, int t) const {
        return get_(c, t, std::make_integer_sequence<int, NQ>{});
    }
    __device__ __forceinline__ void set(int c, int t, real x) {
        set_(c, t, x, std::make_integer_sequence<int, NQ>{});
    }
    // Element (row r, column c), both uniform.
    __device__ __forceinline__ real elem(int c, int r) const {
        return readlane(get(c, r / WAVE), r % WAVE);
    }
    __device__ __forceinline__ void set_elem(int c, int r, real x) {
        if (lane_id() == r % WAVE) {
            set(c, r / WAVE, x);
        }
    }
    __device__ __forceinline__ void swap_cols(int a, int b) {
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            const real va = get(a, t);
            const real vb = get(b, t);
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
    real* base;

    __device__ __forceinline__ real get(int c, int t) const {
        return base[static_cast<long>(c) * kLd + t * WAVE + lane_id()];
    }
    __device__ __forceinline__ void set(int c, int t, real x) {
        base[static_cast<long>(c) * kLd + t * WAVE + lane_id()] = x;
    }
    // A uniform element, read as a vector load of its 64-row segment plus readlane (never
    // through the scalar cache, as the slab is written in the same kernel).
    __device__ __forceinline__ real elem(int c, int r) const {
        const real x = base[static_cast<long>(c) * kLd + (r / WAVE) * WAVE + lane_id()];
        return readlane(x, r % WAVE);
    }
    __device__ __forceinline__ void set_elem(int c, int r, real x) {
        if (lane_id() == r % WAVE) {
            base[static_cast<long>(c) * kLd + r] = x;
        }
    }
    __device__ __forceinline__ void swap_cols(int a, int b) {
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            const real va = get(a, t);
            const real vb = get(b, t);
            set(a, t, vb);
            set(b, t, va);
        }
    }
};

// The wave sum over rows lo <= r < hi of f(t), f evaluated on the lane's slot t.  f is
// evaluated on every slot (its reads stay inside the lane's slots) and the terms outside the
// range are left out by a select, so the lanes do not diverge.
template <int RPL, class F>
__device__ __forceinline__ real wave_sum_f(int lo, int hi, F f) {
    const int l = lane_id();
    real s = real(0.0);
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
        const int r = l + WAVE * t;
        const real term = f(t);
        const real sum = s + term;
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
    nu.fill(real(0.0));
    nd.fill(real(0.0));
    hc.fill(real(0.0));
    trans.fill(0);
    for (int c = 0; c < cols; ++c) {
        const real s = wave_sum_f<RPL>(0, rows, [&](int t) {
            const real x = A.get(c, t);
            return x * x;
        });
        const real nrm = sqrt(s);
        nd.set(c, nrm);
        nu.set(c, nrm);
    }
    real max_norm = real(0.0);
    for (int c = 0; c < cols; ++c) {
        const real v = nu.get(c);
        if (c == 0 || v > max_norm) {
            max_norm = v;
        }
    }
    const real scaled = max_norm * REAL_EPS;
    const real threshold_helper = (scaled * scaled) / static_cast<real>(rows);
    const real norm_downdate_threshold = sqrt(REAL_EPS);
    int nonzero_pivots = size;
    real maxpivot = real(0.0);

    for (int k = 0; k < size; ++k) {
        int biggest = k;
        real bv = nu.get(k);
        for (int c = k + 1; c < cols; ++c) {
            const real v = nu.get(c);
            if (v > bv) {
                bv = v;
                biggest = c;
            }
        }
        const real biggest_sq = bv * bv;
        if (nonzero_pivots == size &&
            biggest_sq < threshold_helper * static_cast<real>(rows - k)) {
            nonzero_pivots = k;
        }
        trans.set(k, biggest);
        if (k != biggest) {
            A.swap_cols(k, biggest);
            const real u1 = nu.get(k);
            const real u2 = nu.get(biggest);
            nu.set(k, u2);
            nu.set(biggest, u1);
            const real d1 = nd.get(k);
            const real d2 = nd.get(biggest);
            nd.set(k, d2);
            nd.set(biggest, d1);
        }

        // Householder vector of column k.
        const real c0 = A.elem(k, k);
        const real tail_sq = wave_sum_f<RPL>(k + 1, rows, [&](int t) {
            const real x = A.get(k, t);
            return x * x;
        });
        real tau;
        real beta;
        if (tail_sq <= REAL_MIN) {
            tau = real(0.0);
            beta = c0;
#pragma unroll
            for (int t = 0; t < RPL; ++t) {
                const int r = l + WAVE * t;
                if (r > k && r < rows) {
                    A.set(k, t, real(0.0));
                }
            }
        } else {
            beta = sqrt(c0 * c0 + tail_sq);
            if (c0 >= real(0.0)) {
                beta = -beta;
            }
            const real denom = c0 - beta;
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
                A.set_elem(c, k, A.elem(c, k) * (real(1.0) - tau));
            }
        } else if (tau != real(0.0)) {
            for (int c = k + 1; c < cols; ++c) {
                real tmp = wave_sum_f<RPL>(k + 1, rows,
                                             [&](int t) { return A.get(k, t) * A.get(c, t); });
                const real top = A.elem(c, k);
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
            const real nuc = nu.get(c);
            if (nuc != real(0.0)) {
                real temp = fabs(A.elem(c, k)) / nuc;
                temp = (real(1.0) + temp) * (real(1.0) - temp);
                temp = temp < real(0.0) ? real(0.0) : temp;
                const real ratio = nuc / nd.get(c);
                const real temp2 = temp * (ratio * ratio);
                if (temp2 <= norm_downdate_threshold) {
                    const real s = wave_sum_f<RPL>(k + 1, rows, [&](int t) {
                        const real x = A.get(c, t);
                        return x * x;
                    });
                    const real nrm = sqrt(s);
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

    const real threshold = static_cast<real>(rows > cols ? rows : cols) * REAL_EPS;
    const real premultiplied = fabs(maxpivot) * threshold;
    int rank = 0;
    for (int i = 0; i < nonzero_pivots; ++i) {
        rank += fabs(A.elem(i, i)) > premultiplied ? 1 : 0;
    }

    // Q^T d over the first `rank` reflectors, H_0 first.
    for (int k = 0; k < rank; ++k) {
        const real tau = hc.get(k);
        if (rows - k == 1) {
            d.set(k, d.get(k) * (real(1.0) - tau));
        } else if (tau != real(0.0)) {
            real tmp = wave_sum_f<RPL>(k + 1, rows, [&](int t) { return A.get(k, t) * d.v[t]; });
            const real top = d.get(k);
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
        real ri = d.get(i);
        if (ri != real(0.0)) {
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
    coef.fill(real(0.0));
    for (int p = 0; p < rank; ++p) {
        coef.set(perm.get(p), d.get(p));
    }
}

// Default NNLS gradient tolerance of a design (nnls.cpp:default_gradient_tolerance):
// 10 eps max_c sum_r |C(r, c)| max(rows, cols), the column sums being wave sums.  P provides
// the columns: P.get(c, t) is column c at the lane's row slot t (c uniform).
template <int NQ, int RPL, class P>
__device__ __forceinline__ real default_tolerance_wave(const P& C, int rows, int cols) {
    real norm1 = real(0.0);
    if (rows * cols > 0) {
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            if (c < cols) {
                const real s = wave_sum_f<RPL>(0, rows, [&](int t) { return fabs(C.get(c, t)); });
                if (c == 0 || s > norm1) {
                    norm1 = s;
                }
            }
        }
    }
    return real(10.0) * REAL_EPS * norm1 * static_cast<real>(rows > cols ? rows : cols);
}

// The generic Lawson-Hanson NNLS on up to NQ columns (solve_nnls_with_limit).  C provides the
// columns (C.get(c, t), c uniform); each least-squares fit copies the active columns into the
// wave's LDS slab `work` (RPL * 64 doubles per column, ncols columns) and runs the pivoted QR
// there.  x receives the solution (uniform).  Returns false when the inner-step limit stopped it.
template <int NQ, int RPL, class P>
__device__ __forceinline__ bool nnls_small(const P& C, int rows, int ncols, const RowVec<RPL>& d,
                                           real tol, int max_inner, real (&x)[NQ],
                                           real* work) {
    bool active[NQ];
    real g[NQ];
    real cand[NQ];
#pragma unroll
    for (int c = 0; c < NQ; ++c) {
        active[c] = false;
        x[c] = real(0.0);
    }
    auto gradient_at = [&](const real (&sol)[NQ], real (&out)[NQ]) {
        real r[RPL];
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            real s = real(0.0);
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
            out[c] = real(0.0);
            if (c < ncols && !active[c]) {
                out[c] = wave_sum_f<RPL>(0, rows, [&](int t) { return C.get(c, t) * r[t]; });
            }
        }
    };
    auto fit_on = [&](real (&out)[NQ]) {
        MemCols<RPL> A{work};
        int ka = 0;
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            out[c] = real(0.0);
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
        real best = real(0.0);
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
                any = any || (c < ncols && active[c] && cand[c] <= real(0.0));
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
            real step = __builtin_inff();
            int blocking = 0;
#pragma unroll
            for (int c = 0; c < NQ; ++c) {
                if (c < ncols && active[c] && cand[c] <= real(0.0)) {
                    const real distance = x[c] - cand[c];
                    const real s = distance > real(0.0) ? x[c] / distance : real(0.0);
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
                    x[c] = real(0.0);
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

// The generic Lawson-Hanson NNLS on up to NQ columns through the normal equations: C^T C and
// C^T d are wave sums formed once; each least-squares fit on the active columns is a Cholesky
// solve of that Gram matrix's active block (uniform values), the gradient C^T d - (C^T C) x.  A
// column whose pivot is not above MIN_NORMAL_PIVOT of its diagonal is left out of the fit (its
// coefficient 0), as the pivoted QR leaves out columns past the numerical rank.  The steps, tie
// rules and limits are nnls_small's; only the fits differ, by rounding (the mixed-precision form's
// fraction fits: a few well-conditioned columns).
constexpr double MIN_NORMAL_PIVOT = 1e-12;

template <int NQ, int RPL, class P>
__device__ __forceinline__ bool nnls_normal(const P& C, int rows, int ncols, const RowVec<RPL>& d,
                                            real tol, int max_inner, real (&x)[NQ]) {
    real gram[NQ][NQ];
    real rhs[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        rhs[i] = real(0);
#pragma unroll
        for (int j = 0; j < NQ; ++j) {
            gram[i][j] = real(0);
        }
        if (i < ncols) {
            rhs[i] = wave_sum_f<RPL>(0, rows, [&](int t) { return C.get(i, t) * d.v[t]; });
#pragma unroll
            for (int j = 0; j <= i; ++j) {
                const real gij = wave_sum_f<RPL>(0, rows, [&](int t) {
                    return C.get(i, t) * C.get(j, t);
                });
                gram[i][j] = gij;
                gram[j][i] = gij;
            }
        }
    }
    bool active[NQ];
    real g[NQ];
    real cand[NQ];
#pragma unroll
    for (int c = 0; c < NQ; ++c) {
        active[c] = false;
        x[c] = real(0);
        g[c] = c < ncols ? rhs[c] : real(0);
    }
    auto gradient_at = [&]() {
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            real v = rhs[i];
#pragma unroll
            for (int j = 0; j < NQ; ++j) {
                v = v - gram[i][j] * x[j];
            }
            g[i] = i < ncols ? v : real(0);
        }
    };
    // cand = the least-squares fit on the active columns (0 elsewhere), by a Cholesky factor of
    // the active block built in column order.
    auto fit_on = [&]() {
        real l[NQ][NQ];
        bool used[NQ];
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            used[i] = false;
            cand[i] = real(0);
#pragma unroll
            for (int j = 0; j < NQ; ++j) {
                l[i][j] = real(0);
            }
        }
#pragma unroll
        for (int j = 0; j < NQ; ++j) {
            if (!(j < ncols && active[j])) {
                continue;
            }
            real pivot = gram[j][j];
#pragma unroll
            for (int k = 0; k < NQ; ++k) {
                if (k < j && used[k]) {
                    pivot = pivot - l[j][k] * l[j][k];
                }
            }
            if (!(pivot > real(MIN_NORMAL_PIVOT) * gram[j][j])) {
                continue;
            }
            const real ljj = sqrt(pivot);
            l[j][j] = ljj;
            used[j] = true;
#pragma unroll
            for (int i = 0; i < NQ; ++i) {
                if (i > j && i < ncols && active[i]) {
                    real v = gram[i][j];
#pragma unroll
                    for (int k = 0; k < NQ; ++k) {
                        if (k < j && used[k]) {
                            v = v - l[i][k] * l[j][k];
                        }
                    }
                    l[i][j] = v / ljj;
                }
            }
        }
        real y[NQ];
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            y[i] = real(0);
            if (used[i]) {
                real v = rhs[i];
#pragma unroll
                for (int k = 0; k < NQ; ++k) {
                    if (k < i && used[k]) {
                        v = v - l[i][k] * y[k];
                    }
                }
                y[i] = v / l[i][i];
            }
        }
#pragma unroll
        for (int i = NQ - 1; i >= 0; --i) {
            if (used[i]) {
                real v = y[i];
#pragma unroll
                for (int k = 0; k < NQ; ++k) {
                    if (k > i && used[k]) {
                        v = v - l[k][i] * cand[k];
                    }
                }
                cand[i] = v / l[i][i];
            }
        }
    };
    int n_inner = 0;
    int n_outer = 0;
    while (true) {
        if (++n_outer > 8 * ncols + 64) {
            return false;
        }
        int entering = -1;
        real best = real(0);
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
        fit_on();
        while (true) {
            bool any = false;
#pragma unroll
            for (int c = 0; c < NQ; ++c) {
                any = any || (c < ncols && active[c] && cand[c] <= real(0));
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
            real step = __builtin_inff();
            int blocking = 0;
#pragma unroll
            for (int c = 0; c < NQ; ++c) {
                if (c < ncols && active[c] && cand[c] <= real(0)) {
                    const real distance = x[c] - cand[c];
                    const real sv = distance > real(0) ? x[c] / distance : real(0);
                    if (sv < step) {
                        step = sv;
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
                    x[c] = real(0);
                    active[c] = false;
                }
                active[c] = active[c] && fabs(x[c]) >= tol;
            }
            fit_on();
        }
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            x[c] = cand[c];
        }
        gradient_at();
    }
    return true;
}

} // namespace LSD_NS
