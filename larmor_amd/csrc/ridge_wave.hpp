// The ridge-regularised Lawson-Hanson NNLS of one voxel on one wavefront.
//
// Restates larmor's solve_ridge_nnls (cpp/src/sd/nnls.cpp:294-350) with its Lawson-Hanson loop
// (nnls.cpp:120-153) and ActiveCholesky (nnls.cpp:169-233):
//  - the gradient is C^T (d - C x) = K^T s - l1 - (K^T K + r^2 I) x, summed over the non-zero
//    coefficients in column order (nnls.cpp:305-313);
//  - the entering column is the inactive one of largest gradient, the first on a tie; the loop
//    ends when it does not exceed the tolerance (a NaN gradient, which only non-finite signals
//    produce, is never chosen);
//  - least-squares fits on the active set come from the Cholesky factor L of the Gram matrix on
//    the active columns, in the order they joined: a joining column appends the row
//    L w = G_A,j, sqrt(G_jj - w.w), refused when that pivot is not above 1e-12 G_jj
//    (nnls.cpp:211-227); a leaving column is removed by shifting the rows below it up and
//    clearing the entry above the diagonal with Givens rotations (remove_cholesky_column,
//    nnls.cpp:237-256);
//  - once a pivot is refused, and always when the design is not well conditioned enough for the
//    normal equations, the fits are pivoted-QR least squares on [K_A; r I] (nnls.cpp:317-338);
//  - inner steps move x towards the candidate until the first active coefficient reaches zero
//    (smallest x / (x - candidate), the first on a tie), drop it and any |x| below the
//    tolerance, and refit; more than 3 n of them end the solve with the last candidate.
//
// Layout: solver columns are spread over lanes as LaneVec slots (column c in lane c % 64, slot
// c / 64: NCPL slots); factor positions likewise (NSK slots).  L lives column-major with an odd
// stride KCAP + 1 (conflict-free for both its row and column walks), in LDS for the fast path or
// in a global scratch slab for the rerun of a voxel whose active set outgrows the LDS factor.
#pragma once

#include "qr_wave.hpp"

namespace LSD_NS {

// Status bits of a voxel solve.
constexpr int ST_NOT_CONVERGED = 1; // the inner-step limit ended an NNLS (result kept)
constexpr int ST_OVERFLOW = 2;      // the active set outgrew the factor capacity: rerun
constexpr int ST_QR_FALLBACK = 4;   // a Cholesky pivot was refused; the QR path finished
constexpr int ST_OUTER_GUARD = 8;   // more entering steps than any finite solve takes: stopped

// Entering steps allowed before a solve is declared cycling.  Each entering step adds a column
// and the reference has no limit; a solve that has not ended after this many can only be
// cycling on rounding, which would never end, so it stops with the current solution.
__device__ __forceinline__ int outer_guard(int n) { return 8 * n + 64; }

constexpr real MIN_RELATIVE_PIVOT = real(1e-12);

// The shared matrices of one ridge design.
struct RidgeDesignDev {
    int m;                // measurements
    int n;                // columns
    const real* kernel; // (m, n) row-major, the weighted anisotropic kernel
    const real* gram;   // (n, n) row-major, K^T K + ridge^2 I, bitwise symmetric
    real ridge;
    int normal_equations; // 1: Cholesky path (with QR fallback); 0: QR only
};

// Rows per lane of the QR fallback's stacked matrix [K_A; r I].
template <int KCAP>
constexpr int qr_rows_per_lane(int max_m) {
    return (max_m + KCAP + WAVE - 1) / WAVE;
}

// QR: whether this instantiation carries the pivoted-QR fits.  The fast (LDS) kernels do not:
// a voxel whose factor refuses a pivot, or whose design needs QR fits throughout, is handed to
// the rerun kernel (reported as ST_OVERFLOW), which carries them, so the fast kernels keep their
// registers for the common path.
// RESID: the gradient is C^T (d - C x) over the measurements (the generic NNLS,
// solve_nnls_with_limit nnls.cpp:264-278: the residual's in-column-order sum, then each column's
// wave sum) instead of corr - G x; such a design has no Gram matrix and fits by pivoted QR
// (normal_equations 0, ridge 0).
template <int NCPL, int KCAP, int RPLQ, bool QR = true, bool RESID = false>
struct RidgeNnlsWave {
    static constexpr int NSK = (KCAP + WAVE - 1) / WAVE; // factor-position slots
    static constexpr int S = KCAP + 1;                   // factor column stride
    static constexpr bool PREFETCH = QR;                 // factor in global memory

    const RidgeDesignDev& D;
    real tol;
    int max_inner;
    real l1;
    real* L;   // factor storage, S * KCAP doubles
    real* qrs; // QR fallback slab: (RPLQ * 64) * KCAP doubles, then RPLQ * 64 for the data

    // Per-lane state.
    real x[NCPL];
    real g[NCPL];
    real cand[NCPL];
    bool act[NCPL];
    int pos[NCPL];     // factor position of the column, -1 when not in the factor
    LaneVec<NSK, int> fcol; // column at each factor position
    int k;             // factor size (uniform)
    bool factor_failed;
    // y = L^-1 corr_A, valid on positions [0, y_valid): kept across appends (the new row of L
    // gives y_k by one sequential dot, the order the full column-oriented solve uses) and cut
    // back to the removed position by a downdate.
    LaneVec<NSK> y;
    int y_valid;
    // diag[2 p], diag[2 p + 1]: the factor's diagonal at position p and its division reciprocal
    // (after the factor and its slack); diag_ok: every diagonal in the divisor range.
    real* diag;
    bool diag_ok;
    // The solve's FP64 operations, counted by formula as its steps run (uniform): the
    // algorithmic work that profiled launches report.
    unsigned long long work;

    __device__ RidgeNnlsWave(const RidgeDesignDev& design, real tol_, int max_inner_, real l1_,
                             real* L_, real* qrs_)
        : D(design), tol(tol_), max_inner(max_inner_), l1(l1_), L(L_), qrs(qrs_),
          diag(L_ + S * KCAP + WAVE) {}

    // L_jj and its reciprocal (one 16-byte read of a uniform address).
    __device__ __forceinline__ real2 diag_at(int j) const {
        return *reinterpret_cast<const real2*>(diag + 2 * j);
    }
    __device__ __forceinline__ int col(int t) const { return lane_id() + WAVE * t; }
    __device__ __forceinline__ real& Lat(int i, int j) { return L[j * S + i]; }
    __device__ __forceinline__ real Luni(int i, int j) const {
        // A uniform factor entry, read as a vector access of the lane's neighbour + readlane.
        const int base = j * S + (i / WAVE) * WAVE;
        return readlane(L[base + lane_id()], i % WAVE);
    }

    // The column-oriented triangular solves update every lane's slot with a select instead of a
    // branch: the factor read L(i, j) of a lane outside the update stays inside the slab (rows
    // up to 64 NSK - 1 of a column lie within the column stride and the slack after the last
    // column) and its product is discarded.

    // a != 0 ? a / L_jj : a from d = (L_jj, its reciprocal); FAST: every diagonal is in the
    // divisor range.
    template <bool FAST>
    __device__ __forceinline__ real divide_by_diag(real a, real2 d) const {
        if (FAST && dividend_in_range(a)) {
            return divide_in_range(a, d.x, d.y);
        }
        return a != real(0.0) ? a / d.x : a;
    }

    // L w = b on positions [p0, kk) when positions [0, p0) of b already hold their final values
    // (p0 = 0: the whole solve); returns the sum of the w_j^2 in order (for p0 = 0).  Positions
    // from p0 on receive the same updates, in the same order, as in the whole solve.  With the
    // factor in global memory (PREFETCH), step j's reads (the diagonal entry and its reciprocal,
    // the factor column) are issued during step j - 1, so their latency is off the chain of
    // dependent steps; from LDS the plain order issues fewer instructions and is faster.
    template <bool FAST>
    __device__ __forceinline__ real forward_impl(LaneVec<NSK>& b, int p0, int kk) {
        const int l = lane_id();
        real sq = real(0.0);
        real2 d_next = diag_at(0);
        real l_next[NSK];
        if constexpr (PREFETCH) {
#pragma unroll
            for (int u = 0; u < NSK; ++u) {
                l_next[u] = Lat(l + WAVE * u, 0);
            }
        }
        for (int j = 0; j < kk; ++j) {
            real2 d;
            real lc[NSK];
            if constexpr (PREFETCH) {
                d = d_next;
#pragma unroll
                for (int u = 0; u < NSK; ++u) {
                    lc[u] = l_next[u];
                }
                const int jn = j + 1 < kk ? j + 1 : j;
                d_next = diag_at(jn);
#pragma unroll
                for (int u = 0; u < NSK; ++u) {
                    l_next[u] = Lat(l + WAVE * u, jn);
                }
            } else {
                d = diag_at(j);
#pragma unroll
                for (int u = 0; u < NSK; ++u) {
                    lc[u] = Lat(l + WAVE * u, j);
                }
            }
            const real bj = b.get(j);
            const real wj = j >= p0 ? divide_by_diag<FAST>(bj, d) : bj;
            sq = fma(wj, wj, sq);
#pragma unroll
            for (int u = 0; u < NSK; ++u) {
                const int i = l + WAVE * u;
                const real updated = fma(-lc[u], wj, b.v[u]);
                const bool below = i > j && i < kk && i >= p0;
                b.v[u] = i == j ? wj : (below ? updated : b.v[u]);
            }
        }
        return sq;
    }

    __device__ __forceinline__ real forward(LaneVec<NSK>& b, int kk_) {
        const int kk = __builtin_amdgcn_readfirstlane(kk_);
        return diag_ok ? forward_impl<true>(b, 0, kk) : forward_impl<false>(b, 0, kk);
    }

    __device__ __forceinline__ void forward_from(LaneVec<NSK>& b, int p0_, int kk_) {
        const int p0 = __builtin_amdgcn_readfirstlane(p0_);
        const int kk = __builtin_amdgcn_readfirstlane(kk_);
        if (p0 >= kk) {
            return;
        }
        if (diag_ok) {
            forward_impl<true>(b, p0, kk);
        } else {
            forward_impl<false>(b, p0, kk);
        }
    }

    // L^T z = w on positions [0, kk), column-oriented.  A lane past the factor reads row 0 of
    // column j instead of the row its index names (outside the slab).  As in the forward solve,
    // with PREFETCH step j's reads are issued during step j + 1.
    template <bool FAST>
    __device__ __forceinline__ void backward_impl(LaneVec<NSK>& b, int kk) {
        const int l = lane_id();
        const int top = kk > 0 ? kk - 1 : 0;
        real2 d_next = diag_at(top);
        real l_next[NSK];
        if constexpr (PREFETCH) {
#pragma unroll
            for (int u = 0; u < NSK; ++u) {
                const int i = l + WAVE * u;
                l_next[u] = Lat(top, i < KCAP ? i : 0);
            }
        }
        for (int j = kk - 1; j >= 0; --j) {
            real2 d;
            real lc[NSK];
            if constexpr (PREFETCH) {
                d = d_next;
#pragma unroll
                for (int u = 0; u < NSK; ++u) {
                    lc[u] = l_next[u];
                }
                const int jn = j > 0 ? j - 1 : 0;
                d_next = diag_at(jn);
#pragma unroll
                for (int u = 0; u < NSK; ++u) {
                    const int i = l + WAVE * u;
                    l_next[u] = Lat(jn, i < KCAP ? i : 0);
                }
            } else {
                d = diag_at(j);
#pragma unroll
                for (int u = 0; u < NSK; ++u) {
                    const int i = l + WAVE * u;
                    lc[u] = Lat(j, i < KCAP ? i : 0);
                }
            }
            const real wj = b.get(j);
            const real zj = divide_by_diag<FAST>(wj, d);
#pragma unroll
            for (int u = 0; u < NSK; ++u) {
                const int i = l + WAVE * u;
                const real updated = fma(-lc[u], zj, b.v[u]);
                b.v[u] = i == j ? zj : (i < j ? updated : b.v[u]);
            }
        }
    }

    __device__ __forceinline__ void backward(LaneVec<NSK>& b, int kk_) {
        const int kk = __builtin_amdgcn_readfirstlane(kk_);
        if (diag_ok) {
            backward_impl<true>(b, kk);
        } else {
            backward_impl<false>(b, kk);
        }
    }

    // The value of a column-spread vector at column c for each lane's factor position.
    __device__ __forceinline__ real gather_cols(const real (&v)[NCPL], int c) const {
        const int src = c % WAVE;
        const int slot = c / WAVE;
        real out = real(0.0);
#pragma unroll
        for (int t = 0; t < NCPL; ++t) {
            const real w = gather(v[t], src);
            out = slot == t ? w : out;
        }
        return out;
    }

    // The value of a column-spread vector at uniform column c.
    __device__ __forceinline__ real uniform_col(const real (&v)[NCPL], int c) const {
        real x = real(0.0);
#pragma unroll
        for (int t = 0; t < NCPL; ++t) {
            x = (c / WAVE) == t ? v[t] : x;
        }
        return readlane(x, c % WAVE);
    }

    // Appends column c to the factor; false when its pivot is refused.  With y valid on every
    // position, also extends y to the new one.
    __device__ __forceinline__ bool append(int c, const real (&corr)[NCPL]) {
        const int l = lane_id();
        LaneVec<NSK> b;
#pragma unroll
        for (int u = 0; u < NSK; ++u) {
            const int p = l + WAVE * u;
            b.v[u] = p < k ? D.gram[static_cast<long>(c) * D.n + fcol.v[u]] : real(0.0);
        }
        const real sq = forward(b, k);
        work += static_cast<unsigned long long>(k) * k + 4ull * k + 4ull;
        const real gcc = D.gram[static_cast<long>(c) * D.n + c];
        const real pivot = gcc - sq;
        if (!(pivot > MIN_RELATIVE_PIVOT * gcc)) {
            return false;
        }
#pragma unroll
        for (int u = 0; u < NSK; ++u) {
            const int p = l + WAVE * u;
            if (p < k) {
                Lat(k, p) = b.v[u];
            }
        }
        const real lkk = sqrt(pivot);
        const real rkk = division_reciprocal(lkk);
        if (l == k % WAVE) {
            Lat(k, k) = lkk;
            diag[2 * k] = lkk;
            diag[2 * k + 1] = rkk;
        }
        const bool lkk_ok = divisor_in_range(lkk);
        if (y_valid == k) {
            // y_k = (corr_c - L_k0 y_0 - L_k1 y_1 - ...) / L_kk, each term a fused
            // multiply-subtract in order, as the column-oriented solve applies them.
            const real cc = uniform_col(corr, c);
            real v = cc;
            for (int j = 0; j < k; ++j) {
                v = fma(-b.get(j), y.get(j), v);
            }
            y.set(k, divide_nonzero_by(v, lkk, rkk, lkk_ok));
            y_valid = k + 1;
        }
        fcol.set(k, c);
#pragma unroll
        for (int t = 0; t < NCPL; ++t) {
            pos[t] = col(t) == c ? k : pos[t];
        }
        diag_ok = diag_ok && lkk_ok;
        k += 1;
        return true;
    }

    // Removes the factor position p (remove_cholesky_column).
    __device__ __forceinline__ void remove(int p) {
        const int l = lane_id();
        // Shift rows p..k-2 up by one; only entries on or left of the superdiagonal are read
        // again.  Slots run in increasing order so each row is read before it is overwritten.
        for (int c = 0; c < k; ++c) {
#pragma unroll
            for (int u = 0; u < NSK; ++u) {
                const int i = l + WAVE * u;
                const bool move = i >= p && i < k - 1 && c <= i + 1;
                const real v = move ? Lat(i + 1, c) : real(0.0);
                if (move) {
                    Lat(i, c) = v;
                }
            }
        }
        for (int column = p; column < k - 1; ++column) {
            const real diagonal = Luni(column, column);
            const real above = Luni(column, column + 1);
            const real length = hypot_fixed(diagonal, above);
            work += 8ull + 6ull * static_cast<unsigned long long>(k - 1 - column);
            const real rl = division_reciprocal(length);
            const real cosine = divide_by(diagonal, length, rl);
            const real sine = divide_by(above, length, rl);
#pragma unroll
            for (int u = 0; u < NSK; ++u) {
                const int i = l + WAVE * u;
                if (i >= column && i < k - 1) {
                    const real left = Lat(i, column);
                    const real right = Lat(i, column + 1);
                    Lat(i, column) = cosine * left + sine * right;
                    Lat(i, column + 1) = cosine * right - sine * left;
                }
            }
        }
        // The rotated diagonal and its reciprocals (each lane its own positions).
        bool out_of_range = false;
#pragma unroll
        for (int u = 0; u < NSK; ++u) {
            const int i = l + WAVE * u;
            if (i < k - 1) {
                real d = diag[2 * i];
                if (i >= p) {
                    d = Lat(i, i);
                    diag[2 * i] = d;
                    diag[2 * i + 1] = division_reciprocal(d);
                }
                out_of_range = out_of_range || !divisor_in_range(d);
            }
        }
        diag_ok = ballot(out_of_range) == 0ull;
        // Positions above p move down one.
        LaneVec<NSK, int> shifted;
#pragma unroll
        for (int u = 0; u < NSK; ++u) {
            const int i = l + WAVE * u;
            // the column at position i + 1
            const int src_lane = (i + 1) % WAVE;
            const int src_slot = (i + 1) / WAVE;
            int from = 0;
#pragma unroll
            for (int s2 = 0; s2 < NSK; ++s2) {
                const int w = gather(fcol.v[s2], src_lane);
                from = src_slot == s2 ? w : from;
            }
            shifted.v[u] = (i >= p && i < k - 1) ? from : fcol.v[u];
        }
        fcol = shifted;
#pragma unroll
        for (int t = 0; t < NCPL; ++t) {
            pos[t] = pos[t] == p ? -1 : (pos[t] > p ? pos[t] - 1 : pos[t]);
        }
        k -= 1;
        y_valid = y_valid < p ? y_valid : p;
    }

    // G_AA z = corr_A from the factor, scattered to the columns (zero elsewhere).
    __device__ __forceinline__ void solve_factor(const real (&corr)[NCPL]) {
        const int l = lane_id();
        LaneVec<NSK> b;
#pragma unroll
        for (int u = 0; u < NSK; ++u) {
            const int p = l + WAVE * u;
            const real v = gather_cols(corr, p < k ? fcol.v[u] : 0);
            b.v[u] = p < k ? (p < y_valid ? static_cast<real>(y.v[u]) : v) : real(0.0);
        }
        if (y_valid < k) {
            const unsigned long long fresh = static_cast<unsigned long long>(k - y_valid);
            work += fresh * fresh + 2ull * static_cast<unsigned long long>(y_valid) * fresh;
        }
        work += static_cast<unsigned long long>(k) * k;
        forward_from(b, y_valid, k);
        y = b;
        y_valid = k;
        backward(b, k);
#pragma unroll
        for (int t = 0; t < NCPL; ++t) {
            const int p = pos[t] < 0 ? 0 : pos[t];
            real z = real(0.0);
#pragma unroll
            for (int u = 0; u < NSK; ++u) {
                const real w = gather(b.v[u], p % WAVE);
                z = (p / WAVE) == u ? w : z;
            }
            cand[t] = pos[t] >= 0 ? z : real(0.0);
        }
    }

    // Least squares on [K_A; r I] by pivoted QR, the fallback fit (nnls.cpp fit_by_qr).
    // Returns false on overflow (more active columns than the slab holds).
    template <int RPL>
    __device__ __forceinline__ bool solve_qr(const RowVec<RPL>& signal) {
        const int l = lane_id();
        // Active columns in increasing order.
        LaneVec<NSK, int> idx;
        idx.fill(0);
        int ka = 0;
#pragma unroll
        for (int t = 0; t < NCPL; ++t) {
            unsigned long long mask = ballot(col(t) < D.n && act[t]);
            while (mask) {
                const int b = __ffsll(static_cast<long long>(mask)) - 1;
                mask &= mask - 1;
                if (ka >= KCAP) {
                    return false;
                }
                idx.set(ka, WAVE * t + b);
                ++ka;
            }
        }
#pragma unroll
        for (int t = 0; t < NCPL; ++t) {
            cand[t] = real(0.0);
        }
        if (ka == 0) {
            return true;
        }
        const bool has_ridge = D.ridge > real(0.0);
        const int rows = D.m + (has_ridge ? ka : 0);
        MemCols<RPLQ> A{qrs};
        RowVec<RPLQ> d;
        for (int p = 0; p < ka; ++p) {
            const int c = idx.get(p);
#pragma unroll
            for (int t = 0; t < RPLQ; ++t) {
                const int r = l + WAVE * t;
                real v = real(0.0);
                if (r < D.m) {
                    v = D.kernel[static_cast<long>(r) * D.n + c];
                } else if (has_ridge && r == D.m + p) {
                    v = D.ridge;
                }
                A.set(p, t, v);
            }
        }
#pragma unroll
        for (int t = 0; t < RPLQ; ++t) {
            const int r = l + WAVE * t;
            real v = real(0.0);
            if (r < D.m) {
#pragma unroll
                for (int s2 = 0; s2 < RPL; ++s2) {
                    v = (s2 == t) ? signal.v[s2] : v;
                }
            } else if (has_ridge && r < rows) {
                v = -l1 / D.ridge;
            }
            d.v[t] = v;
        }
        LaneVec<NSK> coef;
        // Householder QR of the rows x ka stack and the projection of the data.
        work += 2ull * static_cast<unsigned long long>(rows) * ka * ka +
                4ull * static_cast<unsigned long long>(rows) * ka;
        least_squares_wave<NSK>(A, rows, ka, d, coef);
        for (int p = 0; p < ka; ++p) {
            const int c = idx.get(p);
            const real v = coef.get(p);
#pragma unroll
            for (int t = 0; t < NCPL; ++t) {
                cand[t] = col(t) == c ? v : cand[t];
            }
        }
        return true;
    }

    // fit_on: brings the factor to the active set and solves; false on overflow.
    template <int RPL>
    __device__ __forceinline__ bool fit_on(const real (&corr)[NCPL], const RowVec<RPL>& signal, int& status) {
        const int l = lane_id();
        if (!factor_failed) {
            // Remove the positions whose column left, last to first.
            int actbits = 0;
#pragma unroll
            for (int t = 0; t < NCPL; ++t) {
                actbits |= act[t] ? (1 << t) : 0;
            }
            for (int u = NSK - 1; u >= 0; --u) {
                const int p0 = l + WAVE * u;
                const int c = p0 < k ? fcol.v[u] : 0;
                const int bits = gather(actbits, c % WAVE);
                const bool leaves = p0 < k && !((bits >> (c / WAVE)) & 1);
                unsigned long long mask = ballot(leaves);
                while (mask) {
                    const int b = 63 - __clzll(static_cast<long long>(mask));
                    mask &= ~(1ull << b);
                    remove(WAVE * u + b);
                }
            }
            // Append the active columns not in the factor, in increasing order.
            bool failed = false;
#pragma unroll
            for (int t = 0; t < NCPL; ++t) {
                unsigned long long mask = ballot(col(t) < D.n && act[t] && pos[t] < 0);
                while (mask && !failed) {
                    const int b = __ffsll(static_cast<long long>(mask)) - 1;
                    mask &= mask - 1;
                    if (k >= KCAP) {
                        status |= ST_OVERFLOW;
                        return false;
                    }
                    if (!append(WAVE * t + b, corr)) {
                        failed = true;
                    }
                }
            }
            if (failed) {
                factor_failed = true;
                status |= ST_QR_FALLBACK;
            }
        }
        if (factor_failed) {
            if constexpr (QR) {
                if (!this->template solve_qr<RPL>(signal)) {
                    status |= ST_OVERFLOW;
                    return false;
                }
                return true;
            } else {
                status |= ST_OVERFLOW;
                return false;
            }
        }
        solve_factor(corr);
        return true;
    }

    // Runs the solve; corr = K^T s - l1 per column, signal = s (for the QR fallback).
    // Returns the status bits; x holds the solution.
    template <int RPL>
    __device__ __forceinline__ int run(const real (&corr)[NCPL], const RowVec<RPL>& signal,
                                       unsigned long long warm = 0ull) {
        const int l = lane_id();
        int status = 0;
        k = 0;
        y_valid = 0;
        y.fill(real(0.0));
        diag_ok = true;
        work = 0;
        factor_failed = D.normal_equations == 0;
#pragma unroll
        for (int t = 0; t < NCPL; ++t) {
            x[t] = real(0.0);
            g[t] = corr[t];
            act[t] = false;
            pos[t] = -1;
            cand[t] = real(0.0);
        }
        fcol.fill(0);
        int n_inner = 0;
        int n_outer = 0;
        if (NCPL == 1 && warm != 0ull) {
            // Warm start from a previous solve's active set (columns < 64): fit on it, dropping
            // the columns whose coefficient is not positive until the fit is, and start the
            // Lawson-Hanson loop from that feasible point with its gradient.  The loop then ends
            // at the same optimum as from x = 0 (to rounding), usually after no entering step.
            act[0] = col(0) < D.n && ((warm >> col(0)) & 1ull) != 0ull;
            while (ballot(act[0]) != 0ull) {
                if (!this->template fit_on<RPL>(corr, signal, status)) {
                    return status;
                }
                const bool drop = act[0] && !(cand[0] > real(0));
                if (ballot(drop) == 0ull) {
                    break;
                }
                act[0] = act[0] && !drop;
            }
            x[0] = act[0] ? cand[0] : real(0);
            gradient_at_x<RPL>(corr, signal);
        }
        while (true) {
            if (++n_outer > outer_guard(D.n)) {
                status |= ST_OUTER_GUARD | ST_NOT_CONVERGED;
                break;
            }
            bool free_col[NCPL];
#pragma unroll
            for (int t = 0; t < NCPL; ++t) {
                free_col[t] = col(t) < D.n && !act[t];
            }
            real bv = real(0.0);
            int bi = -1;
            wave_argmax_cols<NCPL>(g, free_col, bv, bi);
            if (bi < 0 || bv <= tol) {
                break;
            }
#pragma unroll
            for (int t = 0; t < NCPL; ++t) {
                act[t] = act[t] || col(t) == bi;
            }
            if (!this->template fit_on<RPL>(corr, signal, status)) {
                return status;
            }
            while (true) {
                bool mine = false;
#pragma unroll
                for (int t = 0; t < NCPL; ++t) {
                    mine = mine || (col(t) < D.n && act[t] && cand[t] <= real(0.0));
                }
                if (!ballot(mine)) {
                    break;
                }
                if (++n_inner > max_inner) {
#pragma unroll
                    for (int t = 0; t < NCPL; ++t) {
                        x[t] = cand[t];
                    }
                    return status | ST_NOT_CONVERGED;
                }
                real step[NCPL];
                bool blocking[NCPL];
#pragma unroll
                for (int t = 0; t < NCPL; ++t) {
                    blocking[t] = col(t) < D.n && act[t] && cand[t] <= real(0.0);
                    const real distance = x[t] - cand[t];
                    step[t] = (blocking[t] && distance > real(0.0)) ? x[t] / distance : real(0.0);
                }
                real sv = real(0.0);
                int si = -1;
                wave_argmin_cols<NCPL>(step, blocking, sv, si);
                work += 3ull * static_cast<unsigned long long>(D.n);
#pragma unroll
                for (int t = 0; t < NCPL; ++t) {
                    x[t] = x[t] + sv * (cand[t] - x[t]);
                    if (col(t) == si) {
                        x[t] = real(0.0);
                        act[t] = false;
                    }
                    act[t] = act[t] && fabs(x[t]) >= tol;
                }
                if (!this->template fit_on<RPL>(corr, signal, status)) {
                    return status;
                }
            }
#pragma unroll
            for (int t = 0; t < NCPL; ++t) {
                x[t] = cand[t];
            }
            gradient_at_x<RPL>(corr, signal);
        }
        (void)l;
        return status;
    }

    // The support of the solution as a column mask (columns < 64, NCPL == 1): a warm start for
    // the next solve of the same voxel.
    __device__ __forceinline__ unsigned long long support() const {
        return ballot(col(0) < D.n && x[0] != real(0));
    }

    // Gradient: corr - sum over non-zero x_c, in column order, of x_c G[:, c]; or, RESID,
    // C^T (d - C x) of the inactive columns (the entering choice reads no other).
    template <int RPL>
    __device__ __forceinline__ void gradient_at_x(const real (&corr)[NCPL], const RowVec<RPL>& d) {
        if constexpr (RESID) {
            const int l = lane_id();
            // s = sum over every column in order of C(r, c) x_c, then the residual d - s, row by
            // row.
            real fit[RPL];
#pragma unroll
            for (int t = 0; t < RPL; ++t) {
                fit[t] = real(0);
            }
            for (int c = 0; c < D.n; ++c) {
                const real xc = uniform_col(x, c);
#pragma unroll
                for (int t = 0; t < RPL; ++t) {
                    const int r = l + WAVE * t;
                    const real krc = r < D.m ? D.kernel[static_cast<long>(r) * D.n + c] : real(0);
                    fit[t] = fit[t] + krc * xc;
                }
            }
            real residual[RPL];
#pragma unroll
            for (int t = 0; t < RPL; ++t) {
                residual[t] = d.v[t] - fit[t];
            }
#pragma unroll
            for (int t2 = 0; t2 < NCPL; ++t2) {
                g[t2] = real(0);
            }
            for (int c = 0; c < D.n; ++c) {
                const int slot = c / WAVE;
                bool inactive = false;
#pragma unroll
                for (int t2 = 0; t2 < NCPL; ++t2) {
                    inactive = (t2 == slot) ? !act[t2] : inactive;
                }
                if (!ballot(inactive && l == c % WAVE)) {
                    continue;
                }
                const real gc = wave_sum_f<RPL>(0, D.m, [&](int t) {
                    const int r = l + WAVE * t;
                    const real krc = r < D.m ? D.kernel[static_cast<long>(r) * D.n + c] : real(0);
                    return krc * residual[t];
                });
#pragma unroll
                for (int t2 = 0; t2 < NCPL; ++t2) {
                    g[t2] = (t2 == slot && l == c % WAVE) ? gc : g[t2];
                }
            }
            return;
        }
        (void)d;
#pragma unroll
        for (int t = 0; t < NCPL; ++t) {
            g[t] = corr[t];
        }
#pragma unroll
        for (int t2 = 0; t2 < NCPL; ++t2) {
            unsigned long long mask = ballot(col(t2) < D.n && x[t2] != real(0.0));
            while (mask) {
                const int b = __ffsll(static_cast<long long>(mask)) - 1;
                mask &= mask - 1;
                const int c = WAVE * t2 + b;
                work += 2ull * static_cast<unsigned long long>(D.n);
                const real xc = readlane(x[t2], b);
                const real* grow = D.gram + static_cast<long>(c) * D.n;
#pragma unroll
                for (int t = 0; t < NCPL; ++t) {
                    if (col(t) < D.n) {
                        g[t] = fma(-xc, grow[col(t)], g[t]);
                    }
                }
            }
        }
    }
};

} // namespace LSD_NS
