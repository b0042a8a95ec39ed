// Least squares with inequality constraints (LSI) of one voxel per wavefront.
//
// Restates larmor's solve_lsi (cpp/src/sd/lsi.cpp:245-292) and ldp_multipliers (lsi.cpp:191-240)
// on a design prepared once on the host (q_transposed, P R^-1, the unit-row constraints E, E E^T
// and E's row sums, lsi.cpp:19-69):
//  - r = Q^T s; f = -(E r) / max |E r|; with no positive scale the solution is P R^-1 r;
//  - the least-distance multipliers u >= 0 by Lawson-Hanson on H = E E^T + f f^T: the gradient
//    f (1 - f.u) - sum of u_c (E E^T)[:, c] over the non-zero u_c in column order, the entering
//    column the largest gradient (the first on a tie) above 10 eps norm1 max(n + 1, k), a
//    growing Cholesky factor of H on the active columns in the order they joined (a pivot must
//    exceed 1e-12 H_cc), blocking steps that zero the first blocking coefficient and any below
//    the tolerance, Givens downdates, at most 3 k inner steps;
//  - when that fails, the rerun kernel solves the same NNLS on the explicit matrix [E^T; f^T]
//    with pivoted-QR least squares (solve_nnls_with_limit), its rows summed as wave sums;
//  - x = P R^-1 (scale (E^T u) / (1 - f.u) + r), the sums over the non-zero u_c in order.
//
// Layout: constraint (NNLS column) c in lane c % 64, slot c / 64 (NKS slots); parameter i in
// lane i (n <= 64).  The factor lives in a per-wave global slab, column-major with stride
// KCAP + 1; H has rank at most n + 1 <= 65, so a factor of KCAP = 64 columns only overflows where
// a pivot would be refused anyway, and both go to the rerun kernel.
#pragma once

#include "qr_wave.hpp"

namespace lsd {

struct LsiDev {
    int m;
    int n;
    int k;
    const double* q_t;      // (n, m)
    const double* back;     // (n, n): P R^-1
    const double* e;        // (k, n)
    const double* e_t;      // (n, k)
    const double* gram;     // (k, k): E E^T
    const double* abs_sums; // (k)
};

constexpr int LSI_KCAP = 64;
constexpr int LSI_QR_ROWS = 2;    // row slots of [E^T; f^T]: n + 1 <= 128
constexpr int LSI_QR_COLS = 128;  // active columns the QR fallback holds

struct LsiArgs {
    const double* signals;      // (n_vox, m)
    long n_vox;
    LsiDev d;
    double* x_out;              // (n_vox, n)
    unsigned char* converged;   // (n_vox)
    int* status;                // (n_vox): ST_QR_FALLBACK when the QR path finished the voxel
    const int* list;            // voxels to solve (rerun), or null for all
    const int* list_count;
    int* fallback_list;         // voxels handed to the QR path (fast kernel)
    int* fallback_count;
    double* scratch;            // per-wave slabs
    long per_wave;
};

// The value of a column-spread vector at uniform column c (a select over the slots, so the
// array stays in registers).
template <int NKS>
__device__ __forceinline__ double uniform_at(const double (&a)[NKS], int c) {
    double v = 0.0;
#pragma unroll
    for (int t = 0; t < NKS; ++t) {
        v = (c / WAVE) == t ? a[t] : v;
    }
    return readlane(v, c % WAVE);
}

// r = Q^T s (lane i), f's raw values -(E r) per constraint slot and their largest magnitude.
// Returns scale; NaN when any raw value is NaN.
template <int NKS>
__device__ __forceinline__ double lsi_rotate(const LsiDev& d, const double* s, double& rot,
                                             double (&raw)[NKS]) {
    const int l = lane_id();
    rot = 0.0;
    for (int r = 0; r < d.m; ++r) {
        const double sr = s[r];
        if (l < d.n) {
            rot = rot + d.q_t[static_cast<long>(l) * d.m + r] * sr;
        }
    }
    double mx = 0.0;
    bool any_nan = false;
#pragma unroll
    for (int t = 0; t < NKS; ++t) {
        raw[t] = 0.0;
    }
    for (int i = 0; i < d.n; ++i) {
        const double ri = readlane(rot, i);
#pragma unroll
        for (int t = 0; t < NKS; ++t) {
            const int c = l + WAVE * t;
            if (c < d.k) {
                raw[t] = raw[t] + d.e_t[static_cast<long>(i) * d.k + c] * ri;
            }
        }
    }
#pragma unroll
    for (int t = 0; t < NKS; ++t) {
        const int c = l + WAVE * t;
        if (c < d.k) {
            raw[t] = -raw[t];
            const double a = fabs(raw[t]);
            any_nan = any_nan || isnan(a);
            mx = a > mx ? a : mx;
        }
    }
    if (ballot(any_nan)) {
        return __builtin_nan("");
    }
    return wave_max(mx);
}

// x = P R^-1 z for z in lane i (i < n); returns x in lane i.
__device__ __forceinline__ double lsi_back(const LsiDev& d, double z) {
    const int l = lane_id();
    double x = 0.0;
    for (int j = 0; j < d.n; ++j) {
        const double zj = readlane(z, j);
        if (l < d.n) {
            x = x + d.back[static_cast<long>(l) * d.n + j] * zj;
        }
    }
    return x;
}

// f.u over the non-zero u_c in column order (uniform).
template <int NKS>
__device__ __forceinline__ double lsi_dot(const double (&f)[NKS], const double (&u)[NKS]) {
    double dot = 0.0;
#pragma unroll
    for (int t = 0; t < NKS; ++t) {
        unsigned long long bits = ballot(u[t] != 0.0);
        while (bits) {
            const int b = __ffsll(static_cast<long long>(bits)) - 1;
            bits &= bits - 1;
            dot = dot + readlane(f[t], b) * readlane(u[t], b);
        }
    }
    return dot;
}

// The solution from the multipliers: x = P R^-1 (scale (E^T u) / (1 - f.u) + r).
template <int NKS>
__device__ __forceinline__ double lsi_solution(const LsiDev& d, const double (&f)[NKS],
                                               const double (&u)[NKS], double scale, double rot) {
    const int l = lane_id();
    const double denom = 1.0 - lsi_dot<NKS>(f, u);
    double sum = 0.0;
#pragma unroll
    for (int t = 0; t < NKS; ++t) {
        unsigned long long bits = ballot(u[t] != 0.0);
        while (bits) {
            const int b = __ffsll(static_cast<long long>(bits)) - 1;
            bits &= bits - 1;
            const int c = WAVE * t + b;
            const double uc = readlane(u[t], b);
            if (l < d.n) {
                sum = sum + d.e[static_cast<long>(c) * d.n + l] * uc;
            }
        }
    }
    const double z = (scale * sum) / denom + rot;
    return lsi_back(d, z);
}

// Lawson-Hanson on H = E E^T + f f^T with a growing Cholesky factor (ldp_multipliers).
template <int NKS>
struct LdpWave {
    static constexpr int KCAP = LSI_KCAP;
    static constexpr int S = KCAP + 1;

    const LsiDev& d;
    const double (&f)[NKS];
    double* L;
    double u[NKS];
    double g[NKS];
    double cand[NKS];
    bool act[NKS];
    int pos[NKS];
    int fcol; // column at factor position lane_id() (lane < kf)
    int kf;   // factor size (uniform)

    __device__ LdpWave(const LsiDev& d_, const double (&f_)[NKS], double* L_)
        : d(d_), f(f_), L(L_) {}

    __device__ __forceinline__ int col(int t) const { return lane_id() + WAVE * t; }
    __device__ __forceinline__ double& Lat(int i, int j) { return L[j * S + i]; }
    __device__ __forceinline__ double Luni(int i, int j) const {
        return readlane(L[j * S + lane_id()], i);
    }

    // f at column c (per lane).
    __device__ __forceinline__ double f_at(int c) const {
        const int src = c % WAVE;
        const int slot = c / WAVE;
        double out = 0.0;
#pragma unroll
        for (int t = 0; t < NKS; ++t) {
            const double w = gather(f[t], src);
            out = slot == t ? w : out;
        }
        return out;
    }

    // L w = b on positions [0, kk), column-oriented; returns the sum of the w_j^2 in order.
    __device__ __forceinline__ double forward(double& b, int kk) {
        const int l = lane_id();
        double sq = 0.0;
        for (int j = 0; j < kk; ++j) {
            const double bj = readlane(b, j);
            const double ljj = Luni(j, j);
            const double wj = bj != 0.0 ? bj / ljj : bj;
            sq = sq + wj * wj;
            if (l == j) {
                b = wj;
            } else if (l > j && l < kk) {
                b = b - Lat(l, j) * wj;
            }
        }
        return sq;
    }

    // L^T z = w on positions [0, kk), column-oriented.
    __device__ __forceinline__ void backward(double& b, int kk) {
        const int l = lane_id();
        for (int j = kk - 1; j >= 0; --j) {
            const double wj = readlane(b, j);
            const double ljj = Luni(j, j);
            const double zj = wj != 0.0 ? wj / ljj : wj;
            if (l == j) {
                b = zj;
            } else if (l < j) {
                b = b - Lat(j, l) * zj;
            }
        }
    }

    // Appends column c; false when its pivot is refused or the factor is full.
    __device__ __forceinline__ bool append(int c) {
        const int l = lane_id();
        if (kf >= KCAP) {
            return false;
        }
        const double fc = uniform_at<NKS>(f, c);
        const double f_row = f_at(l < kf ? fcol : 0);
        double b = 0.0;
        if (l < kf) {
            b = d.gram[static_cast<long>(fcol) * d.k + c] + f_row * fc;
        }
        const double sq = forward(b, kf);
        const double hcc = d.gram[static_cast<long>(c) * d.k + c] + fc * fc;
        const double pivot = hcc - sq;
        if (!(pivot > MIN_RELATIVE_PIVOT * hcc)) {
            return false;
        }
        if (l < kf) {
            Lat(kf, l) = b;
        }
        if (l == kf) {
            Lat(kf, kf) = sqrt(pivot);
            fcol = c;
        }
#pragma unroll
        for (int t = 0; t < NKS; ++t) {
            pos[t] = col(t) == c ? kf : pos[t];
        }
        kf += 1;
        return true;
    }

    // Removes factor position p (remove_cholesky_column).
    __device__ __forceinline__ void remove(int p) {
        const int l = lane_id();
        for (int c = 0; c < kf; ++c) {
            const bool move = l >= p && l < kf - 1 && c <= l + 1;
            const double v = move ? Lat(l + 1, c) : 0.0;
            if (move) {
                Lat(l, c) = v;
            }
        }
        for (int column = p; column < kf - 1; ++column) {
            const double diagonal = Luni(column, column);
            const double above = Luni(column, column + 1);
            const double length = hypot_fixed(diagonal, above);
            const double cosine = diagonal / length;
            const double sine = above / length;
            if (l >= column && l < kf - 1) {
                const double left = Lat(l, column);
                const double right = Lat(l, column + 1);
                Lat(l, column) = cosine * left + sine * right;
                Lat(l, column + 1) = cosine * right - sine * left;
            }
        }
        const int from = gather(fcol, (l + 1) % WAVE);
        fcol = (l >= p && l < kf - 1) ? from : fcol;
#pragma unroll
        for (int t = 0; t < NKS; ++t) {
            pos[t] = pos[t] == p ? -1 : (pos[t] > p ? pos[t] - 1 : pos[t]);
        }
        kf -= 1;
    }

    // Drops the positions whose column is inactive, last to first (GrowingCholesky::keep).
    __device__ __forceinline__ void keep() {
        const int l = lane_id();
        int actbits = 0;
#pragma unroll
        for (int t = 0; t < NKS; ++t) {
            actbits |= act[t] ? (1 << t) : 0;
        }
        const int c = l < kf ? fcol : 0;
        const int bits = gather(actbits, c % WAVE);
        const bool leaves = l < kf && !((bits >> (c / WAVE)) & 1);
        unsigned long long mask = ballot(leaves);
        while (mask) {
            const int b = 63 - __clzll(static_cast<long long>(mask));
            mask &= ~(1ull << b);
            remove(b);
        }
    }

    // cand = the solution of H_AA z = f_A, scattered to the columns.
    __device__ __forceinline__ void solve() {
        const int l = lane_id();
        const double fv = f_at(l < kf ? fcol : 0);
        double b = l < kf ? fv : 0.0;
        forward(b, kf);
        backward(b, kf);
#pragma unroll
        for (int t = 0; t < NKS; ++t) {
            const int p = pos[t] < 0 ? 0 : pos[t];
            const double z = gather(b, p);
            cand[t] = pos[t] >= 0 ? z : 0.0;
        }
    }

    // Returns true when the multipliers are found (u), false when the QR path must take over.
    __device__ __forceinline__ bool run() {
        const int l = lane_id();
        kf = 0;
        fcol = 0;
        double norm1 = 0.0;
#pragma unroll
        for (int t = 0; t < NKS; ++t) {
            u[t] = 0.0;
            act[t] = false;
            pos[t] = -1;
            cand[t] = 0.0;
            const int c = col(t);
            if (c < d.k) {
                const double v = d.abs_sums[c] + fabs(f[t]);
                norm1 = v > norm1 ? v : norm1;
            }
        }
        norm1 = wave_max(norm1);
        const double tol = 10.0 * DBL_EPS * norm1 *
                           static_cast<double>(d.n + 1 > d.k ? d.n + 1 : d.k);
        const int max_inner = 3 * d.k;
        int n_inner = 0;
        int n_outer = 0;
        while (true) {
            const double dot = lsi_dot<NKS>(f, u);
#pragma unroll
            for (int t = 0; t < NKS; ++t) {
                g[t] = f[t] * (1.0 - dot);
            }
#pragma unroll
            for (int t2 = 0; t2 < NKS; ++t2) {
                unsigned long long bits = ballot(u[t2] != 0.0);
                while (bits) {
                    const int b = __ffsll(static_cast<long long>(bits)) - 1;
                    bits &= bits - 1;
                    const int c = WAVE * t2 + b;
                    const double uc = readlane(u[t2], b);
                    const double* row = d.gram + static_cast<long>(c) * d.k;
#pragma unroll
                    for (int t = 0; t < NKS; ++t) {
                        if (col(t) < d.k) {
                            g[t] = g[t] - uc * row[col(t)];
                        }
                    }
                }
            }
            double bv = 0.0;
            int bi = -1;
#pragma unroll
            for (int t = 0; t < NKS; ++t) {
                if (col(t) < d.k && !act[t] && (bi < 0 || g[t] > bv)) {
                    bv = g[t];
                    bi = col(t);
                }
            }
            wave_argmax_first(bv, bi);
            if (bi < 0 || bv <= tol) {
                return true;
            }
            if (++n_outer > 8 * d.k + 64) {
                return false;
            }
#pragma unroll
            for (int t = 0; t < NKS; ++t) {
                act[t] = act[t] || col(t) == bi;
            }
            if (!append(bi)) {
                return false;
            }
            solve();
            while (true) {
                double sv = 0.0;
                int si = -1;
#pragma unroll
                for (int t = 0; t < NKS; ++t) {
                    if (col(t) < d.k && act[t] && cand[t] <= 0.0) {
                        const double distance = u[t] - cand[t];
                        const double s = distance > 0.0 ? u[t] / distance : 0.0;
                        if (si < 0 || s < sv) {
                            sv = s;
                            si = col(t);
                        }
                    }
                }
                wave_argmin_first(sv, si);
                if (si < 0) {
                    break;
                }
                if (++n_inner > max_inner) {
                    return false;
                }
#pragma unroll
                for (int t = 0; t < NKS; ++t) {
                    u[t] = u[t] + sv * (cand[t] - u[t]);
                    if (col(t) == si || fabs(u[t]) < tol) {
                        act[t] = false;
                        u[t] = 0.0;
                    }
                }
                keep();
                solve();
            }
#pragma unroll
            for (int t = 0; t < NKS; ++t) {
                u[t] = cand[t];
            }
        }
        (void)l;
    }
};

// Column j of [E^T; f^T] at the lane's row slot t: rows 0..n-1 are E[j, :], row n is f_j.  fj
// comes from the caller, read with readlane where every lane is active (inside a per-lane
// branch a masked source lane would not have its select done).
struct LdpColumns {
    const LsiDev& d;
    __device__ __forceinline__ double get(int j, int t, double fj) const {
        const int r = lane_id() + WAVE * t;
        if (r < d.n) {
            return d.e[static_cast<long>(j) * d.n + r];
        }
        return r == d.n ? fj : 0.0;
    }
};

// solve_nnls_with_limit on [E^T; f^T] (rows n + 1, columns k) with target e_n: the QR fallback
// of ldp_multipliers.  work: LSI_QR_ROWS * 64 * LSI_QR_COLS doubles (global).  Returns whether
// it converged; u receives the multipliers.
template <int NKS>
__device__ __forceinline__ bool ldp_by_qr(const LsiDev& d, const double (&f)[NKS],
                                          double (&u)[NKS], double* work) {
    constexpr int RP = LSI_QR_ROWS;
    const int l = lane_id();
    const int rows = d.n + 1;
    const int cols = d.k;
    const LdpColumns C{d};
    RowVec<RP> target;
#pragma unroll
    for (int t = 0; t < RP; ++t) {
        target.v[t] = (l + WAVE * t) == d.n ? 1.0 : 0.0;
    }
    // Default tolerance: 10 eps max_j (wave sum of |C(:, j)|) max(rows, cols).
    double norm1 = 0.0;
    for (int j = 0; j < cols; ++j) {
        const double fj = uniform_at<NKS>(f, j);
        const double s = wave_sum_f<RP>(0, rows, [&](int t) { return fabs(C.get(j, t, fj)); });
        if (j == 0 || s > norm1) {
            norm1 = s;
        }
    }
    const double tol = 10.0 * DBL_EPS * norm1 * static_cast<double>(rows > cols ? rows : cols);
    const int max_inner = 3 * cols;
    bool act[NKS];
    double g[NKS];
    double cand[NKS];
#pragma unroll
    for (int t = 0; t < NKS; ++t) {
        u[t] = 0.0;
        act[t] = false;
    }
    auto gradient = [&]() {
        // r = d - C u over the non-zero u_j in order, then g_j = wave sum of C(:, j) r.
        double r[RP];
#pragma unroll
        for (int t = 0; t < RP; ++t) {
            r[t] = 0.0;
        }
#pragma unroll
        for (int t2 = 0; t2 < NKS; ++t2) {
            unsigned long long bits = ballot(u[t2] != 0.0);
            while (bits) {
                const int b = __ffsll(static_cast<long long>(bits)) - 1;
                bits &= bits - 1;
                const int j = WAVE * t2 + b;
                const double uj = readlane(u[t2], b);
                const double fj = readlane(f[t2], b);
#pragma unroll
                for (int t = 0; t < RP; ++t) {
                    r[t] = r[t] + C.get(j, t, fj) * uj;
                }
            }
        }
#pragma unroll
        for (int t = 0; t < RP; ++t) {
            r[t] = target.v[t] - r[t];
        }
#pragma unroll
        for (int t = 0; t < NKS; ++t) {
            g[t] = 0.0;
        }
        for (int j = 0; j < cols; ++j) {
            const double fj = uniform_at<NKS>(f, j);
            const double s =
                wave_sum_f<RP>(0, rows, [&](int t) { return C.get(j, t, fj) * r[t]; });
#pragma unroll
            for (int t = 0; t < NKS; ++t) {
                g[t] = (l + WAVE * t) == j ? s : g[t];
            }
        }
    };
    // Least squares on the active columns in increasing order; false on overflow.
    auto fit_on = [&]() -> bool {
        MemCols<RP> A{work};
        int ka = 0;
#pragma unroll
        for (int t2 = 0; t2 < NKS; ++t2) {
            unsigned long long bits = ballot((l + WAVE * t2) < cols && act[t2]);
            while (bits) {
                const int b = __ffsll(static_cast<long long>(bits)) - 1;
                bits &= bits - 1;
                if (ka >= LSI_QR_COLS) {
                    return false;
                }
                const int j = WAVE * t2 + b;
                const double fj = readlane(f[t2], b);
#pragma unroll
                for (int t = 0; t < RP; ++t) {
                    A.set(ka, t, C.get(j, t, fj));
                }
                ++ka;
            }
        }
#pragma unroll
        for (int t = 0; t < NKS; ++t) {
            cand[t] = 0.0;
        }
        if (ka == 0) {
            return true;
        }
        RowVec<RP> dd = target;
        LaneVec<LSI_QR_COLS / WAVE> coef;
        least_squares_wave<LSI_QR_COLS / WAVE>(A, rows, ka, dd, coef);
        int p = 0;
#pragma unroll
        for (int t2 = 0; t2 < NKS; ++t2) {
            unsigned long long bits = ballot((l + WAVE * t2) < cols && act[t2]);
            while (bits) {
                const int b = __ffsll(static_cast<long long>(bits)) - 1;
                bits &= bits - 1;
                const double v = coef.get(p);
                if (l == b) {
                    cand[t2] = v;
                }
                ++p;
            }
        }
        return true;
    };

    gradient();
    int n_inner = 0;
    int n_outer = 0;
    while (true) {
        if (++n_outer > 8 * cols + 64) {
            return false;
        }
        double bv = 0.0;
        int bi = -1;
#pragma unroll
        for (int t = 0; t < NKS; ++t) {
            const int j = l + WAVE * t;
            if (j < cols && !act[t] && (bi < 0 || g[t] > bv)) {
                bv = g[t];
                bi = j;
            }
        }
        wave_argmax_first(bv, bi);
        if (bi < 0 || bv <= tol) {
            return true;
        }
#pragma unroll
        for (int t = 0; t < NKS; ++t) {
            act[t] = act[t] || (l + WAVE * t) == bi;
        }
        if (!fit_on()) {
            return false;
        }
        while (true) {
            double sv = 0.0;
            int si = -1;
#pragma unroll
            for (int t = 0; t < NKS; ++t) {
                const int j = l + WAVE * t;
                if (j < cols && act[t] && cand[t] <= 0.0) {
                    const double distance = u[t] - cand[t];
                    const double s = distance > 0.0 ? u[t] / distance : 0.0;
                    if (si < 0 || s < sv) {
                        sv = s;
                        si = j;
                    }
                }
            }
            wave_argmin_first(sv, si);
            if (si < 0) {
                break;
            }
            if (++n_inner > max_inner) {
#pragma unroll
                for (int t = 0; t < NKS; ++t) {
                    u[t] = cand[t];
                }
                return false;
            }
#pragma unroll
            for (int t = 0; t < NKS; ++t) {
                const int j = l + WAVE * t;
                u[t] = u[t] + sv * (cand[t] - u[t]);
                if (j == si) {
                    u[t] = 0.0;
                    act[t] = false;
                }
                act[t] = act[t] && fabs(u[t]) >= tol;
            }
            if (!fit_on()) {
                return false;
            }
        }
#pragma unroll
        for (int t = 0; t < NKS; ++t) {
            u[t] = cand[t];
        }
        gradient();
    }
}

// The fast path over all voxels: the Cholesky LDP; failures go to fallback_list.
template <int NKS>
__global__ void __launch_bounds__(WAVE * 4) lsi_kernel(LsiArgs A) {
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / WAVE);
    const long gw = static_cast<long>(blockIdx.x) * (blockDim.x / WAVE) + w;
    const long n_waves = static_cast<long>(gridDim.x) * (blockDim.x / WAVE);
    double* slab = A.scratch + gw * A.per_wave;
    const int l = lane_id();
    const LsiDev& d = A.d;
    for (long v = gw; v < A.n_vox; v += n_waves) {
        double rot;
        double raw[NKS];
        const double scale = lsi_rotate<NKS>(d, A.signals + v * d.m, rot, raw);
        if (!(scale > 0.0)) {
            const double x = lsi_back(d, rot);
            if (l < d.n) A.x_out[v * d.n + l] = x;
            if (l == 0) {
                A.converged[v] = 1;
                A.status[v] = 0;
            }
            continue;
        }
        double f[NKS];
#pragma unroll
        for (int t = 0; t < NKS; ++t) {
            f[t] = raw[t] / scale;
        }
        LdpWave<NKS> ldp(d, f, slab);
        if (!ldp.run()) {
            if (l == 0) {
                A.fallback_list[atomicAdd(A.fallback_count, 1)] = static_cast<int>(v);
            }
            continue;
        }
        const double x = lsi_solution<NKS>(d, f, ldp.u, scale, rot);
        if (l < d.n) A.x_out[v * d.n + l] = x;
        if (l == 0) {
            A.converged[v] = 1;
            A.status[v] = 0;
        }
    }
}

// The QR path over the listed voxels (those the fast path handed over).
template <int NKS>
__global__ void __launch_bounds__(WAVE * 4) lsi_qr_kernel(LsiArgs A) {
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / WAVE);
    const long gw = static_cast<long>(blockIdx.x) * (blockDim.x / WAVE) + w;
    const long n_waves = static_cast<long>(gridDim.x) * (blockDim.x / WAVE);
    double* slab = A.scratch + gw * A.per_wave;
    const int l = lane_id();
    const LsiDev& d = A.d;
    const long count = A.list ? static_cast<long>(*A.list_count) : A.n_vox;
    for (long it = gw; it < count; it += n_waves) {
        const long v = A.list ? A.list[it] : it;
        double rot;
        double raw[NKS];
        const double scale = lsi_rotate<NKS>(d, A.signals + v * d.m, rot, raw);
        if (!(scale > 0.0)) {
            const double x = lsi_back(d, rot);
            if (l < d.n) A.x_out[v * d.n + l] = x;
            if (l == 0) {
                A.converged[v] = 1;
                A.status[v] = ST_QR_FALLBACK;
            }
            continue;
        }
        double f[NKS];
#pragma unroll
        for (int t = 0; t < NKS; ++t) {
            f[t] = raw[t] / scale;
        }
        double u[NKS];
        const bool conv = ldp_by_qr<NKS>(d, f, u, slab);
        const double x = lsi_solution<NKS>(d, f, u, scale, rot);
        if (l < d.n) A.x_out[v * d.n + l] = x;
        if (l == 0) {
            A.converged[v] = conv ? 1 : 0;
            A.status[v] = ST_QR_FALLBACK;
        }
    }
}

} // namespace lsd
