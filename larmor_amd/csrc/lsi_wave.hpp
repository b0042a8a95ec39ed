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
//    with pivoted-QR least squares (solve_nnls_with_limit): the QR's row sums are wave sums, the
//    gradient and the tolerance's column sums run over the rows in order, one column per lane;
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
constexpr int LSI_QR_ROWS = 1;    // row slots of [E^T; f^T]: n + 1 <= 64
constexpr int LSI_GRADIENT_BATCH = 2; // rows of E E^T loaded together by the gradient
constexpr int LSI_DIAG_DOUBLES = 2 * LSI_KCAP; // per-wave LDS table of the factor diagonal

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

// Lawson-Hanson on H = E E^T + f f^T with a growing Cholesky factor (ldp_multipliers).  As in
// the ridge solver: the factor's diagonal and its division reciprocals sit in a per-wave LDS table
// (the split division of wave.hpp where every diagonal and the dividend are in range), and
// y = L^-1 f_A is kept across appends (the new row gives y_k by one sequential dot, the order of
// the column-oriented solve) and recomputed from a removed position on.
template <int NKS>
struct LdpWave {
    static constexpr int KCAP = LSI_KCAP;
    static constexpr int S = KCAP + 1;

    const LsiDev& d;
    const double (&f)[NKS];
    double* L;    // factor, global slab, column-major with stride S
    double* diag; // LDS: diag[2 p] = L_pp, diag[2 p + 1] = its division reciprocal
    bool diag_ok; // every diagonal in the divisor range (uniform)
    double u[NKS];
    double g[NKS];
    double cand[NKS];
    bool act[NKS];
    int pos[NKS];
    int fcol;    // column at factor position lane_id() (lane < kf)
    int kf;      // factor size (uniform)
    double y;    // y = L^-1 f_A at the lane's position, valid on positions [0, y_valid)
    int y_valid;

    __device__ LdpWave(const LsiDev& d_, const double (&f_)[NKS], double* L_, double* diag_)
        : d(d_), f(f_), L(L_), diag(diag_) {}

    __device__ __forceinline__ int col(int t) const { return lane_id() + WAVE * t; }
    __device__ __forceinline__ double& Lat(int i, int j) { return L[j * S + i]; }
    __device__ __forceinline__ double Luni(int i, int j) const {
        return readlane(L[j * S + lane_id()], i);
    }
    __device__ __forceinline__ double2 diag_at(int j) const {
        return *reinterpret_cast<const double2*>(diag + 2 * j);
    }
    // a != 0 ? a / L_jj : a from dd = (L_jj, its reciprocal).
    template <bool FAST>
    __device__ __forceinline__ double divide_by_diag(double a, double2 dd) const {
        if (FAST && dividend_in_range(a)) {
            return divide_in_range(a, dd.x, dd.y);
        }
        return a != 0.0 ? a / dd.x : a;
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

    // L w = b on positions [p0, kk) when positions [0, p0) of b already hold their final values
    // (p0 = 0: the whole solve), column-oriented; returns the sum of the w_j^2 in order (for
    // p0 = 0).  The factor lives in global memory: column j + 1 is loaded during step j, so the
    // load latency is off the chain of dependent steps.
    template <bool FAST>
    __device__ __forceinline__ double forward_impl(double& b, int p0, int kk) {
        const int l = lane_id();
        double sq = 0.0;
        double col_next = L[l];
        for (int j = 0; j < kk; ++j) {
            const double colj = col_next;
            const int jn = j + 1 < kk ? j + 1 : j;
            col_next = L[jn * S + l];
            const double2 dd = diag_at(j);
            const double bj = readlane(b, j);
            const double wj = j >= p0 ? divide_by_diag<FAST>(bj, dd) : bj;
            sq = sq + wj * wj;
            const double updated = b - colj * wj;
            b = l == j ? wj : ((l > j && l < kk && l >= p0) ? updated : b);
        }
        return sq;
    }

    __device__ __forceinline__ double forward(double& b, int kk_) {
        const int kk = __builtin_amdgcn_readfirstlane(kk_);
        return diag_ok ? forward_impl<true>(b, 0, kk) : forward_impl<false>(b, 0, kk);
    }

    __device__ __forceinline__ void forward_from(double& b, int p0_, int kk_) {
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

    // L^T z = w on positions [0, kk), column-oriented: step j reads row j (L_{j,l} for the
    // lanes l < j, at stride S), loaded during step j + 1.
    template <bool FAST>
    __device__ __forceinline__ void backward_impl(double& b, int kk) {
        const int l = lane_id();
        const int top = kk > 0 ? kk - 1 : 0;
        const int lr = l < KCAP ? l : 0;
        double row_next = L[lr * S + top];
        for (int j = kk - 1; j >= 0; --j) {
            const double rowj = row_next;
            const int jn = j > 0 ? j - 1 : 0;
            row_next = L[lr * S + jn];
            const double2 dd = diag_at(j);
            const double wj = readlane(b, j);
            const double zj = divide_by_diag<FAST>(wj, dd);
            const double updated = b - rowj * zj;
            b = l == j ? zj : (l < j ? updated : b);
        }
    }

    __device__ __forceinline__ void backward(double& b, int kk_) {
        const int kk = __builtin_amdgcn_readfirstlane(kk_);
        if (diag_ok) {
            backward_impl<true>(b, kk);
        } else {
            backward_impl<false>(b, kk);
        }
    }

    // Appends column c; false when its pivot is refused or the factor is full.  With y valid on
    // every position, also extends y to the new one.
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
        const double lkk = sqrt(pivot);
        const double rkk = division_reciprocal(lkk);
        if (l < kf) {
            Lat(kf, l) = b;
        }
        if (l == kf) {
            Lat(kf, kf) = lkk;
            fcol = c;
            diag[2 * kf] = lkk;
            diag[2 * kf + 1] = rkk;
        }
        const bool lkk_ok = divisor_in_range(lkk);
        if (y_valid == kf) {
            // y_k = (f_c - L_k0 y_0 - L_k1 y_1 - ...) / L_kk, subtracted in order.
            const double prod = b * y;
            double v = fc;
            for (int j = 0; j < kf; ++j) {
                v = v - readlane(prod, j);
            }
            const double yk = divide_nonzero_by(v, lkk, rkk, lkk_ok);
            y = l == kf ? yk : y;
            y_valid = kf + 1;
        }
#pragma unroll
        for (int t = 0; t < NKS; ++t) {
            pos[t] = col(t) == c ? kf : pos[t];
        }
        diag_ok = diag_ok && lkk_ok;
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
            const double rl = division_reciprocal(length);
            const double cosine = divide_by(diagonal, length, rl);
            const double sine = divide_by(above, length, rl);
            if (l >= column && l < kf - 1) {
                const double left = Lat(l, column);
                const double right = Lat(l, column + 1);
                Lat(l, column) = cosine * left + sine * right;
                Lat(l, column + 1) = cosine * right - sine * left;
            }
        }
        // The rotated diagonal and its reciprocals (each lane its own position).
        bool out_of_range = false;
        if (l < kf - 1) {
            double dl = diag[2 * l];
            if (l >= p) {
                dl = Lat(l, l);
                diag[2 * l] = dl;
                diag[2 * l + 1] = division_reciprocal(dl);
            }
            out_of_range = !divisor_in_range(dl);
        }
        diag_ok = ballot(out_of_range) == 0ull;
        const int from = gather(fcol, (l + 1) % WAVE);
        fcol = (l >= p && l < kf - 1) ? from : fcol;
#pragma unroll
        for (int t = 0; t < NKS; ++t) {
            pos[t] = pos[t] == p ? -1 : (pos[t] > p ? pos[t] - 1 : pos[t]);
        }
        kf -= 1;
        y_valid = y_valid < p ? y_valid : p;
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
        double b = l < kf ? (l < y_valid ? y : fv) : 0.0;
        forward_from(b, y_valid, kf);
        y = b;
        y_valid = kf;
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
        y = 0.0;
        y_valid = 0;
        diag_ok = true;
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
            // The rows of E E^T for the non-zero u_c are loaded LSI_GRADIENT_BATCH at a time,
            // so their loads overlap, and subtracted in column order.
#pragma unroll
            for (int t2 = 0; t2 < NKS; ++t2) {
                unsigned long long bits = ballot(u[t2] != 0.0);
                while (bits) {
                    double ub[LSI_GRADIENT_BATCH];
                    double rb[LSI_GRADIENT_BATCH][NKS];
                    bool valid[LSI_GRADIENT_BATCH];
#pragma unroll
                    for (int q = 0; q < LSI_GRADIENT_BATCH; ++q) {
                        valid[q] = bits != 0ull;
                        const int b = valid[q] ? __ffsll(static_cast<long long>(bits)) - 1 : 0;
                        bits &= bits - 1;
                        ub[q] = readlane(u[t2], b);
                        const double* row = d.gram + static_cast<long>(WAVE * t2 + b) * d.k;
#pragma unroll
                        for (int t = 0; t < NKS; ++t) {
                            rb[q][t] = (valid[q] && col(t) < d.k) ? row[col(t)] : 0.0;
                        }
                    }
#pragma unroll
                    for (int q = 0; q < LSI_GRADIENT_BATCH; ++q) {
                        if (valid[q]) {
#pragma unroll
                            for (int t = 0; t < NKS; ++t) {
                                if (col(t) < d.k) {
                                    g[t] = g[t] - ub[q] * rb[q][t];
                                }
                            }
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

// solve_nnls_with_limit on [E^T; f^T] (rows n + 1 <= 64 RP, columns k) with target e_n: the QR
// fallback of ldp_multipliers.  The active columns (at most 64 RP: as many as the rows, which in
// exact arithmetic bound an independent active set) are copied to `work`, RP * 64 * 64 RP
// doubles in LDS or global memory.  Returns whether it converged; u receives the multipliers.
template <int NKS, int RP>
__device__ __forceinline__ bool ldp_by_qr(const LsiDev& d, const double (&f)[NKS],
                                          double (&u)[NKS], double* work) {
    constexpr int QCOLS = RP * WAVE;
    const int l = lane_id();
    const int rows = d.n + 1;
    const int cols = d.k;
    const LdpColumns C{d};
    RowVec<RP> target;
#pragma unroll
    for (int t = 0; t < RP; ++t) {
        target.v[t] = (l + WAVE * t) == d.n ? 1.0 : 0.0;
    }
    // sum over rows r in order of term(C(r, j), res_r) for the lane's columns j: row r < n of
    // column j is E[j, r] = e_t[r, j] (one coalesced load per row and slot), row n is f_j.
    auto column_sums = [&](auto term, const RowVec<RP>& res, double (&out)[NKS]) {
#pragma unroll
        for (int t = 0; t < NKS; ++t) {
            out[t] = 0.0;
        }
        for (int r = 0; r < rows; ++r) {
            const double rr = res.get(r);
#pragma unroll
            for (int t = 0; t < NKS; ++t) {
                const int c = (l + WAVE * t) < cols ? l + WAVE * t : 0;
                const double v = r < d.n ? d.e_t[static_cast<long>(r) * d.k + c] : f[t];
                out[t] = out[t] + term(v, rr);
            }
        }
    };
    // Default tolerance: 10 eps max_j (sum over rows of |C(r, j)|) max(rows, cols).
    double norm1 = 0.0;
    {
        double abs_sums[NKS];
        RowVec<RP> none;
        none.fill(0.0);
        column_sums([](double v, double) { return fabs(v); }, none, abs_sums);
        double lane_max = 0.0;
#pragma unroll
        for (int t = 0; t < NKS; ++t) {
            lane_max = ((l + WAVE * t) < cols && abs_sums[t] > lane_max) ? abs_sums[t] : lane_max;
        }
        norm1 = wave_max(lane_max);
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
        // r = d - C u over the non-zero u_j in order, then g_j = sum over rows in order of
        // C(r, j) r_r.
        RowVec<RP> res;
#pragma unroll
        for (int t = 0; t < RP; ++t) {
            res.v[t] = 0.0;
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
                    res.v[t] = res.v[t] + C.get(j, t, fj) * uj;
                }
            }
        }
#pragma unroll
        for (int t = 0; t < RP; ++t) {
            res.v[t] = target.v[t] - res.v[t];
        }
        column_sums([](double v, double rr) { return v * rr; }, res, g);
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
                if (ka >= QCOLS) {
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
        LaneVec<RP> coef;
        least_squares_wave<RP>(A, rows, ka, dd, coef);
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
    extern __shared__ double lsi_diag[];
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / WAVE);
    const long gw = static_cast<long>(blockIdx.x) * (blockDim.x / WAVE) + w;
    const long n_waves = static_cast<long>(gridDim.x) * (blockDim.x / WAVE);
    double* slab = A.scratch + gw * A.per_wave;
    double* diag = lsi_diag + w * LSI_DIAG_DOUBLES;
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
        LdpWave<NKS> ldp(d, f, slab, diag);
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

// Doubles of the QR path's per-wave column slab: 64 rows (n + 1 <= 64) by 64 columns.
constexpr long LSI_QR_SLAB = static_cast<long>(LSI_QR_ROWS) * WAVE * LSI_QR_ROWS * WAVE;

// The QR path over the listed voxels (those the fast path handed over), one wave per block with
// its column slab (32 KB) in the block's dynamic LDS.
template <int NKS>
__global__ void __launch_bounds__(WAVE) lsi_qr_kernel(LsiArgs A) {
    extern __shared__ double smem[];
    const long gw = blockIdx.x;
    const long n_waves = gridDim.x;
    double* slab = smem;
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
        const bool conv = ldp_by_qr<NKS, LSI_QR_ROWS>(d, f, u, slab);
        const double x = lsi_solution<NKS>(d, f, u, scale, rot);
        if (l < d.n) A.x_out[v * d.n + l] = x;
        if (l == 0) {
            A.converged[v] = conv ? 1 : 0;
            A.status[v] = ST_QR_FALLBACK;
        }
    }
}

} // namespace lsd
