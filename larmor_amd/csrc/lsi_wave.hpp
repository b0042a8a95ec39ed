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
// Layout: constraint (NNLS column) c in lane c % 64, slot c / 64 (NKS slots); parameter i and
// factor position p in lane i % 64 (p % 64), slot i / 64 (PS slots: n <= 63, 127 or 255 for PS =
// 1, 2, 4).  The factor lives in a per-wave global slab, column-major with stride KCAP + 1; H has
// rank at most n + 1 <= 64 PS, so a factor of KCAP = 64 PS columns only overflows where a pivot
// would be refused anyway, and both go to the rerun kernel.  Every lane reads and writes only the
// factor rows of its own positions, except the row shift of a removal, whose slots run one after
// another with the wave's memory operations completed in between.
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

constexpr int LSI_GRADIENT_BATCH = 2; // rows of E E^T loaded together by the gradient
constexpr int LSI_MAX_PARAMS = 4 * WAVE - 1;

// Parameter / position slots for n parameters: 1, 2 or 4 (0: too many).
inline int lsi_position_slots(long n) { return n < WAVE ? 1 : (n < 2 * WAVE ? 2 : (n < 4 * WAVE ? 4 : 0)); }
// The factor's columns, per-wave global slab doubles (stride KCAP + 1) and LDS diagonal table.
constexpr int lsi_kcap(int ps) { return WAVE * ps; }
constexpr long lsi_factor_slab(int ps) { return static_cast<long>(lsi_kcap(ps) + 1) * lsi_kcap(ps); }
constexpr int lsi_diag_doubles(int ps) { return 2 * lsi_kcap(ps); }
// The QR path's column slab: 64 PS rows (n + 1 <= 64 PS) by 64 PS columns.
constexpr long lsi_qr_slab(int ps) { return static_cast<long>(WAVE * ps) * WAVE * ps; }

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

// The value at uniform position p of a position-spread vector (p in lane p % 64, slot p / 64).
template <int PS, class T>
__device__ __forceinline__ T at_position(const T (&v)[PS], int p) {
    T x = v[0];
#pragma unroll
    for (int s = 1; s < PS; ++s) {
        x = (p / WAVE) == s ? v[s] : x;
    }
    return readlane(x, p % WAVE);
}

// The value at per-lane position q of a position-spread vector.
template <int PS>
__device__ __forceinline__ double gather_position(const double (&v)[PS], int q) {
    double x = 0.0;
#pragma unroll
    for (int s = 0; s < PS; ++s) {
        const double z = gather(v[s], q % WAVE);
        x = (q / WAVE) == s ? z : x;
    }
    return x;
}

// r = Q^T s (parameter-spread), f's raw values -(E r) per constraint slot and their largest
// magnitude.  Returns scale; NaN when any raw value is NaN.
template <int NKS, int PS>
__device__ __forceinline__ double lsi_rotate(const LsiDev& d, const double* s, double (&rot)[PS],
                                             double (&raw)[NKS]) {
    const int l = lane_id();
#pragma unroll
    for (int q = 0; q < PS; ++q) {
        rot[q] = 0.0;
    }
    for (int r = 0; r < d.m; ++r) {
        const double sr = s[r];
#pragma unroll
        for (int q = 0; q < PS; ++q) {
            const int i = l + WAVE * q;
            if (i < d.n) {
                rot[q] = rot[q] + d.q_t[static_cast<long>(i) * d.m + r] * sr;
            }
        }
    }
    double mx = 0.0;
    bool any_nan = false;
#pragma unroll
    for (int t = 0; t < NKS; ++t) {
        raw[t] = 0.0;
    }
    for (int i = 0; i < d.n; ++i) {
        const double ri = at_position<PS>(rot, i);
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

// x = P R^-1 z, both parameter-spread.
template <int PS>
__device__ __forceinline__ void lsi_back(const LsiDev& d, const double (&z)[PS], double (&x)[PS]) {
    const int l = lane_id();
#pragma unroll
    for (int q = 0; q < PS; ++q) {
        x[q] = 0.0;
    }
    for (int j = 0; j < d.n; ++j) {
        const double zj = at_position<PS>(z, j);
#pragma unroll
        for (int q = 0; q < PS; ++q) {
            const int i = l + WAVE * q;
            if (i < d.n) {
                x[q] = x[q] + d.back[static_cast<long>(i) * d.n + j] * zj;
            }
        }
    }
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
template <int NKS, int PS>
__device__ __forceinline__ void lsi_solution(const LsiDev& d, const double (&f)[NKS],
                                             const double (&u)[NKS], double scale,
                                             const double (&rot)[PS], double (&x)[PS]) {
    const int l = lane_id();
    const double denom = 1.0 - lsi_dot<NKS>(f, u);
    double sum[PS];
#pragma unroll
    for (int q = 0; q < PS; ++q) {
        sum[q] = 0.0;
    }
#pragma unroll
    for (int t = 0; t < NKS; ++t) {
        unsigned long long bits = ballot(u[t] != 0.0);
        while (bits) {
            const int b = __ffsll(static_cast<long long>(bits)) - 1;
            bits &= bits - 1;
            const int c = WAVE * t + b;
            const double uc = readlane(u[t], b);
#pragma unroll
            for (int q = 0; q < PS; ++q) {
                const int i = l + WAVE * q;
                if (i < d.n) {
                    sum[q] = sum[q] + d.e[static_cast<long>(c) * d.n + i] * uc;
                }
            }
        }
    }
    double z[PS];
#pragma unroll
    for (int q = 0; q < PS; ++q) {
        z[q] = (scale * sum[q]) / denom + rot[q];
    }
    lsi_back<PS>(d, z, x);
}

// Writes the parameter-spread solution of voxel v.
template <int PS>
__device__ __forceinline__ void lsi_store(const LsiDev& d, double* x_out, long v, const double (&x)[PS]) {
    const int l = lane_id();
#pragma unroll
    for (int q = 0; q < PS; ++q) {
        const int i = l + WAVE * q;
        if (i < d.n) {
            x_out[v * d.n + i] = x[q];
        }
    }
}

// Lawson-Hanson on H = E E^T + f f^T with a growing Cholesky factor (ldp_multipliers).  As in
// the ridge solver: the factor's diagonal and its division reciprocals sit in a per-wave LDS table
// (the split division of wave.hpp where every diagonal and the dividend are in range), and
// y = L^-1 f_A is kept across appends (the new row gives y_k by one sequential dot, the order of
// the column-oriented solve) and recomputed from a removed position on.  Factor positions are
// spread over PS lane slots (position p in lane p % 64, slot p / 64).
template <int NKS, int PS>
struct LdpWave {
    static constexpr int KCAP = lsi_kcap(PS);
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
    int fcol[PS]; // column at the lane's factor positions (position < kf)
    int kf;       // factor size (uniform)
    double y[PS]; // y = L^-1 f_A at the lane's positions, valid on positions [0, y_valid)
    int y_valid;

    __device__ LdpWave(const LsiDev& d_, const double (&f_)[NKS], double* L_, double* diag_)
        : d(d_), f(f_), L(L_), diag(diag_) {}

    __device__ __forceinline__ int col(int t) const { return lane_id() + WAVE * t; }
    __device__ __forceinline__ int position(int q) const { return lane_id() + WAVE * q; }
    __device__ __forceinline__ double& Lat(int i, int j) { return L[j * S + i]; }
    // L_ij at uniform (i, j): lane i % 64 reads its own row i.
    __device__ __forceinline__ double Luni(int i, int j) const {
        return readlane(L[j * S + WAVE * (i / WAVE) + lane_id()], i % WAVE);
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
    // load latency is off the chain of dependent steps.  Slots past kk load nothing.
    template <bool FAST>
    __device__ __forceinline__ double forward_impl(double (&b)[PS], int p0, int kk) {
        const int l = lane_id();
        double sq = 0.0;
        double col_next[PS];
#pragma unroll
        for (int q = 0; q < PS; ++q) {
            col_next[q] = (q == 0 || WAVE * q < kk) ? L[WAVE * q + l] : 0.0;
        }
        for (int j = 0; j < kk; ++j) {
            double colj[PS];
#pragma unroll
            for (int q = 0; q < PS; ++q) {
                colj[q] = col_next[q];
            }
            const int jn = j + 1 < kk ? j + 1 : j;
#pragma unroll
            for (int q = 0; q < PS; ++q) {
                if (q == 0 || WAVE * q < kk) {
                    col_next[q] = L[jn * S + WAVE * q + l];
                }
            }
            const double2 dd = diag_at(j);
            const double bj = at_position<PS>(b, j);
            const double wj = j >= p0 ? divide_by_diag<FAST>(bj, dd) : bj;
            sq = sq + wj * wj;
#pragma unroll
            for (int q = 0; q < PS; ++q) {
                const int p = position(q);
                const double updated = b[q] - colj[q] * wj;
                b[q] = p == j ? wj : ((p > j && p < kk && p >= p0) ? updated : b[q]);
            }
        }
        return sq;
    }

    __device__ __forceinline__ double forward(double (&b)[PS], int kk_) {
        const int kk = __builtin_amdgcn_readfirstlane(kk_);
        return diag_ok ? forward_impl<true>(b, 0, kk) : forward_impl<false>(b, 0, kk);
    }

    __device__ __forceinline__ void forward_from(double (&b)[PS], int p0_, int kk_) {
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

    // L^T z = w on positions [0, kk), column-oriented: step j reads row j (L_{j,p} for the
    // positions p < j, at stride S), loaded during step j + 1.
    template <bool FAST>
    __device__ __forceinline__ void backward_impl(double (&b)[PS], int kk) {
        const int top = kk > 0 ? kk - 1 : 0;
        double row_next[PS];
#pragma unroll
        for (int q = 0; q < PS; ++q) {
            row_next[q] = (q == 0 || WAVE * q < kk) ? L[position(q) * S + top] : 0.0;
        }
        for (int j = kk - 1; j >= 0; --j) {
            double rowj[PS];
#pragma unroll
            for (int q = 0; q < PS; ++q) {
                rowj[q] = row_next[q];
            }
            const int jn = j > 0 ? j - 1 : 0;
#pragma unroll
            for (int q = 0; q < PS; ++q) {
                if (q == 0 || WAVE * q < kk) {
                    row_next[q] = L[position(q) * S + jn];
                }
            }
            const double2 dd = diag_at(j);
            const double wj = at_position<PS>(b, j);
            const double zj = divide_by_diag<FAST>(wj, dd);
#pragma unroll
            for (int q = 0; q < PS; ++q) {
                const int p = position(q);
                const double updated = b[q] - rowj[q] * zj;
                b[q] = p == j ? zj : (p < j ? updated : b[q]);
            }
        }
    }

    __device__ __forceinline__ void backward(double (&b)[PS], int kk_) {
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
        if (kf >= KCAP) {
            return false;
        }
        const double fc = uniform_at<NKS>(f, c);
        double b[PS];
#pragma unroll
        for (int q = 0; q < PS; ++q) {
            const int p = position(q);
            const double f_row = f_at(p < kf ? fcol[q] : 0);
            b[q] = 0.0;
            if (p < kf) {
                b[q] = d.gram[static_cast<long>(fcol[q]) * d.k + c] + f_row * fc;
            }
        }
        const double sq = forward(b, kf);
        const double hcc = d.gram[static_cast<long>(c) * d.k + c] + fc * fc;
        const double pivot = hcc - sq;
        if (!(pivot > MIN_RELATIVE_PIVOT * hcc)) {
            return false;
        }
        const double lkk = sqrt(pivot);
        const double rkk = division_reciprocal(lkk);
#pragma unroll
        for (int q = 0; q < PS; ++q) {
            const int p = position(q);
            if (p < kf) {
                Lat(kf, p) = b[q];
            }
            if (p == kf) {
                Lat(kf, kf) = lkk;
                fcol[q] = c;
                diag[2 * kf] = lkk;
                diag[2 * kf + 1] = rkk;
            }
        }
        const bool lkk_ok = divisor_in_range(lkk);
        if (y_valid == kf) {
            // y_k = (f_c - L_k0 y_0 - L_k1 y_1 - ...) / L_kk, subtracted in order.
            double prod[PS];
#pragma unroll
            for (int q = 0; q < PS; ++q) {
                prod[q] = b[q] * y[q];
            }
            double v = fc;
            for (int j = 0; j < kf; ++j) {
                v = v - at_position<PS>(prod, j);
            }
            const double yk = divide_nonzero_by(v, lkk, rkk, lkk_ok);
#pragma unroll
            for (int q = 0; q < PS; ++q) {
                y[q] = position(q) == kf ? yk : y[q];
            }
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
        // Rows p + 1 .. kf - 1 move up one.  Position 64 q + 63 reads row 64 (q + 1), which slot
        // q + 1 overwrites, so the slots run in order with the wave's memory operations complete
        // in between.
#pragma unroll
        for (int q = 0; q < PS; ++q) {
            if (q > 0) {
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            }
            const int r = position(q);
            for (int c = 0; c < kf; ++c) {
                const bool move = r >= p && r < kf - 1 && c <= r + 1;
                const double v = move ? Lat(r + 1, c) : 0.0;
                if (move) {
                    Lat(r, c) = v;
                }
            }
        }
        for (int column = p; column < kf - 1; ++column) {
            const double diagonal = Luni(column, column);
            const double above = Luni(column, column + 1);
            const double length = hypot_fixed(diagonal, above);
            const double rl = division_reciprocal(length);
            const double cosine = divide_by(diagonal, length, rl);
            const double sine = divide_by(above, length, rl);
#pragma unroll
            for (int q = 0; q < PS; ++q) {
                const int r = position(q);
                if (r >= column && r < kf - 1) {
                    const double left = Lat(r, column);
                    const double right = Lat(r, column + 1);
                    Lat(r, column) = cosine * left + sine * right;
                    Lat(r, column + 1) = cosine * right - sine * left;
                }
            }
        }
        // The rotated diagonal and its reciprocals (each lane its own positions).
        bool out_of_range = false;
#pragma unroll
        for (int q = 0; q < PS; ++q) {
            const int r = position(q);
            if (r < kf - 1) {
                double dl = diag[2 * r];
                if (r >= p) {
                    dl = Lat(r, r);
                    diag[2 * r] = dl;
                    diag[2 * r + 1] = division_reciprocal(dl);
                }
                out_of_range = out_of_range || !divisor_in_range(dl);
            }
        }
        diag_ok = ballot(out_of_range) == 0ull;
        // Position r takes position r + 1's column (lane 63 of a slot from lane 0 of the next).
        int from[PS];
#pragma unroll
        for (int q = 0; q < PS; ++q) {
            from[q] = gather(fcol[q], (l + 1) % WAVE);
        }
#pragma unroll
        for (int q = 0; q + 1 < PS; ++q) {
            const int carry = readlane(fcol[q + 1], 0);
            from[q] = l == WAVE - 1 ? carry : from[q];
        }
#pragma unroll
        for (int q = 0; q < PS; ++q) {
            const int r = position(q);
            fcol[q] = (r >= p && r < kf - 1) ? from[q] : fcol[q];
        }
#pragma unroll
        for (int t = 0; t < NKS; ++t) {
            pos[t] = pos[t] == p ? -1 : (pos[t] > p ? pos[t] - 1 : pos[t]);
        }
        kf -= 1;
        y_valid = y_valid < p ? y_valid : p;
    }

    // Drops the positions whose column is inactive, last to first (GrowingCholesky::keep).
    __device__ __forceinline__ void keep() {
        int actbits = 0;
#pragma unroll
        for (int t = 0; t < NKS; ++t) {
            actbits |= act[t] ? (1 << t) : 0;
        }
        unsigned long long mask[PS];
#pragma unroll
        for (int q = 0; q < PS; ++q) {
            const int r = position(q);
            const int c = r < kf ? fcol[q] : 0;
            const int bits = gather(actbits, c % WAVE);
            const bool leaves = r < kf && !((bits >> (c / WAVE)) & 1);
            mask[q] = ballot(leaves);
        }
#pragma unroll
        for (int q = PS - 1; q >= 0; --q) {
            while (mask[q]) {
                const int b = 63 - __clzll(static_cast<long long>(mask[q]));
                mask[q] &= ~(1ull << b);
                remove(WAVE * q + b);
            }
        }
    }

    // cand = the solution of H_AA z = f_A, scattered to the columns.
    __device__ __forceinline__ void solve() {
        double b[PS];
#pragma unroll
        for (int q = 0; q < PS; ++q) {
            const int p = position(q);
            const double fv = f_at(p < kf ? fcol[q] : 0);
            b[q] = p < kf ? (p < y_valid ? y[q] : fv) : 0.0;
        }
        forward_from(b, y_valid, kf);
#pragma unroll
        for (int q = 0; q < PS; ++q) {
            y[q] = b[q];
        }
        y_valid = kf;
        backward(b, kf);
#pragma unroll
        for (int t = 0; t < NKS; ++t) {
            const int p = pos[t] < 0 ? 0 : pos[t];
            const double z = gather_position<PS>(b, p);
            cand[t] = pos[t] >= 0 ? z : 0.0;
        }
    }

    // Returns true when the multipliers are found (u), false when the QR path must take over.
    __device__ __forceinline__ bool run() {
        kf = 0;
#pragma unroll
        for (int q = 0; q < PS; ++q) {
            fcol[q] = 0;
            y[q] = 0.0;
        }
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
template <int NKS, int PS>
__global__ void __launch_bounds__(WAVE * 4) lsi_kernel(LsiArgs A) {
    extern __shared__ double lsi_diag[];
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / WAVE);
    const long gw = static_cast<long>(blockIdx.x) * (blockDim.x / WAVE) + w;
    const long n_waves = static_cast<long>(gridDim.x) * (blockDim.x / WAVE);
    double* slab = A.scratch + gw * A.per_wave;
    double* diag = lsi_diag + w * lsi_diag_doubles(PS);
    const int l = lane_id();
    const LsiDev& d = A.d;
    for (long v = gw; v < A.n_vox; v += n_waves) {
        double rot[PS];
        double raw[NKS];
        const double scale = lsi_rotate<NKS, PS>(d, A.signals + v * d.m, rot, raw);
        if (!(scale > 0.0)) {
            double x[PS];
            lsi_back<PS>(d, rot, x);
            lsi_store<PS>(d, A.x_out, v, x);
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
        LdpWave<NKS, PS> ldp(d, f, slab, diag);
        if (!ldp.run()) {
            if (l == 0) {
                A.fallback_list[atomicAdd(A.fallback_count, 1)] = static_cast<int>(v);
            }
            continue;
        }
        double x[PS];
        lsi_solution<NKS, PS>(d, f, ldp.u, scale, rot, x);
        lsi_store<PS>(d, A.x_out, v, x);
        if (l == 0) {
            A.converged[v] = 1;
            A.status[v] = 0;
        }
    }
}

// The QR path over the listed voxels (those the fast path handed over), one wave per block with
// its column slab (lsi_qr_slab(PS) doubles): in the block's dynamic LDS for PS = 1 (32 KB), else
// in the block's slab of A.scratch (A.per_wave doubles each).  Every lane reads and writes only
// its own rows of the slab.
template <int NKS, int PS>
__global__ void __launch_bounds__(WAVE) lsi_qr_kernel(LsiArgs A) {
    extern __shared__ double smem[];
    const long gw = blockIdx.x;
    const long n_waves = gridDim.x;
    double* slab = PS == 1 ? smem : A.scratch + gw * A.per_wave;
    const int l = lane_id();
    const LsiDev& d = A.d;
    const long count = A.list ? static_cast<long>(*A.list_count) : A.n_vox;
    for (long it = gw; it < count; it += n_waves) {
        const long v = A.list ? A.list[it] : it;
        double rot[PS];
        double raw[NKS];
        const double scale = lsi_rotate<NKS, PS>(d, A.signals + v * d.m, rot, raw);
        if (!(scale > 0.0)) {
            double x[PS];
            lsi_back<PS>(d, rot, x);
            lsi_store<PS>(d, A.x_out, v, x);
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
        const bool conv = ldp_by_qr<NKS, PS>(d, f, u, slab);
        double x[PS];
        lsi_solution<NKS, PS>(d, f, u, scale, rot, x);
        lsi_store<PS>(d, A.x_out, v, x);
        if (l == 0) {
            A.converged[v] = conv ? 1 : 0;
            A.status[v] = ST_QR_FALLBACK;
        }
    }
}

} // namespace lsd
