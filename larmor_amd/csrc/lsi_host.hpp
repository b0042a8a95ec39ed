// Host-side preparation of an LSI design (make_lsi_design, cpp/src/sd/lsi.cpp:19-69).
//
// [A; ridge I] P = Q R by column-pivoted Householder QR (Eigen 5.0.1's ColPivHouseholderQR:
// LAPACK xGEQP3 pivoting with the LAWN 176 norm downdate, beta = -sign(c0) ||x||, rank
// threshold max(rows, cols) eps), then the thin Q's measurement rows transposed, P R^-1, the
// unit-row constraints E = G P R^-1, E E^T and E's row sums.  Sums over rows of the QR are wave
// sums -- row r in lane r % 64, lanes combined by the XOR butterfly -- the order the device QR
// uses; every other sum runs in index order.  The device solves read these matrices, so a fit
// depends on them bit for bit.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

namespace lsd {

namespace host {

template <class Term>
double wave_sum(int64_t begin, int64_t end, Term term) {
    constexpr int W = 64;
    double p[W];
    for (double& x : p) x = 0.0;
    for (int64_t r = begin; r < end; ++r) {
        p[r % W] = p[r % W] + term(r);
    }
    for (int o = W / 2; o >= 1; o /= 2) {
        double q[W];
        for (int l = 0; l < W; ++l) q[l] = p[l] + p[l ^ o];
        std::memcpy(p, q, sizeof(p));
    }
    return p[0];
}

// Column-pivoted Householder QR of a row-major rows x cols matrix, in place.
struct ColPivQr {
    int64_t rows = 0, cols = 0;
    std::vector<double> a;      // row-major: R above the diagonal, reflectors below
    std::vector<double> tau;    // per reflector
    std::vector<int64_t> perm;  // perm[p]: the original column at pivot p
    int64_t nonzero_pivots = 0;
    double maxpivot = 0.0;

    double& at(int64_t r, int64_t c) { return a[r * cols + c]; }
    double at(int64_t r, int64_t c) const { return a[r * cols + c]; }

    void compute() {
        const int64_t size = std::min(rows, cols);
        const double eps = std::numeric_limits<double>::epsilon();
        tau.assign(size, 0.0);
        std::vector<double> nu(cols), nd(cols);
        for (int64_t c = 0; c < cols; ++c) {
            nd[c] = std::sqrt(wave_sum(0, rows, [&](int64_t r) { return at(r, c) * at(r, c); }));
            nu[c] = nd[c];
        }
        double max_norm = 0.0;
        for (int64_t c = 0; c < cols; ++c) {
            if (c == 0 || nu[c] > max_norm) max_norm = nu[c];
        }
        const double scaled = max_norm * eps;
        const double threshold_helper = (scaled * scaled) / static_cast<double>(rows);
        const double downdate_threshold = std::sqrt(eps);
        nonzero_pivots = size;
        maxpivot = 0.0;
        std::vector<int64_t> swaps(size);
        for (int64_t k = 0; k < size; ++k) {
            int64_t big = k;
            for (int64_t c = k + 1; c < cols; ++c) {
                if (nu[c] > nu[big]) big = c;
            }
            if (nonzero_pivots == size &&
                nu[big] * nu[big] < threshold_helper * static_cast<double>(rows - k)) {
                nonzero_pivots = k;
            }
            swaps[k] = big;
            if (big != k) {
                for (int64_t r = 0; r < rows; ++r) std::swap(at(r, k), at(r, big));
                std::swap(nu[k], nu[big]);
                std::swap(nd[k], nd[big]);
            }
            const double c0 = at(k, k);
            const double tail = wave_sum(k + 1, rows, [&](int64_t r) { return at(r, k) * at(r, k); });
            double t;
            double beta;
            if (tail <= std::numeric_limits<double>::min()) {
                t = 0.0;
                beta = c0;
                for (int64_t r = k + 1; r < rows; ++r) at(r, k) = 0.0;
            } else {
                beta = std::sqrt(c0 * c0 + tail);
                if (c0 >= 0.0) beta = -beta;
                const double denom = c0 - beta;
                for (int64_t r = k + 1; r < rows; ++r) at(r, k) = at(r, k) / denom;
                t = (beta - c0) / beta;
            }
            tau[k] = t;
            at(k, k) = beta;
            if (std::fabs(beta) > maxpivot) maxpivot = std::fabs(beta);
            if (rows - k == 1) {
                for (int64_t c = k + 1; c < cols; ++c) at(k, c) = at(k, c) * (1.0 - t);
            } else if (t != 0.0) {
                for (int64_t c = k + 1; c < cols; ++c) {
                    double tmp = wave_sum(k + 1, rows, [&](int64_t r) { return at(r, k) * at(r, c); });
                    tmp = tmp + at(k, c);
                    at(k, c) = at(k, c) - t * tmp;
                    for (int64_t r = k + 1; r < rows; ++r) {
                        at(r, c) = at(r, c) - (t * at(r, k)) * tmp;
                    }
                }
            }
            for (int64_t c = k + 1; c < cols; ++c) {
                if (nu[c] != 0.0) {
                    double temp = std::fabs(at(k, c)) / nu[c];
                    temp = (1.0 + temp) * (1.0 - temp);
                    temp = temp < 0.0 ? 0.0 : temp;
                    const double ratio = nu[c] / nd[c];
                    if (temp * (ratio * ratio) <= downdate_threshold) {
                        nd[c] = std::sqrt(
                            wave_sum(k + 1, rows, [&](int64_t r) { return at(r, c) * at(r, c); }));
                        nu[c] = nd[c];
                    } else {
                        nu[c] = nu[c] * std::sqrt(temp);
                    }
                }
            }
        }
        perm.resize(cols);
        for (int64_t c = 0; c < cols; ++c) perm[c] = c;
        for (int64_t k = 0; k < size; ++k) std::swap(perm[k], perm[swaps[k]]);
    }

    int64_t rank(double threshold) const {
        const double pre = std::fabs(maxpivot) * threshold;
        int64_t r = 0;
        for (int64_t i = 0; i < nonzero_pivots; ++i) r += std::fabs(at(i, i)) > pre ? 1 : 0;
        return r;
    }
};

struct LsiHostDesign {
    int64_t m = 0, n = 0, k = 0;
    std::vector<double> q_t, back, e, e_t, gram, abs_sums;
};

// Builds the design; returns an empty string or the reference's error message.
inline std::string make_lsi_design(const double* a, int64_t m, int64_t n, const double* g,
                                   int64_t k, int64_t g_cols, double ridge, LsiHostDesign& d) {
    if (g_cols != n) {
        return "the constraints have " + std::to_string(g_cols) + " columns but the design has " +
               std::to_string(n);
    }
    if (!(ridge >= 0.0)) {
        return "the ridge must be at least 0";
    }
    const bool has_ridge = ridge > 0.0;
    ColPivQr qr;
    qr.rows = m + (has_ridge ? n : 0);
    qr.cols = n;
    qr.a.assign(static_cast<size_t>(qr.rows * n), 0.0);
    for (int64_t r = 0; r < m; ++r)
        for (int64_t c = 0; c < n; ++c) qr.at(r, c) = a[r * n + c];
    if (has_ridge)
        for (int64_t c = 0; c < n; ++c) qr.at(m + c, c) = ridge;
    qr.compute();
    const double threshold =
        static_cast<double>(std::max(qr.rows, n)) * std::numeric_limits<double>::epsilon();
    if (qr.rows < n || qr.rank(threshold) < n) {
        return "the design does not have full column rank: the measurements cannot determine "
               "all " + std::to_string(n) + " parameters";
    }
    const int64_t rows = qr.rows;
    d.m = m;
    d.n = n;
    d.k = k;
    d.q_t.assign(static_cast<size_t>(n * m), 0.0);
    std::vector<double> col(static_cast<size_t>(rows));
    for (int64_t j = 0; j < n; ++j) {
        std::fill(col.begin(), col.end(), 0.0);
        col[j] = 1.0;
        for (int64_t h = n - 1; h >= 0; --h) {
            const double t = qr.tau[h];
            if (rows - h == 1) {
                col[h] = col[h] * (1.0 - t);
            } else if (t != 0.0) {
                double tmp = wave_sum(h + 1, rows, [&](int64_t r) { return qr.at(r, h) * col[r]; });
                tmp = tmp + col[h];
                col[h] = col[h] - t * tmp;
                for (int64_t r = h + 1; r < rows; ++r) col[r] = col[r] - (t * qr.at(r, h)) * tmp;
            }
        }
        for (int64_t r = 0; r < m; ++r) d.q_t[j * m + r] = col[r];
    }
    std::vector<double> rinv(static_cast<size_t>(n * n), 0.0);
    std::vector<double> x(static_cast<size_t>(n));
    for (int64_t j = 0; j < n; ++j) {
        std::fill(x.begin(), x.end(), 0.0);
        x[j] = 1.0;
        for (int64_t i = n - 1; i >= 0; --i) {
            if (x[i] != 0.0) x[i] = x[i] / qr.at(i, i);
            for (int64_t r = 0; r < i; ++r) x[r] = x[r] - qr.at(r, i) * x[i];
        }
        for (int64_t i = 0; i < n; ++i) rinv[i * n + j] = x[i];
    }
    d.back.assign(static_cast<size_t>(n * n), 0.0);
    for (int64_t p = 0; p < n; ++p)
        for (int64_t j = 0; j < n; ++j) d.back[qr.perm[p] * n + j] = rinv[p * n + j];
    d.e.assign(static_cast<size_t>(k * n), 0.0);
    for (int64_t c = 0; c < k; ++c) {
        for (int64_t j = 0; j < n; ++j) {
            double sum = 0.0;
            for (int64_t i = 0; i < n; ++i) sum = sum + g[c * n + i] * d.back[i * n + j];
            d.e[c * n + j] = sum;
        }
        double sq = 0.0;
        for (int64_t j = 0; j < n; ++j) sq = sq + d.e[c * n + j] * d.e[c * n + j];
        const double norm = std::sqrt(sq);
        if (norm > 0.0)
            for (int64_t j = 0; j < n; ++j) d.e[c * n + j] = d.e[c * n + j] / norm;
    }
    d.e_t.assign(static_cast<size_t>(n * k), 0.0);
    for (int64_t c = 0; c < k; ++c)
        for (int64_t j = 0; j < n; ++j) d.e_t[j * k + c] = d.e[c * n + j];
    d.gram.assign(static_cast<size_t>(k * k), 0.0);
    d.abs_sums.assign(static_cast<size_t>(k), 0.0);
    for (int64_t p = 0; p < k; ++p) {
        for (int64_t q = 0; q < k; ++q) {
            double sum = 0.0;
            for (int64_t j = 0; j < n; ++j) sum = sum + d.e[p * n + j] * d.e[q * n + j];
            d.gram[p * k + q] = sum;
        }
        double s = 0.0;
        for (int64_t j = 0; j < n; ++j) s = s + std::fabs(d.e[p * n + j]);
        d.abs_sums[p] = s;
    }
    return "";
}

} // namespace host

} // namespace lsd
