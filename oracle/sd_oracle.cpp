// Test infrastructure only: the CPU FP64 oracle of larmor's per-voxel spherical deconvolution.
//
// Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load this library, and
// only as the checker or the timed CPU baseline; the product path (larmor_amd) never does.
//
// Each function restates the reference algorithm step by step, with the reference's step order
// and tie rules, written from its source (paths relative to the reference repository):
//   cpp/src/sd/nnls.cpp        Lawson-Hanson NNLS, pivoted-QR least squares, ridge NNLS with a
//                              growing / downdated Cholesky factor and the QR fallback
//   cpp/src/sd/fractions.cpp   median, compute_fractions
//   cpp/src/sd/multi_tissue.cpp fit_multi_tissue_block
//   cpp/src/sd/solver.cpp      L2NnlsSolver::solve_block, DrlSolver::solve_block
//   cpp/src/sd/drl.cpp         damped_richardson_lucy (per voxel)
//   cpp/src/sd/elastic_net.cpp standardize_columns, solve_elastic_net (coordinate descent)
//   cpp/src/sd/csd.cpp         make_csd_problem, solve_csd (LLT restated as below)
//   cpp/src/sh/harmonics.cpp   fill_legendre, Legendre derivatives, derivatives
//   cpp/src/sh/peaks.cpp       climb, voxel_peaks
//   cpp/src/sd/lsi.cpp         make_lsi_design, ldp_multipliers, solve_lsi
//   cpp/src/common/parallel.hpp parallel_for_blocks (blocks of 128 voxels, dynamic)
// Eigen's ColPivHouseholderQR (Eigen 5.0.1, the reference's pinned dependency) is restated from
// its published algorithm (LAPACK xGEQP3 column pivoting with the LAWN 176 norm downdate).
//
// Rounding order.  The reference's sums run in Eigen's vectorised order, which no restatement
// can reproduce, so bit-level agreement with the reference is not claimed.  This oracle fixes
// its own order, the one the HIP kernels use, so that the oracle and the kernels agree to the bit
// on the same inputs:
//  - a sum over the rows (measurements) of a design -- the norms, dot products and gradients of
//    the pivoted QR and of the generic NNLS, and the column sums of the default tolerance -- is a
//    "wave sum": row r belongs to lane r mod 64, each lane adds its rows in increasing order
//    starting from 0, and the 64 lane partials are combined by an XOR butterfly,
//    p[l] = p[l] + p[l ^ o] for o = 32, 16, 8, 4, 2, 1;
//  - every other sum (a correlation K^T s over measurements, a sum over columns or over factor
//    positions) is sequential in index order from 0;
//  - every triangular solve is column-oriented: a value is final once the terms of all earlier
//    (later, for L^T) positions are subtracted, one at a time, in the order the solve meets them.
// hypot is evaluated as m * sqrt(1 + (s / m)^2), m the larger magnitude, here and on the GPU.
//
// Compiled with -O3 -ffp-contract=off (no fused multiply-add), as the reference is
// (cpp/CMakeLists.txt sets -ffp-contract=off).

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <limits>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "sh_series.hpp"

namespace {

using Index = std::int64_t;
constexpr double EPS = std::numeric_limits<double>::epsilon();

// A Cholesky pivot below this share of its diagonal entry counts as a dependent column
// (nnls.cpp MIN_RELATIVE_PIVOT).
constexpr double MIN_RELATIVE_PIVOT = 1e-12;
// The largest (||K||_F^2 + r^2) / r^2 for which the ridge fits use the normal equations
// (nnls.cpp MAX_NORMAL_EQUATIONS_CONDITION).
constexpr double MAX_NORMAL_EQUATIONS_CONDITION = 1e8;
// FOD values this close below the median count as equal to it (fractions.cpp).
constexpr double MEDIAN_TIE_TOLERANCE = 1e-9;
// Keeps the normalisation of the final fractions finite (multi_tissue.cpp).
constexpr double FRACTION_SUM_OFFSET = 1e-16;

thread_local std::string g_error;

// Row-major dense matrix.
struct Mat {
    Index rows = 0;
    Index cols = 0;
    std::vector<double> v;
    Mat() = default;
    Mat(Index r, Index c, double fill = 0.0) : rows(r), cols(c), v(static_cast<size_t>(r * c), fill) {}
    double& operator()(Index r, Index c) { return v[static_cast<size_t>(r * cols + c)]; }
    double operator()(Index r, Index c) const { return v[static_cast<size_t>(r * cols + c)]; }
};
using Vec = std::vector<double>;

Mat from_ptr(const double* p, Index rows, Index cols) {
    Mat m(rows, cols);
    if (rows * cols > 0) {
        std::memcpy(m.v.data(), p, sizeof(double) * static_cast<size_t>(rows * cols));
    }
    return m;
}

double hypot_fixed(double a, double b) {
    const double x = std::fabs(a);
    const double y = std::fabs(b);
    const double big = x > y ? x : y;
    const double small = x > y ? y : x;
    if (big == 0.0) {
        return 0.0;
    }
    const double r = small / big;
    return big * std::sqrt(1.0 + r * r);
}

// The wave sum of term(r) over rows r in [begin, end) (see the header).
constexpr int WAVE = 64;
template <class Term>
double wave_sum(Index begin, Index end, Term term) {
    double partial[WAVE];
    for (int l = 0; l < WAVE; ++l) {
        partial[l] = 0.0;
    }
    for (Index r = begin; r < end; ++r) {
        partial[r % WAVE] += term(r);
    }
    for (int o = WAVE / 2; o >= 1; o /= 2) {
        double next[WAVE];
        for (int l = 0; l < WAVE; ++l) {
            next[l] = partial[l] + partial[l ^ o];
        }
        std::memcpy(partial, next, sizeof(partial));
    }
    return partial[0];
}

// Per-voxel trace counters of one NNLS solve.
struct NnlsTrace {
    Index outer = 0;      // entering columns accepted
    Index inner = 0;      // inner (feasibility) steps
    Index max_active = 0; // largest active-set size seen
    Index qr_fallback = 0;
    Index final_active = 0;
};

// ---------------------------------------------------------------------------------------------
// Column-pivoted Householder QR (Eigen ColPivHouseholderQR, restated).
// ---------------------------------------------------------------------------------------------
struct PivotedQR {
    Mat qr;
    Vec hcoeffs;
    std::vector<Index> perm; // perm[p]: the original column in pivot position p
    Index nonzero_pivots = 0;
    double maxpivot = 0.0;

    void compute(Mat a) {
        qr = std::move(a);
        const Index rows = qr.rows;
        const Index cols = qr.cols;
        const Index size = std::min(rows, cols);
        hcoeffs.assign(static_cast<size_t>(size), 0.0);
        Vec norms_updated(static_cast<size_t>(cols));
        Vec norms_direct(static_cast<size_t>(cols));
        for (Index c = 0; c < cols; ++c) {
            const double s = wave_sum(0, rows, [&](Index r) { return qr(r, c) * qr(r, c); });
            norms_direct[c] = std::sqrt(s);
            norms_updated[c] = norms_direct[c];
        }
        double max_norm = 0.0;
        for (Index c = 0; c < cols; ++c) {
            if (c == 0 || norms_updated[c] > max_norm) {
                max_norm = norms_updated[c];
            }
        }
        const double scaled = max_norm * EPS;
        const double threshold_helper = (scaled * scaled) / static_cast<double>(rows);
        const double norm_downdate_threshold = std::sqrt(EPS);
        nonzero_pivots = size;
        maxpivot = 0.0;
        std::vector<Index> transpositions(static_cast<size_t>(size));

        for (Index k = 0; k < size; ++k) {
            Index biggest = k;
            for (Index c = k + 1; c < cols; ++c) {
                if (norms_updated[c] > norms_updated[biggest]) {
                    biggest = c;
                }
            }
            const double biggest_sq = norms_updated[biggest] * norms_updated[biggest];
            if (nonzero_pivots == size &&
                biggest_sq < threshold_helper * static_cast<double>(rows - k)) {
                nonzero_pivots = k;
            }
            transpositions[k] = biggest;
            if (k != biggest) {
                for (Index r = 0; r < rows; ++r) {
                    std::swap(qr(r, k), qr(r, biggest));
                }
                std::swap(norms_updated[k], norms_updated[biggest]);
                std::swap(norms_direct[k], norms_direct[biggest]);
            }

            // Householder vector of column k, rows k..rows-1, essential part stored below.
            const double c0 = qr(k, k);
            const double tail_sq =
                wave_sum(k + 1, rows, [&](Index r) { return qr(r, k) * qr(r, k); });
            double tau;
            double beta;
            if (tail_sq <= std::numeric_limits<double>::min()) {
                tau = 0.0;
                beta = c0;
                for (Index r = k + 1; r < rows; ++r) {
                    qr(r, k) = 0.0;
                }
            } else {
                beta = std::sqrt(c0 * c0 + tail_sq);
                if (c0 >= 0.0) {
                    beta = -beta;
                }
                const double denom = c0 - beta;
                for (Index r = k + 1; r < rows; ++r) {
                    qr(r, k) = qr(r, k) / denom;
                }
                tau = (beta - c0) / beta;
            }
            hcoeffs[k] = tau;
            qr(k, k) = beta;
            if (std::fabs(beta) > maxpivot) {
                maxpivot = std::fabs(beta);
            }

            // Apply H_k to the columns right of k.
            if (rows - k == 1) {
                for (Index c = k + 1; c < cols; ++c) {
                    qr(k, c) *= (1.0 - tau);
                }
            } else if (tau != 0.0) {
                for (Index c = k + 1; c < cols; ++c) {
                    double tmp =
                        wave_sum(k + 1, rows, [&](Index r) { return qr(r, k) * qr(r, c); });
                    tmp += qr(k, c);
                    qr(k, c) -= tau * tmp;
                    for (Index r = k + 1; r < rows; ++r) {
                        qr(r, c) -= (tau * qr(r, k)) * tmp;
                    }
                }
            }

            // Norm downdate (LAWN 176).
            for (Index c = k + 1; c < cols; ++c) {
                if (norms_updated[c] != 0.0) {
                    double temp = std::fabs(qr(k, c)) / norms_updated[c];
                    temp = (1.0 + temp) * (1.0 - temp);
                    temp = temp < 0.0 ? 0.0 : temp;
                    const double ratio = norms_updated[c] / norms_direct[c];
                    const double temp2 = temp * (ratio * ratio);
                    if (temp2 <= norm_downdate_threshold) {
                        const double s =
                            wave_sum(k + 1, rows, [&](Index r) { return qr(r, c) * qr(r, c); });
                        norms_direct[c] = std::sqrt(s);
                        norms_updated[c] = norms_direct[c];
                    } else {
                        norms_updated[c] *= std::sqrt(temp);
                    }
                }
            }
        }
        perm.resize(static_cast<size_t>(cols));
        for (Index c = 0; c < cols; ++c) {
            perm[c] = c;
        }
        for (Index k = 0; k < size; ++k) {
            std::swap(perm[k], perm[transpositions[k]]);
        }
    }

    Index rank(double threshold) const {
        const double premultiplied = std::fabs(maxpivot) * threshold;
        Index result = 0;
        for (Index i = 0; i < nonzero_pivots; ++i) {
            result += std::fabs(qr(i, i)) > premultiplied ? 1 : 0;
        }
        return result;
    }

    // Q^T d restricted to the first `length` reflectors (H_0 applied first).
    void apply_qt(Vec& d, Index length) const {
        const Index rows = qr.rows;
        for (Index k = 0; k < length; ++k) {
            const double tau = hcoeffs[k];
            if (rows - k == 1) {
                d[k] *= (1.0 - tau);
            } else if (tau != 0.0) {
                double tmp = wave_sum(k + 1, rows, [&](Index r) { return qr(r, k) * d[r]; });
                tmp += d[k];
                d[k] -= tau * tmp;
                for (Index r = k + 1; r < rows; ++r) {
                    d[r] -= (tau * qr(r, k)) * tmp;
                }
            }
        }
    }
};

// least_squares (nnls.cpp): coefficients of data on the columns of `design`; columns past the
// numerical rank of the pivoted QR get zero.
Vec least_squares(const Mat& design, const Vec& data) {
    PivotedQR qr;
    qr.compute(design);
    const double threshold =
        static_cast<double>(std::max(design.rows, design.cols)) * EPS;
    const Index rank = qr.rank(threshold);
    Vec rotated = data;
    qr.apply_qt(rotated, rank);
    // Back substitution on R11, column-oriented (one term subtracted at a time).
    for (Index i = rank - 1; i >= 0; --i) {
        if (rotated[i] != 0.0) {
            rotated[i] = rotated[i] / qr.qr(i, i);
        }
        for (Index r = 0; r < i; ++r) {
            rotated[r] -= qr.qr(r, i) * rotated[i];
        }
    }
    Vec coefficients(static_cast<size_t>(design.cols), 0.0);
    for (Index p = 0; p < rank; ++p) {
        coefficients[qr.perm[p]] = rotated[p];
    }
    return coefficients;
}

// How the generic NNLS sums a column over the rows: as a wave sum, or in row order (the LDP QR
// fallback, where the GPU gives each lane whole columns).
enum class ColumnSum { Wave, InOrder };

template <class Term>
double column_sum(ColumnSum order, Index rows, Term term) {
    if (order == ColumnSum::Wave) {
        return wave_sum(0, rows, term);
    }
    double s = 0.0;
    for (Index r = 0; r < rows; ++r) {
        s = s + term(r);
    }
    return s;
}

double default_gradient_tolerance(const Mat& design, ColumnSum order = ColumnSum::Wave) {
    double norm1 = 0.0;
    if (design.rows * design.cols > 0) {
        for (Index c = 0; c < design.cols; ++c) {
            const double s = column_sum(order, design.rows,
                                        [&](Index r) { return std::fabs(design(r, c)); });
            if (c == 0 || s > norm1) {
                norm1 = s;
            }
        }
    }
    return 10.0 * EPS * norm1 * static_cast<double>(std::max(design.rows, design.cols));
}

// ---------------------------------------------------------------------------------------------
// Lawson-Hanson (nnls.cpp lawson_hanson).
// ---------------------------------------------------------------------------------------------
struct NnlsResult {
    Vec solution;
    bool converged = true;
};

template <class GradientAt, class FitOn>
NnlsResult lawson_hanson(Index n, double tol, Index max_inner, GradientAt gradient_at,
                         FitOn fit_on, NnlsTrace* trace) {
    std::vector<char> active(static_cast<size_t>(n), 0);
    Vec x(static_cast<size_t>(n), 0.0);
    Vec g = gradient_at(x);
    Index n_inner = 0;
    auto count_active = [&] {
        Index k = 0;
        for (Index c = 0; c < n; ++c) {
            k += active[c] ? 1 : 0;
        }
        return k;
    };
    while (true) {
        Index entering = -1;
        for (Index c = 0; c < n; ++c) {
            if (!active[c] && (entering < 0 || g[c] > g[entering])) {
                entering = c;
            }
        }
        if (entering < 0 || g[entering] <= tol) {
            break;
        }
        active[entering] = 1;
        if (trace) {
            trace->outer += 1;
            trace->max_active = std::max(trace->max_active, count_active());
        }
        Vec cand = fit_on(active);
        while (true) {
            bool any_nonpositive = false;
            for (Index c = 0; c < n; ++c) {
                if (active[c] && cand[c] <= 0.0) {
                    any_nonpositive = true;
                    break;
                }
            }
            if (!any_nonpositive) {
                break;
            }
            if (++n_inner > max_inner) {
                if (trace) {
                    trace->inner = n_inner;
                    trace->final_active = count_active();
                }
                return {cand, false};
            }
            double step = std::numeric_limits<double>::infinity();
            Index blocking = 0;
            for (Index c = 0; c < n; ++c) {
                if (active[c] && cand[c] <= 0.0) {
                    const double distance = x[c] - cand[c];
                    const double s = distance > 0.0 ? x[c] / distance : 0.0;
                    if (s < step) {
                        step = s;
                        blocking = c;
                    }
                }
            }
            for (Index c = 0; c < n; ++c) {
                x[c] = x[c] + step * (cand[c] - x[c]);
            }
            x[blocking] = 0.0;
            active[blocking] = 0;
            for (Index c = 0; c < n; ++c) {
                active[c] = (active[c] && std::fabs(x[c]) >= tol) ? 1 : 0;
            }
            cand = fit_on(active);
        }
        x = cand;
        g = gradient_at(x);
    }
    if (trace) {
        trace->inner = n_inner;
        trace->final_active = count_active();
    }
    return {x, true};
}

std::vector<Index> active_indices(const std::vector<char>& active) {
    std::vector<Index> idx;
    for (Index c = 0; c < static_cast<Index>(active.size()); ++c) {
        if (active[c]) {
            idx.push_back(c);
        }
    }
    return idx;
}

// solve_nnls_with_limit (nnls.cpp): the generic NNLS, a pivoted QR per least-squares fit.
NnlsResult solve_nnls_with_limit(const Mat& design, const Vec& data, double tol,
                                 Index max_inner, NnlsTrace* trace,
                                 ColumnSum order = ColumnSum::Wave) {
    const Index n = design.cols;
    const Index m = design.rows;
    auto gradient_at = [&](const Vec& x) {
        Vec r(static_cast<size_t>(m));
        for (Index i = 0; i < m; ++i) {
            double s = 0.0;
            for (Index c = 0; c < n; ++c) {
                s += design(i, c) * x[c];
            }
            r[i] = data[i] - s;
        }
        Vec g(static_cast<size_t>(n));
        for (Index c = 0; c < n; ++c) {
            g[c] = column_sum(order, m, [&](Index i) { return design(i, c) * r[i]; });
        }
        return g;
    };
    auto fit_on = [&](const std::vector<char>& active) {
        const std::vector<Index> idx = active_indices(active);
        Vec full(static_cast<size_t>(n), 0.0);
        if (idx.empty()) {
            return full;
        }
        Mat sub(m, static_cast<Index>(idx.size()));
        for (Index i = 0; i < m; ++i) {
            for (size_t p = 0; p < idx.size(); ++p) {
                sub(i, static_cast<Index>(p)) = design(i, idx[p]);
            }
        }
        const Vec coef = least_squares(sub, data);
        for (size_t p = 0; p < idx.size(); ++p) {
            full[idx[p]] = coef[p];
        }
        return full;
    };
    return lawson_hanson(n, tol, max_inner, gradient_at, fit_on, trace);
}

// ---------------------------------------------------------------------------------------------
// Ridge NNLS (nnls.cpp make_ridge_design, solve_ridge_nnls, ActiveCholesky,
// remove_cholesky_column).
// ---------------------------------------------------------------------------------------------
struct RidgeDesign {
    Mat kernel;
    Mat gram;
    double ridge = 0.0;
    bool normal_equations = false;
};

RidgeDesign make_ridge_design(const Mat& kernel, double ridge) {
    RidgeDesign d;
    d.kernel = kernel;
    const Index m = kernel.rows;
    const Index n = kernel.cols;
    d.gram = Mat(n, n);
    for (Index a = 0; a < n; ++a) {
        for (Index b = 0; b < n; ++b) {
            double s = 0.0;
            for (Index i = 0; i < m; ++i) {
                s += kernel(i, a) * kernel(i, b);
            }
            d.gram(a, b) = s;
        }
    }
    const double r2 = ridge * ridge;
    for (Index a = 0; a < n; ++a) {
        d.gram(a, a) += r2;
    }
    double fro = 0.0;
    for (double value : kernel.v) {
        fro += value * value;
    }
    d.ridge = ridge;
    d.normal_equations = r2 == 0.0 || fro + r2 <= MAX_NORMAL_EQUATIONS_CONDITION * r2;
    return d;
}

// Removes the column at `position` from the lower factor of n_columns columns (nnls.cpp).
void remove_cholesky_column(Mat& factor, Index n_columns, Index position) {
    for (Index row = position; row < n_columns - 1; ++row) {
        for (Index c = 0; c < n_columns; ++c) {
            factor(row, c) = factor(row + 1, c);
        }
    }
    for (Index column = position; column < n_columns - 1; ++column) {
        const double diagonal = factor(column, column);
        const double above = factor(column, column + 1);
        const double length = hypot_fixed(diagonal, above);
        const double cosine = diagonal / length;
        const double sine = above / length;
        for (Index row = column; row < n_columns - 1; ++row) {
            const double left = factor(row, column);
            const double right = factor(row, column + 1);
            factor(row, column) = cosine * left + sine * right;
            factor(row, column + 1) = cosine * right - sine * left;
        }
    }
}

// The ridge solve's growing factor: its triangular solves and the new row's squared norm
// accumulate by fused multiply-adds in order (one rounding per term), as do the correlations
// and the gradient of solve_ridge_nnls.
class ActiveCholesky {
public:
    explicit ActiveCholesky(const Mat& g)
        : gram(g), factor(g.rows, g.rows), in_factor(static_cast<size_t>(g.rows), 0) {}

    bool update(const std::vector<char>& active) {
        for (Index pos = static_cast<Index>(columns.size()) - 1; pos >= 0; --pos) {
            const Index column = columns[pos];
            if (!active[column]) {
                remove_cholesky_column(factor, static_cast<Index>(columns.size()), pos);
                columns.erase(columns.begin() + pos);
                in_factor[column] = 0;
            }
        }
        for (Index column = 0; column < static_cast<Index>(active.size()); ++column) {
            if (active[column] && !in_factor[column] && !append(column)) {
                return false;
            }
        }
        return true;
    }

    Vec solve(const Vec& rhs) const {
        const Index k = static_cast<Index>(columns.size());
        Vec w(static_cast<size_t>(k));
        for (Index p = 0; p < k; ++p) {
            w[p] = rhs[columns[p]];
        }
        forward(w, k);
        // L^T z = w, column-oriented: z_j final, then w_i -= L(j, i) z_j for i < j.
        for (Index j = k - 1; j >= 0; --j) {
            if (w[j] != 0.0) {
                w[j] = w[j] / factor(j, j);
            }
            for (Index i = 0; i < j; ++i) {
                w[i] = std::fma(-factor(j, i), w[j], w[i]);
            }
        }
        Vec full(rhs.size(), 0.0);
        for (Index p = 0; p < k; ++p) {
            full[columns[p]] = w[p];
        }
        return full;
    }

    Index size() const { return static_cast<Index>(columns.size()); }

private:
    // L w = b in place on the first k rows, column-oriented.
    void forward(Vec& w, Index k) const {
        for (Index j = 0; j < k; ++j) {
            if (w[j] != 0.0) {
                w[j] = w[j] / factor(j, j);
            }
            for (Index i = j + 1; i < k; ++i) {
                w[i] = std::fma(-factor(i, j), w[j], w[i]);
            }
        }
    }

    bool append(Index column) {
        const Index pos = static_cast<Index>(columns.size());
        Vec cross(static_cast<size_t>(pos));
        for (Index r = 0; r < pos; ++r) {
            cross[r] = gram(columns[r], column);
        }
        forward(cross, pos);
        double sq = 0.0;
        for (Index r = 0; r < pos; ++r) {
            sq = std::fma(cross[r], cross[r], sq);
        }
        const double pivot = gram(column, column) - sq;
        if (!(pivot > MIN_RELATIVE_PIVOT * gram(column, column))) {
            return false;
        }
        for (Index r = 0; r < pos; ++r) {
            factor(pos, r) = cross[r];
        }
        factor(pos, pos) = std::sqrt(pivot);
        columns.push_back(column);
        in_factor[column] = 1;
        return true;
    }

    const Mat& gram;
    Mat factor;
    std::vector<Index> columns;
    std::vector<char> in_factor;
};

Vec solve_ridge_nnls(const RidgeDesign& design, const Vec& signal, double l1, double tol,
                     NnlsTrace* trace) {
    if (l1 > 0 && design.ridge == 0) {
        throw std::invalid_argument("l1 > 0 needs ridge > 0.");
    }
    const Index m = design.kernel.rows;
    const Index n = design.kernel.cols;
    Vec correlations(static_cast<size_t>(n));
    for (Index c = 0; c < n; ++c) {
        double s = 0.0;
        for (Index i = 0; i < m; ++i) {
            s = std::fma(design.kernel(i, c), signal[i], s);
        }
        correlations[c] = s - l1;
    }
    auto gradient_at = [&](const Vec& x) {
        Vec g = correlations;
        for (Index c = 0; c < n; ++c) {
            if (x[c] != 0.0) {
                const double xc = x[c];
                for (Index i = 0; i < n; ++i) {
                    g[i] = std::fma(-xc, design.gram(i, c), g[i]);
                }
            }
        }
        return g;
    };
    auto fit_by_qr = [&](const std::vector<char>& active) {
        const std::vector<Index> idx = active_indices(active);
        const Index k = static_cast<Index>(idx.size());
        Vec full(static_cast<size_t>(n), 0.0);
        if (k == 0) {
            return full;
        }
        const bool has_ridge = design.ridge > 0.0;
        Mat sub(m + (has_ridge ? k : 0), k);
        for (Index i = 0; i < m; ++i) {
            for (Index p = 0; p < k; ++p) {
                sub(i, p) = design.kernel(i, idx[p]);
            }
        }
        Vec data(static_cast<size_t>(sub.rows), 0.0);
        for (Index i = 0; i < m; ++i) {
            data[i] = signal[i];
        }
        if (has_ridge) {
            for (Index p = 0; p < k; ++p) {
                sub(m + p, p) = design.ridge;
                data[m + p] = -l1 / design.ridge;
            }
        }
        const Vec coef = least_squares(sub, data);
        for (Index p = 0; p < k; ++p) {
            full[idx[p]] = coef[p];
        }
        return full;
    };
    const Index max_inner = 3 * n;
    if (!design.normal_equations) {
        return lawson_hanson(n, tol, max_inner, gradient_at, fit_by_qr, trace).solution;
    }
    ActiveCholesky cholesky(design.gram);
    bool factor_failed = false;
    auto fit_on = [&](const std::vector<char>& active) {
        factor_failed = factor_failed || !cholesky.update(active);
        if (factor_failed && trace) {
            trace->qr_fallback = 1;
        }
        return factor_failed ? fit_by_qr(active) : cholesky.solve(correlations);
    };
    return lawson_hanson(n, tol, max_inner, gradient_at, fit_on, trace).solution;
}

// ---------------------------------------------------------------------------------------------
// Fractions (fractions.cpp) and the multi-tissue loop (multi_tissue.cpp).
// ---------------------------------------------------------------------------------------------
bool less_nan_last(double left, double right) {
    return std::isnan(right) ? !std::isnan(left) : left < right;
}

double median(const Vec& values) {
    Vec sorted = values;
    if (sorted.empty()) {
        return std::numeric_limits<double>::quiet_NaN();
    }
    std::sort(sorted.begin(), sorted.end(), less_nan_last);
    const size_t mid = sorted.size() / 2;
    if (sorted.size() % 2 == 1) {
        return sorted[mid];
    }
    return (sorted[mid - 1] + sorted[mid]) / 2.0;
}

struct TissueKernel {
    Mat anisotropic;        // (m, ca)
    Mat isotropic;          // (m, ci)
    Mat weighted_isotropic; // (m, ci)
    std::vector<std::pair<Index, Index>> blocks;
};

// compute_fractions (fractions.cpp). groups: one range per group column.
Vec compute_fractions(const Vec& fod, const TissueKernel& kernel, const Vec& signal,
                      bool lumped, bool has_tolerance, double tolerance, NnlsTrace* trace) {
    const double threshold = median(fod) * (1.0 - MEDIAN_TIE_TOLERANCE);
    Vec thr(fod.size());
    for (size_t c = 0; c < fod.size(); ++c) {
        thr[c] = fod[c] < threshold ? 0.0 : fod[c];
    }
    std::vector<std::pair<Index, Index>> groups;
    if (lumped) {
        groups.push_back({0, kernel.anisotropic.cols});
    } else {
        groups = kernel.blocks;
    }
    const Index m = kernel.anisotropic.rows;
    const Index ng = static_cast<Index>(groups.size());
    const Index ci = kernel.isotropic.cols;
    Mat columns(m, ng + ci);
    for (Index g = 0; g < ng; ++g) {
        const auto [begin, end] = groups[g];
        Vec gs(static_cast<size_t>(m));
        for (Index i = 0; i < m; ++i) {
            double s = 0.0;
            for (Index c = begin; c < end; ++c) {
                s += kernel.anisotropic(i, c) * thr[c];
            }
            gs[i] = s;
        }
        if (m > 0) {
            // maxCoeff<PropagateNumbers>: the largest non-NaN value.
            double largest = std::numeric_limits<double>::quiet_NaN();
            for (Index i = 0; i < m; ++i) {
                if (std::isnan(largest) || gs[i] > largest) {
                    if (!std::isnan(gs[i])) {
                        largest = gs[i];
                    }
                }
            }
            if (largest > 0.0) {
                for (Index i = 0; i < m; ++i) {
                    gs[i] = gs[i] / largest;
                }
            }
        }
        for (Index i = 0; i < m; ++i) {
            columns(i, g) = gs[i];
        }
    }
    for (Index i = 0; i < m; ++i) {
        for (Index t = 0; t < ci; ++t) {
            columns(i, ng + t) = kernel.isotropic(i, t);
        }
    }
    const double tol = has_tolerance ? tolerance : default_gradient_tolerance(columns);
    return solve_nnls_with_limit(columns, signal, tol, 3 * columns.cols, trace).solution;
}

// Damped Richardson-Lucy of one voxel (drl.cpp:39-99; the block's row moves do not change a
// voxel's arithmetic).  kernel_t is K^T (c, m) and gram K^T K (c, c), as FodSolver.drl takes
// them.  Sums: the mean and the sum of squares of the sample standard deviation, K^T s, each
// f G and the FOD change are sequential in index order.
constexpr double DRL_FOD_FLOOR = 1e-6;           // drl.cpp:13-14
constexpr double DRL_DENOMINATOR_OFFSET = 1e-16; // drl.cpp:16-17
constexpr double DRL_CONVERGENCE_SHARE = 1e-4;   // drl.cpp:19-21

struct DrlSolver {
    Mat kernel_t; // (c, m)
    Mat gram;     // (c, c)
    double nu = 0.0;
    Index n_iter = 0;
    double init = 0.0;
};

double power_8(double x) { // common/maths.hpp:26-30
    const double x2 = x * x;
    const double x4 = x2 * x2;
    return x4 * x4;
}

// Returns the FOD; *iterations (optional) receives the iterations run, and *converged whether the
// change test stopped them.
Vec drl_voxel(const DrlSolver& d, const Vec& signal, Index* iterations, bool* converged) {
    const Index m = static_cast<Index>(signal.size());
    const Index c = d.gram.rows;
    // drl.cpp:23-28, 49-56: damping strength from the sample standard deviation, threshold from
    // the largest (non-NaN) signal value.
    double total = 0.0;
    for (Index i = 0; i < m; ++i) total += signal[i];
    const double mean = total / static_cast<double>(m);
    double squares = 0.0;
    for (Index i = 0; i < m; ++i) squares += (signal[i] - mean) * (signal[i] - mean);
    const double sd = std::sqrt(squares / static_cast<double>(m - 1));
    const double strength = std::fmax(1.0 - 4.0 * sd, 0.0);
    double largest = std::numeric_limits<double>::quiet_NaN();
    for (Index i = 0; i < m; ++i) {
        if (!std::isnan(signal[i]) && (std::isnan(largest) || signal[i] > largest)) {
            largest = signal[i];
        }
    }
    const double threshold = DRL_CONVERGENCE_SHARE * largest;
    const double nu8 = power_8(d.nu);
    // K^T s and f K^T K: each entry accumulated in index order with fused multiply-adds (one
    // rounding per term; the restatement's choice -- Eigen's products in the reference block
    // their sums their own way).
    Vec ks(static_cast<size_t>(c), 0.0);
    for (Index j = 0; j < c; ++j) {
        double acc = 0.0;
        for (Index i = 0; i < m; ++i) acc = std::fma(signal[i], d.kernel_t(j, i), acc);
        ks[j] = acc;
    }
    Vec f(static_cast<size_t>(c), d.init);
    Vec previous = f;
    Vec predicted(static_cast<size_t>(c));
    Index it = 0;
    bool done = false;
    for (; it < d.n_iter && !done; ++it) {
        for (Index j = 0; j < c; ++j) {
            double acc = 0.0;
            for (Index k = 0; k < c; ++k) acc = std::fma(f[k], d.gram(k, j), acc);
            predicted[j] = acc;
        }
        double change = 0.0;
        for (Index j = 0; j < c; ++j) {
            const double a = f[j];
            const double a8 = power_8(a);
            const double damping = 1.0 - strength * (1.0 - a8 / (a8 + nu8));
            const double ratio =
                (ks[j] - predicted[j]) / (predicted[j] + DRL_DENOMINATOR_OFFSET);
            const double updated = std::fmax(a * (1.0 + damping * ratio), DRL_FOD_FLOOR);
            f[j] = updated;
            change += std::fabs(updated - previous[j]);
        }
        if (change <= threshold) {
            done = true;
        } else {
            previous = f;
        }
    }
    if (iterations) *iterations = it;
    if (converged) *converged = done;
    return f;
}

// ---- Elastic net by coordinate descent (cpp/src/sd/elastic_net.cpp) ----------------------------
// The standardised design (elastic_net.cpp:25-37): each column divided by its root mean square
// sqrt(sum_r k(r, j)^2 / m), the sum sequential over rows; a column of zeros keeps scale 1 and
// stays out of the fit.  The sums over rows inside the descent (the correlations) are wave sums.
struct ElasticNet {
    Mat columns; // (m, c) scaled
    Vec scales;  // (c)
    std::vector<char> nonzero;
    double l1_threshold = 0.0; // penalty * l1_ratio
    double shrink = 1.0;       // 1 + penalty (1 - l1_ratio)
    double tol = 0.0;
    Index max_sweeps = 0;
};

ElasticNet make_elastic_net(const Mat& k, double penalty, double l1_ratio, double tol,
                            Index max_sweeps) {
    ElasticNet e;
    const Index m = k.rows;
    const Index c = k.cols;
    e.columns = k;
    e.scales.assign(static_cast<size_t>(c), 1.0);
    e.nonzero.assign(static_cast<size_t>(c), 0);
    for (Index j = 0; j < c; ++j) {
        double squares = 0.0;
        for (Index r = 0; r < m; ++r) {
            squares += k(r, j) * k(r, j);
        }
        const double scale = std::sqrt(squares / static_cast<double>(m));
        e.nonzero[j] = scale > 0.0 ? 1 : 0;
        e.scales[j] = e.nonzero[j] ? scale : 1.0;
        for (Index r = 0; r < m; ++r) {
            e.columns(r, j) = k(r, j) / e.scales[j];
        }
    }
    e.l1_threshold = penalty * l1_ratio;
    e.shrink = 1.0 + penalty * (1.0 - l1_ratio);
    e.tol = tol;
    e.max_sweeps = max_sweeps;
    return e;
}

// elastic_net.cpp:13-21.
double soft_threshold(double value, double threshold) {
    if (value > threshold) {
        return value - threshold;
    }
    if (value < -threshold) {
        return value + threshold;
    }
    return 0.0;
}

// solve_elastic_net (elastic_net.cpp:128-151) with the CoordinateDescent steps (:40-124).
// *sweeps (optional) receives the sweeps run.
Vec elastic_net_voxel(const ElasticNet& e, const Vec& data, bool* converged, Index* sweeps_out) {
    const Index m = e.columns.rows;
    const Index c = e.columns.cols;
    const double row_count = static_cast<double>(m);
    Vec residual = data;
    Vec coef(static_cast<size_t>(c), 0.0);
    auto correlation = [&](Index j) {
        return wave_sum(0, m, [&](Index r) { return e.columns(r, j) * residual[r]; }) / row_count;
    };
    // has_nonzero_solution (:51-57): the largest |correlation| of a non-zero column.
    double largest = 0.0;
    for (Index j = 0; j < c; ++j) {
        const double a = e.nonzero[j] ? std::fabs(correlation(j)) : 0.0;
        largest = std::max(largest, a);
    }
    if (!(largest > e.l1_threshold)) {
        if (converged) *converged = true;
        if (sweeps_out) *sweeps_out = 0;
        return coef; // unscaled zeros
    }
    // update (:104-116): one coefficient to its minimiser; its change relative to 1 + |old|.
    auto update = [&](Index j) {
        const double old_value = coef[j];
        const double corr = correlation(j) + old_value;
        const double new_value = soft_threshold(corr, e.l1_threshold) / e.shrink;
        if (new_value == old_value) {
            return 0.0;
        }
        const double delta = new_value - old_value;
        for (Index r = 0; r < m; ++r) {
            residual[r] = residual[r] - delta * e.columns(r, j);
        }
        coef[j] = new_value;
        return std::fabs(delta) / (1.0 + std::fabs(old_value));
    };
    Index sweeps = 0;
    bool done = false;
    std::vector<Index> active;
    while (sweeps < e.max_sweeps) {
        ++sweeps;
        double change = 0.0;
        for (Index j = 0; j < c; ++j) {
            if (e.nonzero[j]) {
                change = std::max(change, update(j));
            }
        }
        if (change < e.tol) {
            done = true;
            break;
        }
        active.clear();
        for (Index j = 0; j < c; ++j) {
            if (coef[j] != 0.0) {
                active.push_back(j);
            }
        }
        while (sweeps < e.max_sweeps) {
            ++sweeps;
            double inner = 0.0;
            for (Index j : active) {
                inner = std::max(inner, update(j));
            }
            if (inner < e.tol) {
                break;
            }
        }
    }
    for (Index j = 0; j < c; ++j) {
        coef[j] = coef[j] / e.scales[j];
    }
    if (converged) *converged = done;
    if (sweeps_out) *sweeps_out = sweeps;
    return coef;
}

// One stage's FOD solver: ridge NNLS (L2NnlsSolver) or damped Richardson-Lucy (DrlSolver).
struct Solver {
    RidgeDesign design;
    double tol = 0.0;
    bool drl = false;
    DrlSolver rl;
    bool lasso = false; // elastic net (ElasticNetSolver) instead
    ElasticNet net;
};

// The FOD of one signal by a stage's solver (solver.cpp:9-37).
Vec stage_fod(const Solver& solver, const Vec& a, NnlsTrace* tr) {
    if (solver.drl) {
        return drl_voxel(solver.rl, a, nullptr, nullptr);
    }
    if (solver.lasso) {
        return elastic_net_voxel(solver.net, a, nullptr, nullptr);
    }
    return solve_ridge_nnls(solver.design, a, 0.0, solver.tol, tr);
}

struct MtTrace {
    Index rounds = 0;
    NnlsTrace lr_sum;      // summed over rounds
    Index lr_max_active = 0;
    Index lr_qr_fallbacks = 0;
    NnlsTrace hr;
    Index frac_inner = 0;
};

void fit_multi_tissue_block(const Mat& signals, const Vec& weights, const TissueKernel& lr,
                            const Solver& lr_solver, const TissueKernel& hr,
                            const Solver& hr_solver, Index max_rounds, double fraction_tol,
                            double final_tol, double* fods_out, double* fractions_out,
                            MtTrace* traces) {
    const Index nv = signals.rows;
    const Index m = signals.cols;
    const Index ci = lr.isotropic.cols;
    const Index nlf = 1 + ci;
    Mat weighted(nv, m);
    for (Index v = 0; v < nv; ++v) {
        for (Index i = 0; i < m; ++i) {
            weighted(v, i) = signals(v, i) * weights[i];
        }
    }
    Mat iso(nv, ci, 0.0);
    Mat previous(nv, nlf, std::numeric_limits<double>::infinity());
    std::vector<Index> looping;
    for (Index v = 0; v < nv; ++v) {
        looping.push_back(v);
    }
    std::vector<Index> still;
    auto aniso_signal = [&](Index v, const Mat& w_iso) {
        Vec a(static_cast<size_t>(m));
        for (Index i = 0; i < m; ++i) {
            double s = 0.0;
            for (Index t = 0; t < ci; ++t) {
                s += w_iso(i, t) * iso(v, t);
            }
            a[i] = std::fmax(weighted(v, i) - s, 0.0);
        }
        return a;
    };
    auto row = [&](const Mat& mat, Index v) {
        return Vec(mat.v.begin() + v * mat.cols, mat.v.begin() + (v + 1) * mat.cols);
    };
    for (Index round = 0; round < max_rounds && !looping.empty(); ++round) {
        still.clear();
        for (Index v : looping) {
            const Vec a = aniso_signal(v, lr.weighted_isotropic);
            NnlsTrace tr;
            const Vec fod = stage_fod(lr_solver, a, traces ? &tr : nullptr);
            NnlsTrace ftr;
            const Vec fr = compute_fractions(fod, lr, row(signals, v), true, false, 0.0,
                                             traces ? &ftr : nullptr);
            if (traces) {
                MtTrace& t = traces[v];
                t.rounds += 1;
                t.lr_sum.outer += tr.outer;
                t.lr_sum.inner += tr.inner;
                t.lr_sum.max_active += tr.final_active;
                t.lr_max_active = std::max(t.lr_max_active, tr.max_active);
                t.lr_qr_fallbacks += tr.qr_fallback;
                t.frac_inner += ftr.inner;
            }
            for (Index t = 0; t < ci; ++t) {
                iso(v, t) = fr[1 + t];
            }
            bool converged = true;
            for (Index f = 0; f < nlf; ++f) {
                if (!(std::fabs(fr[f] - previous(v, f)) < fraction_tol)) {
                    converged = false;
                }
            }
            if (!converged) {
                for (Index f = 0; f < nlf; ++f) {
                    previous(v, f) = fr[f];
                }
                still.push_back(v);
            }
        }
        std::swap(looping, still);
    }
    const Index ca = hr.anisotropic.cols;
    const Index nf = static_cast<Index>(hr.blocks.size()) + ci;
    for (Index v = 0; v < nv; ++v) {
        const Vec a = aniso_signal(v, hr.weighted_isotropic);
        NnlsTrace tr;
        const Vec fod = stage_fod(hr_solver, a, traces ? &tr : nullptr);
        if (traces) {
            traces[v].hr = tr;
        }
        for (Index c = 0; c < ca; ++c) {
            fods_out[v * ca + c] = fod[c];
        }
    }
    for (Index v = 0; v < nv; ++v) {
        const Vec fod(fods_out + v * ca, fods_out + (v + 1) * ca);
        const Vec fr = compute_fractions(fod, hr, row(signals, v), false, true, final_tol, nullptr);
        double total = 0.0;
        for (double f : fr) {
            total += f;
        }
        total += FRACTION_SUM_OFFSET;
        for (Index f = 0; f < nf; ++f) {
            fractions_out[v * nf + f] = fr[f] / total;
        }
    }
}

// parallel_for_blocks (parallel.hpp): blocks of block_size items taken from a shared counter;
// the first exception is rethrown after all threads finish.
template <class Body>
void parallel_for_blocks(Index n_items, Index block_size, int n_threads, Body body) {
    const Index n_blocks = (n_items + block_size - 1) / block_size;
    if (n_blocks == 0) {
        return;
    }
    std::atomic<Index> next{0};
    std::atomic<bool> failed{false};
    std::exception_ptr first;
    std::mutex mu;
    auto run = [&] {
        try {
            while (!failed.load(std::memory_order_relaxed)) {
                const Index b = next.fetch_add(1);
                if (b >= n_blocks) {
                    break;
                }
                const Index begin = b * block_size;
                const Index end = std::min(begin + block_size, n_items);
                body(begin, end);
            }
        } catch (...) {
            std::lock_guard<std::mutex> lock(mu);
            if (!first) {
                first = std::current_exception();
            }
            failed.store(true);
        }
    };
    const Index workers = std::min<Index>(n_threads, n_blocks);
    std::vector<std::thread> threads;
    for (Index w = 1; w < workers; ++w) {
        threads.emplace_back(run);
    }
    run();
    for (auto& t : threads) {
        t.join();
    }
    if (first) {
        std::rethrow_exception(first);
    }
}

// ---------------------------------------------------------------------------------------------
// Constrained spherical deconvolution (cpp/src/sd/csd.cpp).

struct CsdProblem {
    Mat initial;    // (c0, m)
    Mat forward;    // (m, c)
    Mat constraint; // (k, c), weighted
    Mat gram;       // (c, c)
    double threshold = 0.0;
    Index max_iterations = 0;
};

// make_csd_problem (csd.cpp:11-21): gram = F^T F, each entry summed over the measurements in
// order, then norm_weight added to the diagonal.
CsdProblem make_csd_problem(Mat initial, Mat forward, Mat constraint, double norm_weight,
                            double threshold, Index max_iterations) {
    const Index c = forward.cols;
    Mat gram(c, c);
    for (Index i = 0; i < c; ++i) {
        for (Index j = 0; j < c; ++j) {
            double sum = 0.0;
            for (Index r = 0; r < forward.rows; ++r) {
                sum = sum + forward(r, i) * forward(r, j);
            }
            gram(i, j) = sum;
        }
    }
    for (Index i = 0; i < c; ++i) {
        gram(i, i) = gram(i, i) + norm_weight;
    }
    CsdProblem p;
    p.initial = std::move(initial);
    p.forward = std::move(forward);
    p.constraint = std::move(constraint);
    p.gram = std::move(gram);
    p.threshold = threshold;
    p.max_iterations = max_iterations;
    return p;
}

// The Cholesky factor of a symmetric matrix (lower triangle used), left-looking:
// L_ij = (A_ij - L_i0 L_j0 - L_i1 L_j1 - ... - L_i(j-1) L_j(j-1)) / L_jj, subtracted in order,
// and L_jj = sqrt(A_jj - L_j0^2 - ...).  Eigen's LLT (csd.cpp:34,60) factors the same matrix.
Mat cholesky_lower(const Mat& a) {
    const Index n = a.rows;
    Mat l(n, n);
    for (Index j = 0; j < n; ++j) {
        double d = a(j, j);
        for (Index p = 0; p < j; ++p) {
            d = d - l(j, p) * l(j, p);
        }
        const double ljj = std::sqrt(d);
        l(j, j) = ljj;
        for (Index i = j + 1; i < n; ++i) {
            double v = a(i, j);
            for (Index p = 0; p < j; ++p) {
                v = v - l(i, p) * l(j, p);
            }
            l(i, j) = v / ljj;
        }
    }
    return l;
}

// L L^T x = b: forward then backward substitution; a zero right-hand entry stays zero (as
// Eigen's triangular solves skip it), the terms subtracted in the order the column-oriented
// solves meet them (forward: earlier positions first; backward: later positions first).
Vec cholesky_solve(const Mat& l, const Vec& b) {
    const Index n = l.rows;
    Vec y(b);
    for (Index i = 0; i < n; ++i) {
        double v = y[i];
        for (Index p = 0; p < i; ++p) {
            v = v - l(i, p) * y[p];
        }
        y[i] = v != 0.0 ? v / l(i, i) : v;
    }
    Vec x(y);
    for (Index i = n - 1; i >= 0; --i) {
        double v = x[i];
        for (Index p = n - 1; p > i; --p) {
            v = v - l(p, i) * x[p];
        }
        x[i] = v != 0.0 ? v / l(i, i) : v;
    }
    return x;
}

struct CsdOut {
    Vec fod;
    bool converged = true;
    Index solves = 0;
};

// solve_csd (csd.cpp:23-65): the linear start initial * s (first c0 coefficients), then solves of
// (gram + sum over constrained directions d, in increasing d, of r_d r_d^T) f = F^T s until the
// set {d : (C f)_d < threshold} repeats or max_iterations solves have run.  Sums over the
// measurements and over the coefficients run in index order.
CsdOut solve_csd(const CsdProblem& p, const Vec& s) {
    const Index m = p.forward.rows;
    const Index c = p.forward.cols;
    const Index k = p.constraint.rows;
    CsdOut out;
    out.fod.assign(static_cast<size_t>(c), 0.0);
    for (Index i = 0; i < p.initial.rows; ++i) {
        double sum = 0.0;
        for (Index r = 0; r < m; ++r) {
            sum = sum + p.initial(i, r) * s[r];
        }
        out.fod[i] = sum;
    }
    Vec projected(static_cast<size_t>(c));
    for (Index i = 0; i < c; ++i) {
        double sum = 0.0;
        for (Index r = 0; r < m; ++r) {
            sum = sum + p.forward(r, i) * s[r];
        }
        projected[i] = sum;
    }
    std::vector<Index> constrained, previous;
    bool first = true;
    out.converged = p.max_iterations == 0;
    for (Index it = 0; it < p.max_iterations; ++it) {
        constrained.clear();
        for (Index d = 0; d < k; ++d) {
            double a = 0.0;
            for (Index j = 0; j < c; ++j) {
                a = a + p.constraint(d, j) * out.fod[j];
            }
            if (a < p.threshold) {
                constrained.push_back(d);
            }
        }
        if (!first && constrained == previous) {
            out.converged = true;
            break;
        }
        first = false;
        Mat system = p.gram;
        for (Index d : constrained) {
            for (Index i = 0; i < c; ++i) {
                for (Index j = 0; j < c; ++j) {
                    system(i, j) = system(i, j) + p.constraint(d, i) * p.constraint(d, j);
                }
            }
        }
        out.fod = cholesky_solve(cholesky_lower(system), projected);
        out.solves += 1;
        std::swap(previous, constrained);
    }
    return out;
}

// ---------------------------------------------------------------------------------------------
// Peaks of SH series (cpp/src/sh/harmonics.cpp, cpp/src/sh/peaks.cpp).

using namespace oracle_sh;

// The peaks of one voxel (peaks.cpp voxel_peaks): a climb from every vertex whose positive
// amplitude no neighbour exceeds, in vertex order; a new peak within 0.99 in cosine (|cos| unless
// full) of a kept one replaces it when larger, else is dropped; then the kept peaks by amplitude,
// stable, largest first, cut at threshold and relative_threshold * largest; symmetric peaks are
// turned to z >= 0.  Unused slots are NaN, all of them when on_sphere[0] is not finite.
void sh_voxel_peaks(const double* c, Index lmax, bool full, const double* on_sphere, Index nv,
                    const double* vertices, const std::int64_t* neighbours, Index n_peaks,
                    double relative_threshold, double threshold, double* dirs, double* amps) {
    const double nan = std::numeric_limits<double>::quiet_NaN();
    for (Index i = 0; i < 3 * n_peaks; ++i) dirs[i] = nan;
    for (Index i = 0; i < n_peaks; ++i) amps[i] = nan;
    if (nv == 0 || !std::isfinite(on_sphere[0])) {
        return;
    }
    std::vector<ShPeak> found;
    for (Index v = 0; v < nv; ++v) {
        const double here = on_sphere[v];
        bool local_max = here > 0.0;
        for (int q = 0; q < 6 && local_max; ++q) {
            local_max = on_sphere[neighbours[v * 6 + q]] <= here;
        }
        if (!local_max) {
            continue;
        }
        ShPeak peak;
        if (!sh_climb(c, lmax, full, vertices + 3 * v, 50, 1e-4, 0.2, peak)) {
            continue;
        }
        if (!(peak.amplitude > 0.0)) {
            continue;
        }
        ShPeak* match = nullptr;
        for (ShPeak& other : found) {
            const double cosine = peak.dir[0] * other.dir[0] + peak.dir[1] * other.dir[1] +
                                  peak.dir[2] * other.dir[2];
            if ((full ? cosine : std::fabs(cosine)) > 0.99) {
                match = &other;
                break;
            }
        }
        if (match == nullptr) {
            found.push_back(peak);
        } else if (peak.amplitude > match->amplitude) {
            *match = peak;
        }
    }
    if (found.empty()) {
        return;
    }
    std::stable_sort(found.begin(), found.end(),
                     [](const ShPeak& a, const ShPeak& b) { return a.amplitude > b.amplitude; });
    const double largest = found.front().amplitude;
    const Index kept = std::min<Index>(n_peaks, static_cast<Index>(found.size()));
    for (Index r = 0; r < kept; ++r) {
        const ShPeak& p = found[static_cast<size_t>(r)];
        if (p.amplitude < threshold || p.amplitude < relative_threshold * largest) {
            break;
        }
        const bool flip = !full && p.dir[2] < 0.0;
        for (int a = 0; a < 3; ++a) {
            dirs[3 * r + a] = flip ? -p.dir[a] : p.dir[a];
        }
        amps[r] = p.amplitude;
    }
}

// ---------------------------------------------------------------------------------------------
// Least squares with inequality constraints (cpp/src/sd/lsi.cpp).

struct LsiDesign {
    Index m = 0, n = 0, k = 0;
    Mat q_transposed;      // (n, m)
    Mat back_substitution; // (n, n): P R^-1
    Mat constraints;       // (k, n): E = G P R^-1, unit rows
    Mat constraint_gram;   // (k, k): E E^T
    Vec constraint_abs_sums;
};

// make_lsi_design (lsi.cpp:19-69): the pivoted QR of [A; ridge I] (the QR above, wave-sum order),
// the thin Q's measurement rows transposed, P R^-1 (R^-1 by column-oriented back substitution of
// the identity), E = G P R^-1 with unit rows, E E^T and the rows' absolute sums.  Every other sum
// runs in index order.
LsiDesign make_lsi_design(const Mat& a, const Mat& g, double ridge) {
    const Index m = a.rows;
    const Index n = a.cols;
    if (g.cols != n) {
        throw std::invalid_argument("the constraints have " + std::to_string(g.cols) +
                                    " columns but the design has " + std::to_string(n));
    }
    if (!(ridge >= 0.0)) {
        throw std::invalid_argument("the ridge must be at least 0");
    }
    const bool has_ridge = ridge > 0.0;
    const Index rows = m + (has_ridge ? n : 0);
    Mat stacked(rows, n);
    for (Index r = 0; r < m; ++r) {
        for (Index c = 0; c < n; ++c) {
            stacked(r, c) = a(r, c);
        }
    }
    if (has_ridge) {
        for (Index c = 0; c < n; ++c) {
            stacked(m + c, c) = ridge;
        }
    }
    PivotedQR qr;
    qr.compute(stacked);
    const double threshold = static_cast<double>(std::max(rows, n)) * EPS;
    if (rows < n || qr.rank(threshold) < n) {
        throw std::invalid_argument(
            "the design does not have full column rank: the measurements cannot determine all " +
            std::to_string(n) + " parameters");
    }
    LsiDesign d;
    d.m = m;
    d.n = n;
    d.k = g.rows;
    // Thin Q: Q e_j for j < n, the reflectors applied last to first.
    d.q_transposed = Mat(n, m);
    for (Index j = 0; j < n; ++j) {
        Vec col(static_cast<size_t>(rows), 0.0);
        col[j] = 1.0;
        for (Index h = n - 1; h >= 0; --h) {
            const double tau = qr.hcoeffs[h];
            if (rows - h == 1) {
                col[h] *= (1.0 - tau);
            } else if (tau != 0.0) {
                double tmp = wave_sum(h + 1, rows, [&](Index r) { return qr.qr(r, h) * col[r]; });
                tmp += col[h];
                col[h] -= tau * tmp;
                for (Index r = h + 1; r < rows; ++r) {
                    col[r] -= (tau * qr.qr(r, h)) * tmp;
                }
            }
        }
        for (Index r = 0; r < m; ++r) {
            d.q_transposed(j, r) = col[r];
        }
    }
    // R^-1, one identity column at a time, then P R^-1.
    Mat rinv(n, n);
    for (Index j = 0; j < n; ++j) {
        Vec x(static_cast<size_t>(n), 0.0);
        x[j] = 1.0;
        for (Index i = n - 1; i >= 0; --i) {
            if (x[i] != 0.0) {
                x[i] = x[i] / qr.qr(i, i);
            }
            for (Index r = 0; r < i; ++r) {
                x[r] -= qr.qr(r, i) * x[i];
            }
        }
        for (Index i = 0; i < n; ++i) {
            rinv(i, j) = x[i];
        }
    }
    d.back_substitution = Mat(n, n);
    for (Index p = 0; p < n; ++p) {
        for (Index j = 0; j < n; ++j) {
            d.back_substitution(qr.perm[p], j) = rinv(p, j);
        }
    }
    d.constraints = Mat(d.k, n);
    for (Index c = 0; c < d.k; ++c) {
        for (Index j = 0; j < n; ++j) {
            double sum = 0.0;
            for (Index i = 0; i < n; ++i) {
                sum = sum + g(c, i) * d.back_substitution(i, j);
            }
            d.constraints(c, j) = sum;
        }
        double sq = 0.0;
        for (Index j = 0; j < n; ++j) {
            sq = sq + d.constraints(c, j) * d.constraints(c, j);
        }
        const double norm = std::sqrt(sq);
        if (norm > 0.0) {
            for (Index j = 0; j < n; ++j) {
                d.constraints(c, j) = d.constraints(c, j) / norm;
            }
        }
    }
    d.constraint_gram = Mat(d.k, d.k);
    d.constraint_abs_sums.assign(static_cast<size_t>(d.k), 0.0);
    for (Index a2 = 0; a2 < d.k; ++a2) {
        for (Index b = 0; b < d.k; ++b) {
            double sum = 0.0;
            for (Index j = 0; j < n; ++j) {
                sum = sum + d.constraints(a2, j) * d.constraints(b, j);
            }
            d.constraint_gram(a2, b) = sum;
        }
        double abs_sum = 0.0;
        for (Index j = 0; j < n; ++j) {
            abs_sum = abs_sum + std::fabs(d.constraints(a2, j));
        }
        d.constraint_abs_sums[a2] = abs_sum;
    }
    return d;
}

// The growing Cholesky factor of H = E E^T + f f^T on the active columns in the order they
// joined (lsi.cpp GrowingCholesky): append by a forward solve (a zero right-hand entry stays
// zero) with the pivot test of nnls.cpp, removal by the Givens downdate, solves forward then
// backward, column-oriented.
class LdpCholesky {
public:
    LdpCholesky(const LsiDesign& d, const Vec& f) : design(d), bounds(f), factor(d.k, d.k) {}

    double h(Index row, Index column) const {
        return design.constraint_gram(row, column) + bounds[row] * bounds[column];
    }

    bool append(Index column) {
        const Index pos = static_cast<Index>(columns.size());
        Vec cross(static_cast<size_t>(pos));
        for (Index r = 0; r < pos; ++r) {
            cross[r] = h(columns[r], column);
        }
        forward(cross, pos);
        double sq = 0.0;
        for (Index r = 0; r < pos; ++r) {
            sq += cross[r] * cross[r];
        }
        const double hcc = h(column, column);
        const double pivot = hcc - sq;
        if (!(pivot > MIN_RELATIVE_PIVOT * hcc)) {
            return false;
        }
        for (Index r = 0; r < pos; ++r) {
            factor(pos, r) = cross[r];
        }
        factor(pos, pos) = std::sqrt(pivot);
        columns.push_back(column);
        return true;
    }

    void keep(const std::vector<char>& active) {
        for (Index pos = static_cast<Index>(columns.size()) - 1; pos >= 0; --pos) {
            if (!active[columns[pos]]) {
                remove_cholesky_column(factor, static_cast<Index>(columns.size()), pos);
                columns.erase(columns.begin() + pos);
            }
        }
    }

    Vec solve(const Vec& rhs) const {
        const Index k = static_cast<Index>(columns.size());
        Vec w(static_cast<size_t>(k));
        for (Index p = 0; p < k; ++p) {
            w[p] = rhs[columns[p]];
        }
        forward(w, k);
        for (Index j = k - 1; j >= 0; --j) {
            if (w[j] != 0.0) {
                w[j] = w[j] / factor(j, j);
            }
            for (Index i = 0; i < j; ++i) {
                w[i] -= factor(j, i) * w[j];
            }
        }
        Vec full(rhs.size(), 0.0);
        for (Index p = 0; p < k; ++p) {
            full[columns[p]] = w[p];
        }
        return full;
    }

    Index size() const { return static_cast<Index>(columns.size()); }

private:
    void forward(Vec& w, Index k) const {
        for (Index j = 0; j < k; ++j) {
            if (w[j] != 0.0) {
                w[j] = w[j] / factor(j, j);
            }
            for (Index i = j + 1; i < k; ++i) {
                w[i] -= factor(i, j) * w[j];
            }
        }
    }

    const LsiDesign& design;
    const Vec& bounds;
    Mat factor;
    std::vector<Index> columns;
};

// Entering steps after which an LDP solve counts as failed (handed to the QR path): each step
// adds a column, and a finite solve takes far fewer.
Index ldp_outer_guard(Index k) { return 8 * k + 64; }

struct LsiTrace {
    int fallback = 0; // 1 when the QR-based NNLS finished the voxel
    Index outer = 0;
    Index inner = 0;
    Index max_factor = 0;
};

// ldp_multipliers (lsi.cpp:191-240): Lawson-Hanson on H, the gradient f (1 - f.u) minus
// u_c (E E^T)[:, c] over the non-zero u_c in column order (f.u over the non-zero u_c in order).
// Returns false when a pivot is refused, the inner steps run out or the outer guard is hit.
bool ldp_multipliers(const LsiDesign& d, const Vec& f, Vec& u, LsiTrace* tr) {
    const Index k = d.k;
    double norm1 = 0.0;
    for (Index c = 0; c < k; ++c) {
        const double v = d.constraint_abs_sums[c] + std::fabs(f[c]);
        if (c == 0 || v > norm1) {
            norm1 = v;
        }
    }
    const double tol = 10.0 * EPS * norm1 * static_cast<double>(std::max(d.n + 1, k));
    const Index max_inner = 3 * k;
    std::vector<char> active(static_cast<size_t>(k), 0);
    LdpCholesky chol(d, f);
    u.assign(static_cast<size_t>(k), 0.0);
    Index n_inner = 0;
    Index n_outer = 0;
    while (true) {
        double dot = 0.0;
        for (Index c = 0; c < k; ++c) {
            if (u[c] != 0.0) {
                dot = dot + f[c] * u[c];
            }
        }
        Vec g(static_cast<size_t>(k));
        for (Index j = 0; j < k; ++j) {
            g[j] = f[j] * (1.0 - dot);
        }
        for (Index c = 0; c < k; ++c) {
            if (u[c] != 0.0) {
                for (Index j = 0; j < k; ++j) {
                    g[j] = g[j] - u[c] * d.constraint_gram(c, j);
                }
            }
        }
        Index entering = -1;
        for (Index c = 0; c < k; ++c) {
            if (!active[c] && (entering < 0 || g[c] > g[entering])) {
                entering = c;
            }
        }
        if (entering < 0 || g[entering] <= tol) {
            return true;
        }
        if (++n_outer > ldp_outer_guard(k)) {
            return false;
        }
        if (tr) tr->outer += 1;
        active[entering] = 1;
        if (!chol.append(entering)) {
            return false;
        }
        if (tr) tr->max_factor = std::max(tr->max_factor, chol.size());
        Vec cand = chol.solve(f);
        while (true) {
            double step = std::numeric_limits<double>::infinity();
            Index blocking = -1;
            for (Index c = 0; c < k; ++c) {
                if (!active[c] || cand[c] > 0.0) {
                    continue;
                }
                const double distance = u[c] - cand[c];
                const double s = distance > 0.0 ? u[c] / distance : 0.0;
                if (s < step) {
                    step = s;
                    blocking = c;
                }
            }
            if (blocking < 0) {
                break;
            }
            if (++n_inner > max_inner) {
                return false;
            }
            if (tr) tr->inner += 1;
            for (Index c = 0; c < k; ++c) {
                u[c] = u[c] + step * (cand[c] - u[c]);
            }
            u[blocking] = 0.0;
            for (Index c = 0; c < k; ++c) {
                if (c == blocking || std::fabs(u[c]) < tol) {
                    active[c] = 0;
                    u[c] = 0.0;
                }
            }
            chol.keep(active);
            cand = chol.solve(f);
        }
        u = cand;
    }
}

struct LsiOut {
    Vec x;
    bool converged = true;
};

// solve_lsi (lsi.cpp:245-292): r = Q^T s; f = -(E r) / max|E r| (all sums in index order); the
// LDP multipliers u by the Cholesky NNLS or, when that fails, by solve_nnls_with_limit (its
// gradient and tolerance summing each column over the rows in order) on
// [E^T; f^T] with target e_n; x = P R^-1 (scale (E^T u) / (1 - f.u) + r), the sums over the
// non-zero u_c in order.
LsiOut solve_lsi(const LsiDesign& d, const Vec& s, LsiTrace* tr, bool only_qr = false) {
    const Index n = d.n;
    const Index k = d.k;
    Vec rotated(static_cast<size_t>(n));
    for (Index i = 0; i < n; ++i) {
        double sum = 0.0;
        for (Index r = 0; r < d.m; ++r) {
            sum = sum + d.q_transposed(i, r) * s[r];
        }
        rotated[i] = sum;
    }
    // scale = max |raw_c|, NaN when any raw_c is NaN.
    Vec raw(static_cast<size_t>(k));
    double scale = 0.0;
    bool any_nan = false;
    for (Index c = 0; c < k; ++c) {
        double sum = 0.0;
        for (Index i = 0; i < n; ++i) {
            sum = sum + d.constraints(c, i) * rotated[i];
        }
        raw[c] = -sum;
        const double a = std::fabs(raw[c]);
        any_nan = any_nan || std::isnan(a);
        scale = a > scale ? a : scale;
    }
    if (any_nan) {
        scale = std::numeric_limits<double>::quiet_NaN();
    }
    LsiOut out;
    auto back = [&](const Vec& z) {
        Vec x(static_cast<size_t>(n));
        for (Index i = 0; i < n; ++i) {
            double sum = 0.0;
            for (Index j = 0; j < n; ++j) {
                sum = sum + d.back_substitution(i, j) * z[j];
            }
            x[i] = sum;
        }
        return x;
    };
    if (!(scale > 0.0)) {
        out.x = back(rotated);
        return out;
    }
    Vec f(static_cast<size_t>(k));
    for (Index c = 0; c < k; ++c) {
        f[c] = raw[c] / scale;
    }
    Vec u;
    if (only_qr || !ldp_multipliers(d, f, u, tr)) {
        if (tr) tr->fallback = 1;
        Mat c(n + 1, k);
        for (Index j = 0; j < k; ++j) {
            for (Index i = 0; i < n; ++i) {
                c(i, j) = d.constraints(j, i);
            }
            c(n, j) = f[j];
        }
        Vec target(static_cast<size_t>(n + 1), 0.0);
        target[n] = 1.0;
        NnlsResult r = solve_nnls_with_limit(
            c, target, default_gradient_tolerance(c, ColumnSum::InOrder), 3 * k, nullptr,
            ColumnSum::InOrder);
        u = std::move(r.solution);
        out.converged = r.converged;
    }
    double dot = 0.0;
    for (Index c = 0; c < k; ++c) {
        if (u[c] != 0.0) {
            dot = dot + f[c] * u[c];
        }
    }
    const double denom = 1.0 - dot;
    Vec z(static_cast<size_t>(n));
    for (Index j = 0; j < n; ++j) {
        double sum = 0.0;
        for (Index c = 0; c < k; ++c) {
            if (u[c] != 0.0) {
                sum = sum + d.constraints(c, j) * u[c];
            }
        }
        z[j] = (scale * sum) / denom + rotated[j];
    }
    out.x = back(z);
    return out;
}

TissueKernel make_kernel(const double* k_aniso, const double* k_iso, const double* w_k_iso,
                         Index m, Index ca, Index ci, const std::int64_t* blocks,
                         Index n_blocks) {
    TissueKernel k;
    k.anisotropic = from_ptr(k_aniso, m, ca);
    k.isotropic = from_ptr(k_iso, m, ci);
    k.weighted_isotropic = from_ptr(w_k_iso, m, ci);
    for (Index b = 0; b < n_blocks; ++b) {
        k.blocks.push_back({blocks[2 * b], blocks[2 * b + 1]});
    }
    return k;
}

template <class F>
int guarded(F f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return 1;
    }
}

// ---------------------------------------------------------------------------------------------
// Log-signal fit (log_fit.cpp:28-62 reweighted_fit, 159-185 fit_log_signal_block), without
// constraints, on a prepared design: the unit-column design transposed xt (p, n), its
// pseudo-inverse (p, n) and the column norms.  The weighted linear least squares of Veraart et
// al. (NeuroImage 81, 2013): the ordinary fit, then refits weighted by the predicted signal
// squared, until no predicted log signal moves by more than tol.
// ---------------------------------------------------------------------------------------------

// e^x for x <= 0: Cody-Waite reduction x = k ln2 + r (ln2 in a 32-bit head and its remainder)
// and the degree-13 Taylor series of e^r by Horner's rule, scaled by 2^k.
double exp_nonpositive(double x) {
    if (!(x > -746.0)) {
        return x != x ? x : 0.0;
    }
    const double k = std::rint(x * 0x1.71547652b82fep+0);
    const double r = (x - k * 0x1.62e42fee00000p-1) - k * 0x1.a39ef35793c76p-33;
    static const double inverse_factorials[] = {
        0x1.6124613a86d09p-33, 0x1.1eed8eff8d898p-29, 0x1.ae64567f544e4p-26,
        0x1.27e4fb7789f5cp-22, 0x1.71de3a556c734p-19, 0x1.a01a01a01a01ap-16,
        0x1.a01a01a01a01ap-13, 0x1.6c16c16c16c17p-10, 0x1.1111111111111p-7,
        0x1.5555555555555p-5,  0x1.5555555555555p-3,  0.5, 1.0, 1.0};
    double e = inverse_factorials[0];
    for (int i = 1; i < 14; ++i) {
        e = e * r + inverse_factorials[i];
    }
    return std::ldexp(e, static_cast<int>(k));
}

// Solves A q = b by LU with partial pivoting (the first row of largest magnitude as pivot);
// false when a pivot is 0.  aug is (p, p + 1) row-major, [A | b].
bool lu_solve_augmented(std::vector<double>& aug, Index p, Vec& q) {
    auto at = [&](Index i, Index j) -> double& { return aug[static_cast<size_t>(i * (p + 1) + j)]; };
    for (Index k = 0; k < p; ++k) {
        Index best = k;
        double best_abs = std::fabs(at(k, k));
        for (Index i = k + 1; i < p; ++i) {
            if (std::fabs(at(i, k)) > best_abs) {
                best = i;
                best_abs = std::fabs(at(i, k));
            }
        }
        const double pivot = at(best, k);
        if (pivot == 0.0) {
            return false;
        }
        if (best != k) {
            for (Index j = 0; j <= p; ++j) {
                std::swap(at(k, j), at(best, j));
            }
        }
        for (Index i = k + 1; i < p; ++i) {
            const double lik = at(i, k) / pivot;
            for (Index j = k + 1; j <= p; ++j) {
                at(i, j) = at(i, j) - lik * at(k, j);
            }
        }
    }
    q.assign(static_cast<size_t>(p), 0.0);
    for (Index k = p - 1; k >= 0; --k) {
        double s = at(k, p);
        for (Index j = k + 1; j < p; ++j) {
            s = s - at(k, j) * q[j];
        }
        q[k] = s / at(k, k);
    }
    return true;
}

bool all_finite(const Vec& v) {
    for (double x : v) {
        if (!std::isfinite(x)) {
            return false;
        }
    }
    return true;
}

// One voxel; returns the refits kept (-1 when a refit failed and the ordinary fit stands).
int fit_log_signal_voxel(const double* xt, const double* pinv, const double* scales, Index n,
                         Index p, const double* y, bool weighted, Index max_reweightings,
                         double tol, double* params) {
    Vec ordinary(static_cast<size_t>(p));
    for (Index a = 0; a < p; ++a) {
        ordinary[a] = wave_sum(0, n, [&](Index r) { return pinv[a * n + r] * y[r]; });
    }
    Vec cur = ordinary;
    int refits = 0;
    auto predict = [&](const Vec& q) {
        Vec out(static_cast<size_t>(n));
        for (Index r = 0; r < n; ++r) {
            double s = 0.0;
            for (Index a = 0; a < p; ++a) {
                s = s + xt[a * n + r] * q[a];
            }
            out[r] = s;
        }
        return out;
    };
    if (weighted && all_finite(ordinary)) {
        Vec pred = predict(cur);
        for (Index it = 0; it < max_reweightings; ++it) {
            double top = -std::numeric_limits<double>::infinity();
            for (Index r = 0; r < n; ++r) {
                top = pred[r] > top ? pred[r] : top;
            }
            Vec w(static_cast<size_t>(n));
            for (Index r = 0; r < n; ++r) {
                w[r] = exp_nonpositive(2.0 * (pred[r] - top));
            }
            std::vector<double> aug(static_cast<size_t>(p * (p + 1)));
            for (Index a = 0; a < p; ++a) {
                Vec wx(static_cast<size_t>(n));
                for (Index r = 0; r < n; ++r) {
                    wx[r] = w[r] * xt[a * n + r];
                }
                aug[a * (p + 1) + p] = wave_sum(0, n, [&](Index r) { return wx[r] * y[r]; });
                for (Index b = a; b < p; ++b) {
                    const double s = wave_sum(0, n, [&](Index r) { return wx[r] * xt[b * n + r]; });
                    aug[a * (p + 1) + b] = s;
                    aug[b * (p + 1) + a] = s;
                }
            }
            Vec next;
            if (!lu_solve_augmented(aug, p, next) || !all_finite(next)) {
                refits = -1;
                break;
            }
            const Vec next_pred = predict(next);
            double change = 0.0;
            for (Index r = 0; r < n; ++r) {
                const double d = std::fabs(next_pred[r] - pred[r]);
                change = d > change ? d : change;
            }
            cur = next;
            pred = next_pred;
            refits = static_cast<int>(it + 1);
            if (!(change > tol)) {
                break;
            }
        }
        if (refits < 0) {
            cur = ordinary;
        }
    }
    for (Index a = 0; a < p; ++a) {
        params[a] = cur[a] / scales[a];
    }
    return refits;
}

} // namespace

extern "C" {

const char* oracle_last_error() { return g_error.c_str(); }

double oracle_default_gradient_tolerance(const double* design, std::int64_t rows,
                                         std::int64_t cols) {
    return default_gradient_tolerance(from_ptr(design, rows, cols));
}

// Generic NNLS; returns 1 when converged, 0 when the inner-step limit stopped it, -1 on error.
int oracle_nnls(const double* design, std::int64_t rows, std::int64_t cols, const double* data,
                double tol, std::int64_t max_inner, double* solution) {
    NnlsResult r;
    const int status = guarded([&] {
        const Mat c = from_ptr(design, rows, cols);
        const Vec d(data, data + rows);
        r = solve_nnls_with_limit(c, d, tol, max_inner, nullptr);
    });
    if (status) {
        return -1;
    }
    std::copy(r.solution.begin(), r.solution.end(), solution);
    return r.converged ? 1 : 0;
}

int oracle_ridge_nnls(const double* kernel, std::int64_t rows, std::int64_t cols,
                      const double* signal, double ridge, double l1, double tol,
                      double* solution, std::int64_t* trace) {
    return guarded([&] {
        const RidgeDesign d = make_ridge_design(from_ptr(kernel, rows, cols), ridge);
        NnlsTrace tr;
        const Vec x = solve_ridge_nnls(d, Vec(signal, signal + rows), l1, tol, &tr);
        std::copy(x.begin(), x.end(), solution);
        if (trace) {
            trace[0] = tr.outer;
            trace[1] = tr.inner;
            trace[2] = tr.max_active;
            trace[3] = tr.qr_fallback;
            trace[4] = tr.final_active;
        }
    });
}

int oracle_ridge_normal_equations(const double* kernel, std::int64_t rows, std::int64_t cols,
                                  double ridge) {
    return make_ridge_design(from_ptr(kernel, rows, cols), ridge).normal_equations ? 1 : 0;
}

// The ridge design's Gram matrix K^T K + r^2 I, (cols, cols).
void oracle_ridge_gram(const double* kernel, std::int64_t rows, std::int64_t cols, double ridge,
                       double* gram) {
    const RidgeDesign d = make_ridge_design(from_ptr(kernel, rows, cols), ridge);
    std::copy(d.gram.v.begin(), d.gram.v.end(), gram);
}

// remove_cholesky_column on a (n, n) row-major factor, in place.
void oracle_remove_cholesky_column(double* factor, std::int64_t n, std::int64_t position) {
    Mat f = from_ptr(factor, n, n);
    remove_cholesky_column(f, n, position);
    std::copy(f.v.begin(), f.v.end(), factor);
}

double oracle_median(const double* values, std::int64_t n) {
    return median(Vec(values, values + n));
}

int oracle_least_squares(const double* design, std::int64_t rows, std::int64_t cols,
                         const double* data, double* coefficients) {
    return guarded([&] {
        const Vec x = least_squares(from_ptr(design, rows, cols), Vec(data, data + rows));
        std::copy(x.begin(), x.end(), coefficients);
    });
}

int oracle_compute_fractions(const double* fod, const double* k_aniso, const double* k_iso,
                             std::int64_t m, std::int64_t ca, std::int64_t ci,
                             const std::int64_t* blocks, std::int64_t n_blocks,
                             const double* signal, int lumped, int has_tol, double tol,
                             double* fractions) {
    return guarded([&] {
        const TissueKernel k = make_kernel(k_aniso, k_iso, k_iso, m, ca, ci, blocks, n_blocks);
        const Vec f = compute_fractions(Vec(fod, fod + ca), k, Vec(signal, signal + m),
                                        lumped != 0, has_tol != 0, tol, nullptr);
        std::copy(f.begin(), f.end(), fractions);
    });
}

// fit_single_tissue: one ridge NNLS per voxel with a prepared solver.
int oracle_fit_single_tissue(const double* signals, std::int64_t n_voxels, std::int64_t m,
                             const double* kernel, std::int64_t n, double ridge, double tol,
                             int n_threads, double* fods) {
    return guarded([&] {
        const RidgeDesign d = make_ridge_design(from_ptr(kernel, m, n), ridge);
        parallel_for_blocks(n_voxels, 128, n_threads, [&](Index begin, Index end) {
            for (Index v = begin; v < end; ++v) {
                const Vec s(signals + v * m, signals + (v + 1) * m);
                const Vec x = solve_ridge_nnls(d, s, 0.0, tol, nullptr);
                std::copy(x.begin(), x.end(), fods + v * n);
            }
        });
    });
}

// fit_multi_tissue. trace (optional): per voxel 10 int64s
// [rounds, lr_outer_sum, lr_inner_sum, lr_final_active_sum, lr_max_active, lr_qr_fallbacks,
//  hr_outer, hr_inner, hr_max_active, hr_qr_fallback].
int oracle_fit_multi_tissue(const double* signals, std::int64_t n_voxels, std::int64_t m,
                            const double* weights, const double* lr_k_aniso,
                            const double* lr_k_iso, const double* lr_w_k_iso,
                            const double* lr_w_k_aniso, std::int64_t lr_ca,
                            const std::int64_t* lr_blocks, std::int64_t lr_n_blocks,
                            double lr_ridge, double lr_tol, const double* hr_k_aniso,
                            const double* hr_k_iso, const double* hr_w_k_iso,
                            const double* hr_w_k_aniso, std::int64_t hr_ca,
                            const std::int64_t* hr_blocks, std::int64_t hr_n_blocks,
                            double hr_ridge, double hr_tol, std::int64_t ci,
                            std::int64_t max_rounds, double fraction_tol, double final_tol,
                            int n_threads, double* fods, double* fractions,
                            std::int64_t* trace) {
    return guarded([&] {
        const TissueKernel lr =
            make_kernel(lr_k_aniso, lr_k_iso, lr_w_k_iso, m, lr_ca, ci, lr_blocks, lr_n_blocks);
        const TissueKernel hr =
            make_kernel(hr_k_aniso, hr_k_iso, hr_w_k_iso, m, hr_ca, ci, hr_blocks, hr_n_blocks);
        Solver lr_solver;
        lr_solver.design = make_ridge_design(from_ptr(lr_w_k_aniso, m, lr_ca), lr_ridge);
        lr_solver.tol = lr_tol;
        Solver hr_solver;
        hr_solver.design = make_ridge_design(from_ptr(hr_w_k_aniso, m, hr_ca), hr_ridge);
        hr_solver.tol = hr_tol;
        const Vec w(weights, weights + m);
        const Index nf = hr_n_blocks + ci;
        parallel_for_blocks(n_voxels, 128, n_threads, [&](Index begin, Index end) {
            const Index nb = end - begin;
            const Mat block = from_ptr(signals + begin * m, nb, m);
            std::vector<MtTrace> tr(trace ? static_cast<size_t>(nb) : 0);
            fit_multi_tissue_block(block, w, lr, lr_solver, hr, hr_solver, max_rounds,
                                   fraction_tol, final_tol, fods + begin * hr_ca,
                                   fractions + begin * nf, trace ? tr.data() : nullptr);
            if (trace) {
                for (Index v = 0; v < nb; ++v) {
                    std::int64_t* t = trace + (begin + v) * 10;
                    t[0] = tr[v].rounds;
                    t[1] = tr[v].lr_sum.outer;
                    t[2] = tr[v].lr_sum.inner;
                    t[3] = tr[v].lr_sum.max_active;
                    t[4] = tr[v].lr_max_active;
                    t[5] = tr[v].lr_qr_fallbacks;
                    t[6] = tr[v].hr.outer;
                    t[7] = tr[v].hr.inner;
                    t[8] = tr[v].hr.max_active;
                    t[9] = tr[v].hr.qr_fallback;
                }
            }
        });
    });
}

// fit_single_tissue with FodSolver.drl (solver.cpp:9-12, drl.cpp:39-99): kernel_t (c, m),
// gram (c, c).  iterations (optional, int64 per voxel): the iterations each voxel ran, negated
// when the iteration limit stopped it.
int oracle_fit_single_tissue_drl(const double* signals, std::int64_t n_voxels, std::int64_t m,
                                 const double* kernel_t, std::int64_t c, const double* gram,
                                 double nu, std::int64_t n_iter, double init, int n_threads,
                                 double* fods, std::int64_t* iterations) {
    return guarded([&] {
        const DrlSolver d{from_ptr(kernel_t, c, m), from_ptr(gram, c, c), nu, n_iter, init};
        parallel_for_blocks(n_voxels, 128, n_threads, [&](Index begin, Index end) {
            for (Index v = begin; v < end; ++v) {
                const Vec sig(signals + v * m, signals + (v + 1) * m);
                Index it = 0;
                bool converged = false;
                const Vec f = drl_voxel(d, sig, &it, &converged);
                std::copy(f.begin(), f.end(), fods + v * c);
                if (iterations) iterations[v] = converged ? it : -it;
            }
        });
    });
}

// fit_multi_tissue with FodSolver.drl at both resolutions (the GRL method): each stage's
// kernel_t (c, m), gram (c, c), nu, iteration limit and starting amplitude.  rounds (optional,
// int64 per voxel).
int oracle_fit_multi_tissue_drl(
    const double* signals, std::int64_t n_voxels, std::int64_t m, const double* weights,
    const double* lr_k_aniso, const double* lr_k_iso, const double* lr_w_k_iso, std::int64_t lr_ca,
    const std::int64_t* lr_blocks, std::int64_t lr_n_blocks, const double* lr_kernel_t,
    const double* lr_gram, double lr_nu, std::int64_t lr_n_iter, double lr_init,
    const double* hr_k_aniso, const double* hr_k_iso, const double* hr_w_k_iso, std::int64_t hr_ca,
    const std::int64_t* hr_blocks, std::int64_t hr_n_blocks, const double* hr_kernel_t,
    const double* hr_gram, double hr_nu, std::int64_t hr_n_iter, double hr_init, std::int64_t ci,
    std::int64_t max_rounds, double fraction_tol, double final_tol, int n_threads, double* fods,
    double* fractions, std::int64_t* rounds) {
    return guarded([&] {
        const TissueKernel lr =
            make_kernel(lr_k_aniso, lr_k_iso, lr_w_k_iso, m, lr_ca, ci, lr_blocks, lr_n_blocks);
        const TissueKernel hr =
            make_kernel(hr_k_aniso, hr_k_iso, hr_w_k_iso, m, hr_ca, ci, hr_blocks, hr_n_blocks);
        Solver lr_solver;
        lr_solver.drl = true;
        lr_solver.rl = DrlSolver{from_ptr(lr_kernel_t, lr_ca, m), from_ptr(lr_gram, lr_ca, lr_ca),
                                 lr_nu, lr_n_iter, lr_init};
        Solver hr_solver;
        hr_solver.drl = true;
        hr_solver.rl = DrlSolver{from_ptr(hr_kernel_t, hr_ca, m), from_ptr(hr_gram, hr_ca, hr_ca),
                                 hr_nu, hr_n_iter, hr_init};
        const Vec w(weights, weights + m);
        const Index nf = hr_n_blocks + ci;
        parallel_for_blocks(n_voxels, 128, n_threads, [&](Index begin, Index end) {
            const Index nb = end - begin;
            const Mat block = from_ptr(signals + begin * m, nb, m);
            std::vector<MtTrace> tr(rounds ? static_cast<size_t>(nb) : 0);
            fit_multi_tissue_block(block, w, lr, lr_solver, hr, hr_solver, max_rounds,
                                   fraction_tol, final_tol, fods + begin * hr_ca,
                                   fractions + begin * nf, rounds ? tr.data() : nullptr);
            if (rounds) {
                for (Index v = 0; v < nb; ++v) rounds[begin + v] = tr[v].rounds;
            }
        });
    });
}

// The standardised elastic-net design: scaled columns (m, c) and scales (c).
int oracle_elastic_net_design(const double* kernel, std::int64_t m, std::int64_t c,
                              double* columns, double* scales) {
    return guarded([&] {
        const ElasticNet e = make_elastic_net(from_ptr(kernel, m, c), 0.0, 1.0, 1.0, 1);
        std::copy(e.columns.v.begin(), e.columns.v.end(), columns);
        std::copy(e.scales.begin(), e.scales.end(), scales);
    });
}

// fit_single_tissue with FodSolver.lasso (ElasticNetSolver::solve_block, solver.cpp:23-35):
// sweeps (optional, int64 per voxel) holds the sweeps run, negative when max_sweeps stopped it.
int oracle_fit_single_tissue_lasso(const double* signals, std::int64_t n_voxels, std::int64_t m,
                                   const double* kernel, std::int64_t c, double penalty,
                                   double l1_ratio, double tol, std::int64_t max_sweeps,
                                   int n_threads, double* fods, std::int64_t* sweeps) {
    return guarded([&] {
        const ElasticNet e = make_elastic_net(from_ptr(kernel, m, c), penalty, l1_ratio, tol,
                                              max_sweeps);
        parallel_for_blocks(n_voxels, 128, n_threads, [&](Index begin, Index end) {
            for (Index v = begin; v < end; ++v) {
                const Vec sig(signals + v * m, signals + (v + 1) * m);
                bool converged = false;
                Index n = 0;
                const Vec f = elastic_net_voxel(e, sig, &converged, &n);
                std::copy(f.begin(), f.end(), fods + v * c);
                if (sweeps) sweeps[v] = converged ? n : -n;
            }
        });
    });
}

// fit_multi_tissue with FodSolver.lasso at both resolutions: each stage's weighted anisotropic
// kernel (m, ca) and its settings.  rounds (optional, int64 per voxel).
int oracle_fit_multi_tissue_lasso(
    const double* signals, std::int64_t n_voxels, std::int64_t m, const double* weights,
    const double* lr_k_aniso, const double* lr_k_iso, const double* lr_w_k_iso,
    const double* lr_w_k_aniso, std::int64_t lr_ca, const std::int64_t* lr_blocks,
    std::int64_t lr_n_blocks, const double* hr_k_aniso, const double* hr_k_iso,
    const double* hr_w_k_iso, const double* hr_w_k_aniso, std::int64_t hr_ca,
    const std::int64_t* hr_blocks, std::int64_t hr_n_blocks, double penalty, double l1_ratio,
    double tol, std::int64_t max_sweeps, std::int64_t ci, std::int64_t max_rounds,
    double fraction_tol, double final_tol, int n_threads, double* fods, double* fractions,
    std::int64_t* rounds) {
    return guarded([&] {
        const TissueKernel lr =
            make_kernel(lr_k_aniso, lr_k_iso, lr_w_k_iso, m, lr_ca, ci, lr_blocks, lr_n_blocks);
        const TissueKernel hr =
            make_kernel(hr_k_aniso, hr_k_iso, hr_w_k_iso, m, hr_ca, ci, hr_blocks, hr_n_blocks);
        Solver lr_solver;
        lr_solver.lasso = true;
        lr_solver.net = make_elastic_net(from_ptr(lr_w_k_aniso, m, lr_ca), penalty, l1_ratio, tol,
                                         max_sweeps);
        Solver hr_solver;
        hr_solver.lasso = true;
        hr_solver.net = make_elastic_net(from_ptr(hr_w_k_aniso, m, hr_ca), penalty, l1_ratio, tol,
                                         max_sweeps);
        const Vec w(weights, weights + m);
        const Index nf = hr_n_blocks + ci;
        parallel_for_blocks(n_voxels, 128, n_threads, [&](Index begin, Index end) {
            const Index nb = end - begin;
            const Mat block = from_ptr(signals + begin * m, nb, m);
            std::vector<MtTrace> tr(rounds ? static_cast<size_t>(nb) : 0);
            fit_multi_tissue_block(block, w, lr, lr_solver, hr, hr_solver, max_rounds,
                                   fraction_tol, final_tol, fods + begin * hr_ca,
                                   fractions + begin * nf, rounds ? tr.data() : nullptr);
            if (rounds) {
                for (Index v = 0; v < nb; ++v) rounds[begin + v] = tr[v].rounds;
            }
        });
    });
}

// fit_csd (cpp/src/python/sd.cpp:171-206): converged (uint8) and solves (int32, optional) per voxel.
int oracle_fit_csd(const double* signals, std::int64_t n_voxels, std::int64_t m,
                   const double* initial, std::int64_t c0, const double* forward, std::int64_t c,
                   const double* constraint, std::int64_t k, double norm_weight, double threshold,
                   std::int64_t max_iterations, int n_threads, double* fods,
                   std::uint8_t* converged, std::int32_t* solves) {
    return guarded([&] {
        const CsdProblem p =
            make_csd_problem(from_ptr(initial, c0, m), from_ptr(forward, m, c),
                             from_ptr(constraint, k, c), norm_weight, threshold, max_iterations);
        parallel_for_blocks(n_voxels, 128, n_threads, [&](Index begin, Index end) {
            for (Index v = begin; v < end; ++v) {
                const CsdOut r = solve_csd(p, Vec(signals + v * m, signals + (v + 1) * m));
                std::copy(r.fod.begin(), r.fod.end(), fods + v * c);
                converged[v] = r.converged ? 1 : 0;
                if (solves) {
                    solves[v] = static_cast<std::int32_t>(r.solves);
                }
            }
        });
    });
}

// The CSD Gram matrix F^T F + norm_weight I, for checking the library's.
int oracle_csd_gram(const double* forward, std::int64_t m, std::int64_t c, double norm_weight,
                    double* gram) {
    return guarded([&] {
        const CsdProblem p =
            make_csd_problem(Mat(0, m), from_ptr(forward, m, c), Mat(0, c), norm_weight, 0.0, 0);
        std::copy(p.gram.v.begin(), p.gram.v.end(), gram);
    });
}

// sh_derivatives: out = (value, d_theta, d_phi, d_theta_theta, d_theta_phi, d_phi_phi).
void oracle_sh_derivatives(const double* coeffs, std::int64_t lmax, int full, double theta,
                           double phi, double* out) {
    const ShDerivatives r = sh_derivatives(coeffs, lmax, full != 0, theta, phi);
    out[0] = r.value;
    out[1] = r.d_theta;
    out[2] = r.d_phi;
    out[3] = r.d_theta_theta;
    out[4] = r.d_theta_phi;
    out[5] = r.d_phi_phi;
}

// climb (cpp/src/python/sh.cpp:86-108): 1 and the peak, or 0 (NaN outputs) when it did not stop.
int oracle_climb(const double* coeffs, std::int64_t lmax, int full, const double* start,
                 std::int64_t max_iters, double tol, double max_step, double* direction,
                 double* amplitude) {
    ShPeak p;
    if (sh_climb(coeffs, lmax, full != 0, start, max_iters, tol, max_step, p)) {
        for (int a = 0; a < 3; ++a) direction[a] = p.dir[a];
        *amplitude = p.amplitude;
        return 1;
    }
    const double nan = std::numeric_limits<double>::quiet_NaN();
    for (int a = 0; a < 3; ++a) direction[a] = nan;
    *amplitude = nan;
    return 0;
}

// peaks_chunk (cpp/src/python/sh.cpp:112-175): directions (n, n_peaks, 3), amplitudes (n, n_peaks).
int oracle_peaks_chunk(const double* coeffs, std::int64_t n_voxels, std::int64_t n_coeffs,
                       std::int64_t lmax, int full, const double* on_sphere, std::int64_t nv,
                       const double* vertices, const std::int64_t* neighbours,
                       std::int64_t n_peaks, double relative_threshold, double threshold,
                       int n_threads, double* directions, double* amplitudes) {
    return guarded([&] {
        parallel_for_blocks(n_voxels, 16, n_threads, [&](Index begin, Index end) {
            for (Index v = begin; v < end; ++v) {
                sh_voxel_peaks(coeffs + v * n_coeffs, lmax, full != 0, on_sphere + v * nv, nv,
                               vertices, neighbours, n_peaks, relative_threshold, threshold,
                               directions + v * n_peaks * 3, amplitudes + v * n_peaks);
            }
        });
    });
}

// LsiDesign(a, g, ridge) then fit_lsi (cpp/src/python/sd.cpp:360-394): x (n_voxels, n),
// converged (uint8), trace (optional, 4 int64 per voxel: fallback, outer, inner, max factor).
int oracle_fit_lsi(const double* signals, std::int64_t n_voxels, const double* a, std::int64_t m,
                   std::int64_t n, const double* g, std::int64_t k, double ridge, int n_threads,
                   double* x, std::uint8_t* converged, std::int64_t* trace) {
    return guarded([&] {
        const LsiDesign d = make_lsi_design(from_ptr(a, m, n), from_ptr(g, k, n), ridge);
        parallel_for_blocks(n_voxels, 128, n_threads, [&](Index begin, Index end) {
            for (Index v = begin; v < end; ++v) {
                LsiTrace tr;
                const LsiOut o = solve_lsi(d, Vec(signals + v * m, signals + (v + 1) * m), &tr);
                std::copy(o.x.begin(), o.x.end(), x + v * n);
                converged[v] = o.converged ? 1 : 0;
                if (trace) {
                    trace[4 * v + 0] = tr.fallback;
                    trace[4 * v + 1] = tr.outer;
                    trace[4 * v + 2] = tr.inner;
                    trace[4 * v + 3] = tr.max_factor;
                }
            }
        });
    });
}

// The design matrices: q_t (n, m), back (n, n), e (k, n), gram (k, k), abs_sums (k).
int oracle_lsi_design(const double* a, std::int64_t m, std::int64_t n, const double* g,
                      std::int64_t k, double ridge, double* q_t, double* back, double* e,
                      double* gram, double* abs_sums) {
    return guarded([&] {
        const LsiDesign d = make_lsi_design(from_ptr(a, m, n), from_ptr(g, k, n), ridge);
        std::copy(d.q_transposed.v.begin(), d.q_transposed.v.end(), q_t);
        std::copy(d.back_substitution.v.begin(), d.back_substitution.v.end(), back);
        std::copy(d.constraints.v.begin(), d.constraints.v.end(), e);
        std::copy(d.constraint_gram.v.begin(), d.constraint_gram.v.end(), gram);
        std::copy(d.constraint_abs_sums.begin(), d.constraint_abs_sums.end(), abs_sums);
    });
}

// solve_lsi_by_qr (lsi.cpp:298-300) of one data vector: the QR-based NNLS only.
int oracle_solve_lsi_by_qr(const double* signal, const double* a, std::int64_t m, std::int64_t n,
                           const double* g, std::int64_t k, double ridge, double* x,
                           std::uint8_t* converged) {
    return guarded([&] {
        const LsiDesign d = make_lsi_design(from_ptr(a, m, n), from_ptr(g, k, n), ridge);
        const LsiOut o = solve_lsi(d, Vec(signal, signal + m), nullptr, true);
        std::copy(o.x.begin(), o.x.end(), x);
        *converged = o.converged ? 1 : 0;
    });
}

double oracle_exp_nonpositive(double x) { return exp_nonpositive(x); }

// The log-signal fits of n_vox voxels (log_s (n_vox, n)) on a prepared design.
int oracle_fit_log_signal(const double* xt, const double* pinv, const double* scales,
                          std::int64_t n, std::int64_t p, const double* log_s, std::int64_t n_vox,
                          int weighted, std::int64_t max_reweightings, double tol, double* params,
                          std::int32_t* refits) {
    return guarded([&] {
        for (std::int64_t v = 0; v < n_vox; ++v) {
            refits[v] = fit_log_signal_voxel(xt, pinv, scales, n, p, log_s + v * n, weighted != 0,
                                             max_reweightings, tol, params + v * p);
        }
    });
}

} // extern "C"
