// Elastic-net FOD solves by coordinate descent (cpp/src/sd/elastic_net.cpp:40-151,
// ElasticNetSolver::solve_block cpp/src/sd/solver.cpp:23-35), one wave per voxel.
//
// The wave holds its voxel's residual as a RowVec (row r in lane r % 64, slot r / 64) and its
// coefficients as a LaneVec (column j in lane j % 64, slot j / 64).  A coordinate update reads
// the scaled column j (rows contiguous: RPL coalesced loads, L2-resident across the grid), takes
// its wave sum with the residual, soft-thresholds, and, when the coefficient moves, updates the
// residual in registers.  The sweep over the non-zero coefficients walks a snapshot of them kept
// as one ballot mask per slot (scalar registers), in increasing column order as the reference's
// list.  Every decision is wave-uniform.
#pragma once

#include "kernels.hpp"

namespace lsd {

struct LassoArgs {
    const double* signals;    // (n_vox, m)
    int m;
    int multi;                // 1: the anisotropic signal max(w s - W_iso p, 0); 0: the signal
    const double* weights;    // (m), multi only
    const double* w_iso;      // (m, ci), multi only
    const double* iso;        // (n_vox, ci), multi only
    int ci;
    const double* columns_t;  // (c, m) the scaled columns, one column's rows contiguous
    const double* scales;     // (c) root mean square of each column (1 for a zero column)
    const int* nonzero;       // (c) 1 for a column with a non-zero entry
    int c;
    double l1_threshold;      // penalty * l1_ratio
    double shrink;            // 1 + penalty (1 - l1_ratio)
    double tol;               // relative change that ends the descent
    long max_sweeps;
    const int* list;          // voxels by list position, or null for all
    const int* list_count;
    long n_vox;
    double* x_out;            // (n_vox, c)
    int* status;              // (n_vox): ST_NOT_CONVERGED OR-ed in when max_sweeps stopped it
    unsigned long long* unconverged; // the solver's count of such solves
    unsigned long long* work; // profiled launches: the solves' FP64 operations
};

inline bool attach_work(LassoArgs& a, unsigned long long* slot) {
    a.work = slot;
    return slot != nullptr;
}

// elastic_net.cpp:13-21.
__device__ __forceinline__ double soft_threshold(double value, double threshold) {
    return value > threshold ? value - threshold : (value < -threshold ? value + threshold : 0.0);
}

template <int NCPL, int RPL>
__global__ void __launch_bounds__(WAVE * WAVES_PER_BLOCK) lasso_kernel(LassoArgs A) {
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / WAVE);
    const int l = lane_id();
    const int m = A.m;
    const int c = A.c;
    const double row_count = static_cast<double>(m);
    // Both divisors are fixed for the launch: their refined reciprocals once, then each quotient
    // is the division's last three operations (bit-identical to a / d where the operands are in
    // range, a plain division elsewhere).
    const double y_rows = division_reciprocal(row_count);
    const bool rows_ok = divisor_in_range(row_count);
    const double y_shrink = division_reciprocal(A.shrink);
    const bool shrink_ok = divisor_in_range(A.shrink);
    const long count = A.list ? static_cast<long>(*A.list_count) : A.n_vox;
    const long stride = static_cast<long>(gridDim.x) * WAVES_PER_BLOCK;
    unsigned long long wave_work = 0;
    for (long it = static_cast<long>(blockIdx.x) * WAVES_PER_BLOCK + w; it < count; it += stride) {
        const long v = A.list ? static_cast<long>(__builtin_amdgcn_readfirstlane(A.list[it])) : it;
        double p[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            p[q] = (A.multi && q < A.ci) ? A.iso[v * A.ci + q] : 0.0;
        }
        // The data: the signal or the anisotropic signal (multi_tissue.cpp:20-24), rows past m 0.
        double res[RPL];
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            const int r = l + WAVE * t;
            double a = 0.0;
            if (r < m) {
                const double s = A.signals[v * m + r];
                a = s;
                if (A.multi) {
                    const double ws = s * A.weights[r];
                    double wp = 0.0;
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        if (q < A.ci) {
                            wp = wp + A.w_iso[static_cast<long>(r) * A.ci + q] * p[q];
                        }
                    }
                    a = fmax(ws - wp, 0.0);
                }
            }
            res[t] = a;
        }
        auto load_column = [&](int j, double (&col)[RPL]) {
            const double* cp = A.columns_t + static_cast<long>(j) * m;
#pragma unroll
            for (int t = 0; t < RPL; ++t) {
                const int r = l + WAVE * t;
                col[t] = r < m ? cp[r] : 0.0;
            }
        };
        // The correlation of column col with the residual, divided by m: a wave sum.
        auto correlation = [&](const double (&col)[RPL]) {
            double part = 0.0;
#pragma unroll
            for (int t = 0; t < RPL; ++t) {
                if (l + WAVE * t < m) {
                    part = part + col[t] * res[t];
                }
            }
            return divide_nonzero_by(butterfly_sum(part), row_count, y_rows, rows_ok);
        };
        // has_nonzero_solution (elastic_net.cpp:51-57).
        double largest = 0.0;
        for (int j = 0; j < c; ++j) {
            if (A.nonzero[j]) {
                double col[RPL];
                load_column(j, col);
                largest = fmax(largest, fabs(correlation(col)));
            }
        }
        wave_work += 2ull * m * c;
        LaneVec<NCPL> coef;
        coef.fill(0.0);
        bool done = true;
        if (largest > A.l1_threshold) {
            // update (elastic_net.cpp:104-116): coefficient j to its minimiser with the others
            // fixed; returns its change relative to 1 + its old size.
            auto update = [&](int j) {
                double col[RPL];
                load_column(j, col);
                const double old_value = coef.get(j);
                const double corr = correlation(col) + old_value;
                const double new_value = divide_nonzero_by(soft_threshold(corr, A.l1_threshold),
                                                           A.shrink, y_shrink, shrink_ok);
                wave_work += 2ull * m;
                if (new_value == old_value) {
                    return 0.0;
                }
                const double delta = new_value - old_value;
#pragma unroll
                for (int t = 0; t < RPL; ++t) {
                    res[t] = res[t] - delta * col[t];
                }
                coef.set(j, new_value);
                wave_work += 2ull * m;
                return fabs(delta) / (1.0 + fabs(old_value));
            };
            // solve_elastic_net (elastic_net.cpp:128-151): full sweeps, each followed by sweeps
            // over the coefficients it left non-zero until they settle.
            long sweeps = 0;
            done = false;
            while (sweeps < A.max_sweeps) {
                ++sweeps;
                double change = 0.0;
                for (int j = 0; j < c; ++j) {
                    if (A.nonzero[j]) {
                        change = fmax(change, update(j));
                    }
                }
                if (change < A.tol) {
                    done = true;
                    break;
                }
                unsigned long long active[NCPL];
#pragma unroll
                for (int t = 0; t < NCPL; ++t) {
                    active[t] = ballot(coef.v[t] != 0.0);
                }
                while (sweeps < A.max_sweeps) {
                    ++sweeps;
                    double inner = 0.0;
#pragma unroll
                    for (int t = 0; t < NCPL; ++t) {
                        unsigned long long bits = active[t];
                        while (bits != 0) {
                            const int j = t * WAVE + __builtin_ctzll(bits);
                            bits &= bits - 1;
                            inner = fmax(inner, update(j));
                        }
                    }
                    if (inner < A.tol) {
                        break;
                    }
                }
            }
        }
        // The coefficients of the unscaled columns.
#pragma unroll
        for (int t = 0; t < NCPL; ++t) {
            const int j = l + WAVE * t;
            if (j < c) {
                A.x_out[v * c + j] = coef.v[t] / A.scales[j];
            }
        }
        if (!done && l == 0) {
            A.status[v] |= ST_NOT_CONVERGED;
            atomicAdd(A.unconverged, 1ull);
        }
    }
    if (A.work != nullptr && l == 0 && wave_work != 0) {
        atomicAdd(A.work, wave_work);
    }
}

// Launches lasso_kernel<ncpl, rpl> (inst_lasso.hip) on at most `waves` waves.
int launch_lasso(const char* name, int ncpl, int rpl, const LassoArgs& a, int waves,
                 hipStream_t stream);

} // namespace lsd
