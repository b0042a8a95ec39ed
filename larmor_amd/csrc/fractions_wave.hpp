// Tissue fractions of one voxel on one wavefront: the median threshold of the FOD, the group
// signal columns and the NNLS fraction fit.
//
// Restates larmor's median (cpp/src/sd/fractions.cpp:31-43; NaN sorted last, the mean of the
// two middle values for an even count) and compute_fractions (fractions.cpp:45-76): FOD values
// below median * (1 - 1e-9) are zeroed, each group's signal K_group f is scaled to a largest
// value of 1 (when positive), and [groups, isotropic kernels] are fitted to the unweighted
// signal by the generic NNLS with the given tolerance or the default one.
#pragma once

#include "qr_wave.hpp"

namespace LSD_NS {

constexpr real MEDIAN_TIE_TOLERANCE = real(1e-9);

__device__ __forceinline__ bool less_nan_last(real a, real b) {
    return isnan(b) ? !isnan(a) : a < b;
}

constexpr int pow2_ceil(int v) {
    int p = 1;
    while (p < v) {
        p *= 2;
    }
    return p;
}

// The median of the n values spread as columns (lane c % 64, slot c / 64), by a bitonic sort of
// the wave's 64 * NP slots with NaN padding.
template <int NCPL>
__device__ __forceinline__ real median_cols(const real (&v)[NCPL], int n) {
    constexpr int NP = pow2_ceil(NCPL);
    constexpr int TOTAL = NP * WAVE;
    const int l = lane_id();
    rvec<NP> a;
#pragma unroll
    for (int t = 0; t < NP; ++t) {
        real x = __builtin_nan("");
        if (t < NCPL) {
            x = (l + WAVE * t) < n ? v[t < NCPL ? t : 0] : x;
        }
        a[t] = x;
    }
#pragma unroll
    for (int kk = 2; kk <= TOTAL; kk <<= 1) {
#pragma unroll
        for (int j = kk >> 1; j > 0; j >>= 1) {
            if (j >= WAVE) {
                const int js = j / WAVE;
#pragma unroll
                for (int t = 0; t < NP; ++t) {
                    const int tp = t ^ js;
                    if (tp > t) {
                        const int idx = t * WAVE + l;
                        const bool ascending = (idx & kk) == 0;
                        const real lo = a[t];
                        const real hi = a[tp];
                        const bool swap = ascending ? less_nan_last(hi, lo) : less_nan_last(lo, hi);
                        a[t] = swap ? hi : lo;
                        a[tp] = swap ? lo : hi;
                    }
                }
            } else {
#pragma unroll
                for (int t = 0; t < NP; ++t) {
                    const real other = xor_lanes(a[t], j);
                    const int idx = t * WAVE + l;
                    const bool ascending = (idx & kk) == 0;
                    const bool lower = (l & j) == 0;
                    const bool keep_min = lower == ascending;
                    const bool other_less = less_nan_last(other, a[t]);
                    const bool other_greater = less_nan_last(a[t], other);
                    a[t] = keep_min ? (other_less ? other : a[t]) : (other_greater ? other : a[t]);
                }
            }
        }
    }
    if (n == 0) {
        return __builtin_nan("");
    }
    auto at = [&](int i) { return readlane(static_cast<real>(a[i / WAVE]), i % WAVE); };
    const int mid = n / 2;
    if (n % 2 == 1) {
        return at(mid);
    }
    return (at(mid - 1) + at(mid)) / real(2.0);
}

// Shared matrices of one resolution's fraction fit.
struct FractionKernelDev {
    int m;
    int ca;                 // anisotropic columns
    int ci;                 // isotropic tissues
    const real* aniso_t;  // (ca, m): the unweighted anisotropic kernel, transposed
    const real* iso_t;    // (ci, m): the unweighted isotropic kernel, transposed
    const int* blocks;      // (n_blocks, 2) column ranges
    int n_blocks;
};

// The columns of a fraction fit: up to NG group signals in registers, then the isotropic
// kernel columns, read from the block's LDS copy (column c at c * RPL * 64, rows past m zero).
// Reading them from global memory instead lets the compiler hoist every load out of the NNLS
// loop and hold them all in registers.
template <int NG, int RPL>
struct FractionDesign {
    rvec<RPL> grp[NG];
    int n_groups;
    const real* iso_lds;

    __device__ __forceinline__ real get(int c, int t) const {
        if (c < n_groups) {
            real x = real(0.0);
#pragma unroll
            for (int q = 0; q < NG; ++q) {
                x = c == q ? static_cast<real>(grp[q][t]) : x;
            }
            return x;
        }
        return iso_lds[(c - n_groups) * (RPL * WAVE) + lane_id() + WAVE * t];
    }
};

constexpr int MAX_GROUPS = 3; // anisotropic blocks of a per-block fraction fit
constexpr int GROUP_BATCH = 4; // kernel rows loaded together for a group signal

// compute_fractions of one voxel: fod in column slots, s the unweighted signal.  lumped takes
// all anisotropic columns as one group; NG group columns are held (1 suffices when lumped).  fr
// receives the fractions, groups first.  work is the wave's LDS slab for the pivoted QR (RPL * 64
// doubles per column of the fit), iso_lds the block's LDS copy of the isotropic columns.
// Returns false when the NNLS hit its inner-step limit.
// NORMAL (the mixed-precision form's loop fits): the NNLS fits through the normal equations
// (nnls_normal) instead of the reference's pivoted QR.
template <int NCPL, int RPL, int NQ, int NG = MAX_GROUPS, bool NORMAL = false>
__device__ __forceinline__ bool fractions_wave(const FractionKernelDev& K, const real (&fod)[NCPL],
                                               const RowVec<RPL>& s, bool lumped, bool has_tol,
                                               real tol_in, real (&fr)[NQ], real* work,
                                               const real* iso_lds) {
    const int l = lane_id();
    const real threshold = median_cols<NCPL>(fod, K.ca) * (real(1.0) - MEDIAN_TIE_TOLERANCE);
    real thr[NCPL];
#pragma unroll
    for (int t = 0; t < NCPL; ++t) {
        thr[t] = fod[t] < threshold ? real(0.0) : fod[t];
    }
    FractionDesign<NG, RPL> C;
    C.n_groups = lumped ? 1 : K.n_blocks;
    C.iso_lds = iso_lds;
#pragma unroll
    for (int grp = 0; grp < NG; ++grp) {
        rvec<RPL> gs = real(0.0);
        if (grp < C.n_groups) {
            const int begin = lumped ? 0 : K.blocks[2 * grp];
            const int end = lumped ? K.ca : K.blocks[2 * grp + 1];
            // Zero terms add exactly nothing, so only the non-zero thresholded values are
            // visited, in column order; their kernel rows are loaded GROUP_BATCH at a time, so
            // the loads overlap, and then added in column order.
#pragma unroll
            for (int t2 = 0; t2 < NCPL; ++t2) {
                const int c0 = l + WAVE * t2;
                unsigned long long mask = ballot(c0 >= begin && c0 < end && thr[t2] != real(0.0));
                while (mask) {
                    real fb[GROUP_BATCH];
                    real kb[GROUP_BATCH][RPL];
                    bool valid[GROUP_BATCH];
#pragma unroll
                    for (int q = 0; q < GROUP_BATCH; ++q) {
                        valid[q] = mask != 0ull;
                        const int b = valid[q] ? __ffsll(static_cast<long long>(mask)) - 1 : 0;
                        mask &= mask - 1;
                        fb[q] = readlane(thr[t2], b);
                        const real* krow = K.aniso_t + static_cast<long>(WAVE * t2 + b) * K.m;
#pragma unroll
                        for (int t = 0; t < RPL; ++t) {
                            const int r = l + WAVE * t;
                            kb[q][t] = (valid[q] && r < K.m) ? krow[r] : real(0.0);
                        }
                    }
#pragma unroll
                    for (int q = 0; q < GROUP_BATCH; ++q) {
                        if (valid[q]) {
#pragma unroll
                            for (int t = 0; t < RPL; ++t) {
                                const int r = l + WAVE * t;
                                if (r < K.m) {
                                    gs[t] = gs[t] + kb[q][t] * fb[q];
                                }
                            }
                        }
                    }
                }
            }
            if (K.m > 0) {
                real largest = __builtin_nan("");
#pragma unroll
                for (int t = 0; t < RPL; ++t) {
                    const int r = l + WAVE * t;
                    if (r < K.m) {
                        largest = fmax(largest, static_cast<real>(gs[t]));
                    }
                }
#pragma unroll
                for (int o = WAVE / 2; o >= 1; o >>= 1) {
                    largest = fmax(largest, xor_lanes(largest, o));
                }
                largest = readlane(largest, 0);
                if (largest > real(0.0)) {
#pragma unroll
                    for (int t = 0; t < RPL; ++t) {
                        gs[t] = gs[t] / largest;
                    }
                }
            }
        }
        C.grp[grp] = gs;
    }
    const int ncols = C.n_groups + K.ci;
    const real tol = has_tol ? tol_in : default_tolerance_wave<NQ, RPL>(C, K.m, ncols);
    if constexpr (NORMAL) {
        (void)work;
        return nnls_normal<NQ, RPL>(C, K.m, ncols, s, tol, 3 * ncols, fr);
    } else {
        return nnls_small<NQ, RPL>(C, K.m, ncols, s, tol, 3 * ncols, fr, work);
    }
}

} // namespace LSD_NS
