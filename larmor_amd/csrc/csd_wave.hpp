// Constrained spherical deconvolution (CSD) of one voxel per wavefront.
//
// Restates larmor's solve_csd (cpp/src/sd/csd.cpp:23-65): the FOD starts from the linear
// estimate initial * s (its first c0 coefficients); each iteration evaluates the weighted FOD on
// the constraint directions, a = C f, constrains the directions with a_d < threshold, and, unless
// that set equals the previous iteration's, solves
//     (G + sum over constrained d, in increasing d, of c_d c_d^T) f = F^T s
// by Cholesky (G = F^T F + norm_weight I, made once on the host).  At most max_iter solves; a
// voxel converges when its set repeats (or max_iter is 0).
//
// Layout: lane i holds row i of the N x N system in N registers (N <= 64 coefficients); the
// factor overwrites it in place (L_ij in lane i, register j, for j <= i).  The Cholesky is
// right-looking: at step j lane j's diagonal and the column L_kj are broadcast with readlane and
// every lane updates its own row, so each entry receives the terms L_ip L_jp in increasing p --
// the order of the left-looking restatement in the oracle.  The backward solve needs column i
// of L in lane i, so after the forward solve each lane writes its row of L into a per-wave LDS
// copy of the packed lower triangle, and lane i reads L_ji from there.
//
// Orders (shared with the oracle, which makes the results bit-identical): every sum over
// measurements, coefficients or constrained directions runs in increasing index order; the
// forward solve subtracts earlier positions first, the backward solve later positions first, a
// zero right-hand entry stays zero.
#pragma once

#include "wave.hpp"

namespace lsd {

struct CsdArgs {
    const double* signals;      // (n_vox, m)
    long n_vox;
    int m;
    int c0;                     // rows of the linear start
    const double* initial_t;    // (m, c0): the linear start, transposed
    const double* forward;      // (m, N)
    const double* constraint;   // (k, N), weighted
    const double* constraint_t; // (N, k)
    const double* gram;         // (N, N)
    int k;                      // constraint directions
    double threshold;
    int max_iter;
    double* fods;               // (n_vox, N)
    unsigned char* converged;   // (n_vox)
    int* solves;                // (n_vox), or null
};

// Doubles of the per-wave LDS copy of the factor: the packed lower triangle.
constexpr int csd_lds_doubles(int n) { return n * (n + 1) / 2; }

// Offset of column i in the packed lower triangle (column-major: column i holds rows i..n-1).
__device__ __forceinline__ int packed_column(int n, int i) { return i * n - (i * (i - 1)) / 2; }

template <int N, int NK>
__global__ void __launch_bounds__(WAVE * 4) csd_kernel(CsdArgs A) {
    static_assert(N <= WAVE, "one system row per lane");
    extern __shared__ double smem[];
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / WAVE);
    const long gw = static_cast<long>(blockIdx.x) * (blockDim.x / WAVE) + w;
    const long n_waves = static_cast<long>(gridDim.x) * (blockDim.x / WAVE);
    double* lt = smem + w * csd_lds_doubles(N);
    const int l = lane_id();
    const bool row_lane = l < N;
    const int my_column = packed_column(N, l < N ? l : 0);

    for (long v = gw; v < A.n_vox; v += n_waves) {
        const double* s = A.signals + v * A.m;
        // F^T s (every solve's right-hand side) and the linear start, summed over measurements.
        double proj = 0.0;
        double f = 0.0;
        for (int r = 0; r < A.m; ++r) {
            const double sr = s[r];
            if (row_lane) {
                proj = proj + A.forward[static_cast<long>(r) * N + l] * sr;
            }
            if (l < A.c0) {
                f = f + A.initial_t[static_cast<long>(r) * A.c0 + l] * sr;
            }
        }
        bool first = true;
        bool converged = A.max_iter == 0;
        int solves = 0;
        unsigned long long prev[NK];
#pragma unroll
        for (int t = 0; t < NK; ++t) {
            prev[t] = 0ull;
        }
        for (int it = 0; it < A.max_iter; ++it) {
            // The weighted amplitudes on the constraint directions (direction d in lane d % 64,
            // slot d / 64), summed over the coefficients in order.
            double a[NK];
#pragma unroll
            for (int t = 0; t < NK; ++t) {
                a[t] = 0.0;
            }
#pragma unroll 3
            for (int j = 0; j < N; ++j) {
                const double fj = readlane(f, j);
#pragma unroll
                for (int t = 0; t < NK; ++t) {
                    const int d = l + WAVE * t;
                    if (d < A.k) {
                        a[t] = a[t] + A.constraint_t[static_cast<long>(j) * A.k + d] * fj;
                    }
                }
            }
            unsigned long long mask[NK];
            bool same = true;
#pragma unroll
            for (int t = 0; t < NK; ++t) {
                const int d = l + WAVE * t;
                mask[t] = ballot(d < A.k && a[t] < A.threshold);
                same = same && mask[t] == prev[t];
            }
            if (!first && same) {
                converged = true;
                break;
            }
            first = false;

            // Row l of G + sum of c_d c_d^T over the constrained directions in increasing d.
            double row[N];
#pragma unroll
            for (int j = 0; j < N; ++j) {
                row[j] = row_lane ? A.gram[l * N + j] : 0.0;
            }
#pragma unroll
            for (int t = 0; t < NK; ++t) {
                unsigned long long bits = mask[t];
                while (bits) {
                    const int b = __ffsll(static_cast<long long>(bits)) - 1;
                    bits &= bits - 1;
                    const double* cd = A.constraint + static_cast<long>(WAVE * t + b) * N;
                    // c_d[l] in lane l; c_d[j] broadcast from lane j.
                    const double ci = row_lane ? cd[l] : 0.0;
#pragma unroll
                    for (int j = 0; j < N; ++j) {
                        row[j] = row[j] + ci * readlane(ci, j);
                    }
                }
            }

            // Cholesky, right-looking; lane i keeps L_ij in row[j] for j <= i (the entries
            // right of the diagonal are never read).
#pragma unroll
            for (int j = 0; j < N; ++j) {
                const double ljj = sqrt(readlane(row[j], j));
                row[j] = l == j ? ljj : row[j] / ljj;
#pragma unroll
                for (int q = j + 1; q < N; ++q) {
                    const double lqj = readlane(row[j], q);
                    row[q] = row[q] - row[j] * lqj;
                }
            }

            // L y = F^T s, column-oriented.
            double b = proj;
#pragma unroll
            for (int j = 0; j < N; ++j) {
                const double bj = readlane(b, j);
                const double yj = bj != 0.0 ? bj / readlane(row[j], j) : bj;
                b = l == j ? yj : (l > j ? b - row[j] * yj : b);
            }

            // Lane j publishes row j of L (entries L_ji, i <= j) into the packed triangle.
            if (row_lane) {
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    if (i <= l) {
                        lt[packed_column(N, i) + (l - i)] = row[i];
                    }
                }
            }
            wave_lds_sync();

            // L^T f = y, column-oriented from the last position.
#pragma unroll
            for (int j = N - 1; j >= 0; --j) {
                const double yj = readlane(b, j);
                const double fj = yj != 0.0 ? yj / readlane(row[j], j) : yj;
                const double lji = l < j ? lt[my_column + (j - l)] : 0.0;
                b = l == j ? fj : (l < j ? b - lji * fj : b);
            }
            wave_lds_sync();
            f = row_lane ? b : 0.0;
            solves += 1;
#pragma unroll
            for (int t = 0; t < NK; ++t) {
                prev[t] = mask[t];
            }
        }
        if (row_lane) {
            A.fods[v * N + l] = f;
        }
        if (l == 0) {
            A.converged[v] = converged ? 1 : 0;
            if (A.solves) {
                A.solves[v] = solves;
            }
        }
    }
}

} // namespace lsd
