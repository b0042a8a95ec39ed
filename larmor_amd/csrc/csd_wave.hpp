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
//
// csd_kernel holds a row per lane and so stops at 64 coefficients with a compile-time N;
// csd_wide_kernel below takes any N at run time with the system in memory.
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
    int n;                      // coefficients (csd_wide_kernel)
    double* sys_slabs;          // csd_wide_kernel with the system in global memory: one packed
                                // triangle per workgroup
};

// The most coefficients a CSD fit takes: lmax 30 (496 coefficients), whose system's triangle
// (~1 MB) goes to global memory.
constexpr int CSD_MAX_COEFFICIENTS = 496;

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

// The wide form: any number of coefficients N, one voxel per workgroup of CSD_WIDE_WAVES waves.
// The packed lower triangle of the system (column-major, csd_lds_doubles(N) doubles) lives in the
// workgroup's LDS, or, when it does not fit there (GLOBAL_SYS), in the workgroup's own slab of
// global memory; the FOD, F^T s, the solve's right-hand side and the constrained directions'
// list live in LDS.  Each step gives every entry the same terms in the same order as csd_kernel
// and the oracle:
//  - system: entry (i, j), i >= j, is G_ij plus c_d[i] c_d[j] over the constrained d in
//    increasing d (columns j over the waves, rows i over the lanes);
//  - Cholesky, right-looking: at step j the column below the diagonal is divided by
//    sqrt(A_jj), then every trailing entry (i, q) loses L_ij L_qj (columns q over the waves), so
//    entry (i, q) receives its terms in increasing j, as the left-looking oracle subtracts them;
//  - forward and backward solves column-oriented (earlier, then later positions first), on the
//    first wave.
// Every wave evaluates the constraint amplitudes and the constrained set itself (the same
// arithmetic, so the same decisions); the workgroup synchronises between steps (its barrier also
// orders the global slab's accesses at workgroup scope).
constexpr int CSD_WIDE_WAVES = 4;

// LDS bytes of a csd_wide_kernel workgroup: the triangle (unless in global memory), three
// vectors of N doubles and k direction indices.
inline size_t csd_wide_lds_bytes(int n, int k, bool global_sys) {
    const size_t tri = global_sys ? 0 : sizeof(double) * static_cast<size_t>(csd_lds_doubles(n));
    return tri + sizeof(double) * 3 * static_cast<size_t>(n) + sizeof(int) * static_cast<size_t>(k > 0 ? k : 1);
}

template <int NK, bool GLOBAL_SYS>
__global__ void __launch_bounds__(WAVE * CSD_WIDE_WAVES) csd_wide_kernel(CsdArgs A) {
    extern __shared__ double smem[];
    const int n = A.n;
    const long tri = static_cast<long>(n) * (n + 1) / 2;
    double* sys = GLOBAL_SYS ? A.sys_slabs + static_cast<long>(blockIdx.x) * tri : smem;
    double* f = GLOBAL_SYS ? smem : smem + tri;
    double* proj = f + n;
    double* b = proj + n;
    int* list = reinterpret_cast<int*>(b + n);
    const int l = lane_id();
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / WAVE);
    const int n_waves = static_cast<int>(blockDim.x) / WAVE;
    const int tid = static_cast<int>(threadIdx.x);
    const int n_threads = static_cast<int>(blockDim.x);
    auto column = [n](int i) { return static_cast<long>(i) * n - static_cast<long>(i) * (i - 1) / 2; };

    for (long v = blockIdx.x; v < A.n_vox; v += gridDim.x) {
        const double* s = A.signals + v * A.m;
        for (int i = tid; i < n; i += n_threads) {
            double p = 0.0;
            double x = 0.0;
            for (int r = 0; r < A.m; ++r) {
                const double sr = s[r];
                p = p + A.forward[static_cast<long>(r) * n + i] * sr;
                if (i < A.c0) {
                    x = x + A.initial_t[static_cast<long>(r) * A.c0 + i] * sr;
                }
            }
            proj[i] = p;
            f[i] = x;
        }
        __syncthreads();
        bool first = true;
        bool converged = A.max_iter == 0;
        int solves = 0;
        unsigned long long prev[NK];
#pragma unroll
        for (int t = 0; t < NK; ++t) {
            prev[t] = 0ull;
        }
        for (int it = 0; it < A.max_iter; ++it) {
            double a[NK];
#pragma unroll
            for (int t = 0; t < NK; ++t) {
                a[t] = 0.0;
            }
            for (int j = 0; j < n; ++j) {
                const double fj = f[j];
#pragma unroll
                for (int t = 0; t < NK; ++t) {
                    const int d = l + WAVE * t;
                    if (d < A.k) {
                        a[t] = a[t] + A.constraint_t[static_cast<long>(j) * A.k + d] * fj;
                    }
                }
            }
            bool same = true;
            int n_c = 0;
            unsigned long long mask[NK];
#pragma unroll
            for (int t = 0; t < NK; ++t) {
                const int d = l + WAVE * t;
                const bool on = d < A.k && a[t] < A.threshold;
                mask[t] = ballot(on);
                same = same && mask[t] == prev[t];
                if (on && w == 0) {
                    list[n_c + __popcll(mask[t] & ((1ull << l) - 1ull))] = d;
                }
                n_c += __popcll(mask[t]);
            }
            if (!first && same) {
                converged = true;
                break;
            }
            first = false;
            __syncthreads();

            for (int j = w; j < n; j += n_waves) {
                const long cj = column(j);
                for (int i = j + l; i < n; i += WAVE) {
                    double acc = A.gram[static_cast<long>(i) * n + j];
                    for (int q = 0; q < n_c; ++q) {
                        const double* cd = A.constraint + static_cast<long>(list[q]) * n;
                        acc = acc + cd[i] * cd[j];
                    }
                    sys[cj + (i - j)] = acc;
                }
            }
            __syncthreads();

            for (int j = 0; j < n; ++j) {
                const long cj = column(j);
                const double ljj = sqrt(sys[cj]);
                __syncthreads();
                for (int i = j + 1 + tid; i < n; i += n_threads) {
                    sys[cj + (i - j)] = sys[cj + (i - j)] / ljj;
                }
                if (tid == 0) {
                    sys[cj] = ljj;
                }
                __syncthreads();
                for (int q = j + 1 + w; q < n; q += n_waves) {
                    const double lqj = sys[cj + (q - j)];
                    const long cq = column(q);
                    for (int i = q + l; i < n; i += WAVE) {
                        sys[cq + (i - q)] = sys[cq + (i - q)] - sys[cj + (i - j)] * lqj;
                    }
                }
                __syncthreads();
            }

            if (w == 0) {
                // L y = F^T s, column-oriented.
                for (int i = l; i < n; i += WAVE) {
                    b[i] = proj[i];
                }
                wave_lds_sync();
                for (int j = 0; j < n; ++j) {
                    const long cj = column(j);
                    const double bj = b[j];
                    const double yj = bj != 0.0 ? bj / sys[cj] : bj;
                    for (int i = j + 1 + l; i < n; i += WAVE) {
                        b[i] = b[i] - sys[cj + (i - j)] * yj;
                    }
                    if (l == 0) {
                        b[j] = yj;
                    }
                    wave_lds_sync();
                }
                // L^T f = y, column-oriented from the last position.
                for (int j = n - 1; j >= 0; --j) {
                    const double yj = b[j];
                    const double xj = yj != 0.0 ? yj / sys[column(j)] : yj;
                    for (int i = l; i < j; i += WAVE) {
                        b[i] = b[i] - sys[column(i) + (j - i)] * xj;
                    }
                    if (l == 0) {
                        b[j] = xj;
                    }
                    wave_lds_sync();
                }
                for (int i = l; i < n; i += WAVE) {
                    f[i] = b[i];
                }
            }
            __syncthreads();
            solves += 1;
#pragma unroll
            for (int t = 0; t < NK; ++t) {
                prev[t] = mask[t];
            }
        }
        for (int i = tid; i < n; i += n_threads) {
            A.fods[v * n + i] = f[i];
        }
        if (tid == 0) {
            A.converged[v] = converged ? 1 : 0;
            if (A.solves) {
                A.solves[v] = solves;
            }
        }
        // The next voxel overwrites the vectors, the list and the triangle.
        __syncthreads();
    }
}

} // namespace lsd
