// Damped Richardson-Lucy of a list of voxels, one voxel per lane (cpp/src/sd/drl.cpp:39-99,
// DrlSolver::solve_block in cpp/src/sd/solver.cpp:9-12).
//
// Per voxel s (the signal, or in a multi-tissue fit the anisotropic signal
// max(w s - W_iso p, 0) of multi_tissue.cpp:20-24): the damping strength max(1 - 4 sd(s), 0)
// with the sample standard deviation, the stopping threshold 1e-4 max(s), K^T s, then from
// f = init, up to n_iter times: p = f K^T K and per direction
//   f_j <- max(f_j (1 + (1 - strength (1 - f_j^8 / (f_j^8 + nu^8))) (ks_j - p_j) / (p_j + 1e-16)),
//              1e-6),
// stopping once the FOD changed in total by no more than the threshold.  Every sum (the mean and
// sum of squares of sd, each K^T s and f K^T K entry, the change) runs in index order, one lane
// per voxel, the K^T s and f K^T K entries by fused multiply-adds as the oracle accumulates them,
// so the result is the oracle's to the bit.
//
// The per-voxel vectors live in HBM by list position, position-minor ((index, ld) layouts), so
// a wave's 64 lanes read and write 512 contiguous bytes per index; the design rows are
// wave-uniform scalar operands read through the constant address space.  The products f K^T K
// and K^T s run in passes of NCB directions held in registers.
#pragma once

#include "kernels.hpp"

namespace lsd {

constexpr double DRL_FOD_FLOOR = 1e-6;           // drl.cpp:13-14
constexpr double DRL_DENOMINATOR_OFFSET = 1e-16; // drl.cpp:16-17
constexpr double DRL_CONVERGENCE_SHARE = 1e-4;   // drl.cpp:19-21
constexpr int DRL_COLUMNS = 32;                  // directions per pass of the products

struct DrlArgs {
    const double* signals;    // (n_vox, m)
    int m;
    int multi;                // 1: the anisotropic signal of iso; 0: the signal itself
    const double* weights;    // (m), multi only
    const double* w_iso;      // (m, ci), multi only
    const double* iso;        // (n_vox, ci), multi only
    int ci;
    const double* kernel;     // K (m, c), then DRL_COLUMNS readable doubles
    const double* gram;       // K^T K (c, c), then DRL_COLUMNS readable doubles
    int c;
    double nu8;               // nu^8
    long n_iter;
    double init;
    const int* list;          // voxels by list position, or null for all
    const int* list_count;
    long n_vox;
    long ld;                  // leading dimension of the position buffers (>= the list size)
    double* a;                // (m, ld) the signal by position
    double* ks;               // (c, ld) K^T s
    double* f;                // (c, ld) the FOD
    double* prev;             // (c, ld) the FOD of the previous iteration
    double* pred;             // (c, ld) f K^T K
    double* x_out;            // (n_vox, c)
    int* status;              // (n_vox): ST_NOT_CONVERGED OR-ed in when the limit stopped it
    unsigned long long* work; // profiled launches: the solves' FP64 operations
};

inline bool attach_work(DrlArgs& a, unsigned long long* slot) {
    a.work = slot;
    return slot != nullptr;
}

__device__ __forceinline__ double power_8(double x) { // common/maths.hpp:26-30
    const double x2 = x * x;
    const double x4 = x2 * x2;
    return x4 * x4;
}

template <int NCB>
__global__ void __launch_bounds__(WAVE * WAVES_PER_BLOCK) drl_kernel(DrlArgs A) {
    __shared__ double tile[WAVES_PER_BLOCK][CORR_ROWS][WAVE + 1];
    using ConstRow = const __attribute__((address_space(4))) double*;
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / WAVE);
    const int l = lane_id();
    const long count = A.list ? static_cast<long>(*A.list_count) : A.n_vox;
    const long first = (static_cast<long>(blockIdx.x) * WAVES_PER_BLOCK + w) * WAVE;
    const long stride = static_cast<long>(gridDim.x) * WAVES_PER_BLOCK * WAVE;
    const long ld = A.ld;
    for (long base = first; base < count; base += stride) {
        const long p = base + l;
        const bool on = p < count;
        const long v = on ? (A.list ? A.list[p] : p) : 0;
        double iso[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            iso[q] = (on && A.multi && q < A.ci) ? A.iso[v * A.ci + q] : 0.0;
        }
        // The signal by position (staged through LDS with coalesced loads), its sum and its
        // largest non-NaN value.
        double total = 0.0;
        double largest = __builtin_nan("");
        for (int r0 = 0; r0 < A.m; r0 += CORR_ROWS) {
            const int rows = min(CORR_ROWS, A.m - r0);
#pragma unroll 4
            for (int k = 0; k < CORR_ROWS; ++k) {
                const int e = k * WAVE + l;
                const int j = e / CORR_ROWS;
                const int rr = e % CORR_ROWS;
                const long vj = __shfl(v, j);
                const bool load = base + j < count && rr < rows;
                tile[w][rr][j] = load ? A.signals[vj * A.m + r0 + rr] : 0.0;
            }
            wave_lds_sync();
            for (int rr = 0; rr < rows; ++rr) {
                const int r = r0 + rr;
                const double sr = tile[w][rr][l];
                double ar = sr;
                if (A.multi) {
                    const double ws = sr * A.weights[r];
                    double wp = 0.0;
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        if (q < A.ci) {
                            wp = wp + A.w_iso[static_cast<long>(r) * A.ci + q] * iso[q];
                        }
                    }
                    ar = fmax(ws - wp, 0.0);
                }
                if (on) {
                    A.a[r * ld + p] = ar;
                }
                total = total + ar;
                largest = (!isnan(ar) && (isnan(largest) || ar > largest)) ? ar : largest;
            }
            wave_lds_sync();
        }
        const double mean = total / static_cast<double>(A.m);
        double squares = 0.0;
        for (int r = 0; r < A.m; ++r) {
            const double d = (on ? A.a[r * ld + p] : 0.0) - mean;
            squares = squares + d * d;
        }
        const double sd = sqrt(squares / static_cast<double>(A.m - 1));
        const double strength = fmax(1.0 - 4.0 * sd, 0.0);
        const double threshold = DRL_CONVERGENCE_SHARE * largest;
        // K^T s.  Directions past c read the next row (or the padding) and are not stored.
        for (int j0 = 0; j0 < A.c; j0 += NCB) {
            double acc[NCB];
#pragma unroll
            for (int jj = 0; jj < NCB; ++jj) {
                acc[jj] = 0.0;
            }
            for (int r = 0; r < A.m; ++r) {
                const double ar = on ? A.a[r * ld + p] : 0.0;
                const ConstRow krow = (ConstRow)(A.kernel + static_cast<long>(r) * A.c + j0);
#pragma unroll
                for (int jj = 0; jj < NCB; ++jj) {
                    acc[jj] = __fma_rn(ar, krow[jj], acc[jj]);
                }
            }
            if (on) {
#pragma unroll
                for (int jj = 0; jj < NCB; ++jj) {
                    if (j0 + jj < A.c) {
                        A.ks[(j0 + jj) * ld + p] = acc[jj];
                    }
                }
            }
        }
        if (on) {
            for (int j = 0; j < A.c; ++j) {
                A.f[j * ld + p] = A.init;
                A.prev[j * ld + p] = A.init;
            }
        }
        bool active = on;
        bool converged = false;
        long ran = 0;
        for (long iter = 0; iter < A.n_iter; ++iter) {
            if (ballot(active) == 0ull) {
                break;
            }
            // predicted = f K^T K, each entry summed over k in order.
            for (int j0 = 0; j0 < A.c; j0 += NCB) {
                double acc[NCB];
#pragma unroll
                for (int jj = 0; jj < NCB; ++jj) {
                    acc[jj] = 0.0;
                }
                int k = 0;
                for (; k + 4 <= A.c; k += 4) {
                    double fq[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        fq[u] = active ? A.f[(k + u) * ld + p] : 0.0;
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const ConstRow grow =
                            (ConstRow)(A.gram + static_cast<long>(k + u) * A.c + j0);
#pragma unroll
                        for (int jj = 0; jj < NCB; ++jj) {
                            acc[jj] = __fma_rn(fq[u], grow[jj], acc[jj]);
                        }
                    }
                }
                for (; k < A.c; ++k) {
                    const double fk = active ? A.f[k * ld + p] : 0.0;
                    const ConstRow grow = (ConstRow)(A.gram + static_cast<long>(k) * A.c + j0);
#pragma unroll
                    for (int jj = 0; jj < NCB; ++jj) {
                        acc[jj] = __fma_rn(fk, grow[jj], acc[jj]);
                    }
                }
                if (active) {
#pragma unroll
                    for (int jj = 0; jj < NCB; ++jj) {
                        if (j0 + jj < A.c) {
                            A.pred[(j0 + jj) * ld + p] = acc[jj];
                        }
                    }
                }
            }
            if (active) {
                double change = 0.0;
                for (int j = 0; j < A.c; ++j) {
                    const double amp = A.f[j * ld + p];
                    const double amp8 = power_8(amp);
                    const double damping = 1.0 - strength * (1.0 - amp8 / (amp8 + A.nu8));
                    const double pj = A.pred[j * ld + p];
                    const double ratio = (A.ks[j * ld + p] - pj) / (pj + DRL_DENOMINATOR_OFFSET);
                    const double updated = fmax(amp * (1.0 + damping * ratio), DRL_FOD_FLOOR);
                    A.f[j * ld + p] = updated;
                    change = change + fabs(updated - A.prev[j * ld + p]);
                    A.prev[j * ld + p] = updated;
                }
                ran = iter + 1;
                if (change <= threshold) {
                    converged = true;
                    active = false;
                }
            }
        }
        if (on) {
            for (int j = 0; j < A.c; ++j) {
                A.x_out[v * A.c + j] = A.f[j * ld + p];
            }
            if (!converged && A.n_iter > 0) {
                A.status[v] |= ST_NOT_CONVERGED;
            }
            if (A.work != nullptr) {
                // Per voxel: the signal statistics (4m), K^T s (2mc), and per iteration f K^T K
                // (2c^2) and the update (16c: power, damping, ratio, step, change).
                const unsigned long long c = static_cast<unsigned long long>(A.c);
                const unsigned long long mm = static_cast<unsigned long long>(A.m);
                atomicAdd(A.work, 4ull * mm + 2ull * mm * c +
                                      static_cast<unsigned long long>(ran) * (2ull * c * c + 16ull * c));
            }
        }
    }
}

} // namespace lsd
