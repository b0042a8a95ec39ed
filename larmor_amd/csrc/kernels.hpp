// Kernels of liblarmor_sd.so and the launch helpers shared by its translation units.
//
// The C-ABI replaces the compiled core that larmor's Python fit API calls for this path
// (cpp/src/python/sd.cpp: FodSolver.l2nnls, fit_single_tissue, DeconvKernel,
// fit_multi_tissue) with device-resident solver and kernel objects and stream-ordered entry
// points on device buffers.  Declarations and per-function reference citations are in
// include/larmor_sd.h.
//
// Execution model: one 64-lane wavefront per voxel, voxels strided over a grid sized to the
// chip.  A multi-tissue fit runs two kernels: the low-resolution loop (FOD NNLS <-> fraction
// NNLS rounds, giving each voxel's isotropic fractions) and the final high-resolution FOD and
// fractions.  A voxel whose active set outgrows the LDS Cholesky factor is appended to a device
// list and solved again by a rerun kernel that keeps its factor in global memory; the rerun
// kernels read the list length on the device, so a fit is a fixed sequence of launches with no
// host synchronisation.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/larmor_sd.h"
#include "fractions_wave.hpp"
#include "launch.hpp"
#include "ridge_wave.hpp"

namespace LSD_NS {

#ifdef LSD_FP32
using namespace lsd; // the shared launch helpers
#endif

constexpr int NQ = 6;             // most fraction columns (groups + isotropic tissues)
// The fraction and correlation kernels are also built for at most 4 columns and 2 blocks (the
// usual tissue models), which keeps their column arrays in fewer registers.
constexpr int NQ_SMALL = 4;
constexpr int GROUPS_SMALL = 2;
inline bool fraction_columns_small(int n_blocks, int ci) {
    return n_blocks + ci <= NQ_SMALL && 1 + ci <= NQ_SMALL && n_blocks <= GROUPS_SMALL;
}
// LDS factor capacity (active columns) of the fast FOD kernels: 20 for the low-resolution solve
// (its active sets stay below 20 in all but about 0.1% of voxels, which take the rerun kernel),
// so that 8 waves per SIMD fit in LDS; 32 for the others.
constexpr int kcap_lds(int ncpl) { return ncpl == 1 ? 20 : 32; }
constexpr int RERUN_WAVES = 1024; // waves of a rerun launch (4 per CU; the overflow list is a few percent)
constexpr int RERUN_WAVES_ALL = 512; // waves when every voxel needs the global-memory kernel
constexpr real FRACTION_SUM_OFFSET = real(1e-16);

// Per-wave slab of the FOD kernels: the Cholesky factor (stride kcap + 1, plus one wave of slack
// for the uniform reads), then its diagonal and the diagonal's reciprocals (2 kcap).  An even
// count keeps every slab 16-byte aligned.
constexpr int lds_factor_doubles(int kcap, int /*rpl*/) {
    return (kcap + 1) * kcap + WAVE + 2 * kcap;
}

// ---------------------------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------------------------

// corr_c = sum over rows, in order, of K(r, c) a_r by fused multiply-adds, minus l1 (nnls.cpp:302).
template <int NCPL, int RPL>
__device__ __forceinline__ void correlations(const RidgeDesignDev& D, const RowVec<RPL>& a, real l1,
                             real (&corr)[NCPL]) {
    const int l = lane_id();
#pragma unroll
    for (int t = 0; t < NCPL; ++t) {
        corr[t] = real(0.0);
    }
#pragma unroll
    for (int ts = 0; ts < RPL; ++ts) {
        const int rows_here = D.m - WAVE * ts;
        const int count = rows_here < WAVE ? rows_here : WAVE;
        for (int ll = 0; ll < count; ++ll) {
            const int r = WAVE * ts + ll;
            const real ar = readlane(a.v[ts], ll);
            const real* krow = D.kernel + static_cast<long>(r) * D.n;
#pragma unroll
            for (int t = 0; t < NCPL; ++t) {
                const int c = l + WAVE * t;
                if (c < D.n) {
                    corr[t] = fma(krow[c], ar, corr[t]);
                }
            }
        }
    }
#pragma unroll
    for (int t = 0; t < NCPL; ++t) {
        corr[t] = corr[t] - l1;
    }
}

template <int RPL>
__device__ __forceinline__ void load_signal(const real* signals, long v, int m, RowVec<RPL>& s) {
    const int l = lane_id();
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
        const int r = l + WAVE * t;
        s.v[t] = r < m ? signals[v * m + r] : real(0.0);
    }
}

// max(w s - W_iso p, 0) per row (multi_tissue.cpp:20-24), p the isotropic fractions.
template <int RPL>
__device__ __forceinline__ void anisotropic_signal(const RowVec<RPL>& s, const real* weights,
                                   const real* w_iso, int m, int ci, const real (&p)[NQ],
                                   RowVec<RPL>& a) {
    const int l = lane_id();
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
        const int r = l + WAVE * t;
        real out = real(0.0);
        if (r < m) {
            const real ws = s.v[t] * weights[r];
            real acc = real(0.0);
#pragma unroll
            for (int tt = 0; tt < NQ; ++tt) {
                if (tt < ci) {
                    acc = acc + w_iso[static_cast<long>(r) * ci + tt] * p[tt];
                }
            }
            out = fmax(ws - acc, real(0.0));
        }
        a.v[t] = out;
    }
}

struct SolveScratch {
    real* base;     // this launch's slab array
    long per_wave;    // doubles per wave
};

// ---------------------------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------------------------

struct VoxelSet {
    long n_vox;               // all voxels, when list is null
    const int* list;          // voxel indices to process, or null
    const int* list_count;    // device count of list
    int* overflow_list;       // where voxels that overflow are recorded (null: none expected)
    int* overflow_count;
};

__device__ __forceinline__ long voxel_count(const VoxelSet& vs) {
    return vs.list ? static_cast<long>(*vs.list_count) : vs.n_vox;
}

__device__ __forceinline__ void record_overflow(const VoxelSet& vs, long v) {
    if (lane_id() == 0 && vs.overflow_count != nullptr) {
        const int at = atomicAdd(vs.overflow_count, 1);
        vs.overflow_list[at] = static_cast<int>(v);
    }
}

// ---------------------------------------------------------------------------------------------
// Batched launches: the voxels of up to MAX_JOBS volumes, each with its own measurements and
// design, in one launch per step of the fit.  The batch numbers its voxels job after job (job j's
// voxels are first_j .. first_j + n_vox_j - 1); per-voxel state is indexed by that number, and a
// job's list of voxels for a round lives in its own region (first_j ..) of the list buffer, with
// its own count.  A wave finds the job of a work item with one ballot: lane j holds job j's start
// in the launch's job-major item space (a wave prefix sum of the jobs' item counts).
// ---------------------------------------------------------------------------------------------
constexpr int MAX_JOBS = WAVE;

// One job's view of a stage (the low-resolution rounds or the high-resolution solve).
struct JobStage {
    long first;            // its first voxel in the batch numbering
    long n_vox;
    const real* signals;   // (n_vox, m)
    int m;
    const real* weights;   // (m)
    const real* w_iso;     // (m, ci)
    RidgeDesignDev design; // the stage's ridge design
    real tol;
    FractionKernelDev frac;
};

struct JobTable {
    const JobStage* jobs; // device array, n_jobs entries
    int n_jobs;
    const int* counts;    // per-job list counts of this launch (null: every voxel of every job)
    int by_voxel;         // 1: the launch walks a flat list of batch voxel numbers (a rerun)
};

// Inclusive wave prefix sum (lane l: the sum over lanes 0..l).
__device__ __forceinline__ int wave_inclusive_sum(int x) {
    const int l = lane_id();
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const int y = __shfl_up(x, o);
        x = l >= o ? x + y : x;
    }
    return x;
}

// The launch's item space: each lane's job start (exclusive prefix of the jobs' counts, in units
// of `unit` items rounded up) and the total.
struct JobSpace {
    int start; // lane j: job j's first unit
    int total; // units over all jobs (uniform)
};

__device__ __forceinline__ JobSpace job_space(const JobTable& T, int unit) {
    const int l = lane_id();
    int count = 0;
    if (l < T.n_jobs) {
        count = T.counts != nullptr ? T.counts[l] : static_cast<int>(T.jobs[l].n_vox);
    }
    const int units = (count + unit - 1) / unit;
    const int incl = wave_inclusive_sum(units);
    JobSpace js;
    js.start = incl - units;
    js.total = __builtin_amdgcn_readfirstlane(__shfl(incl, WAVE - 1));
    return js;
}

// The job holding unit u (uniform) and u's offset in it.
__device__ __forceinline__ int job_of_unit(const JobTable& T, const JobSpace& js, int u, int& local) {
    const unsigned long long before = ballot(lane_id() < T.n_jobs && js.start <= u);
    const int j = __builtin_amdgcn_readfirstlane(__popcll(before) - 1);
    local = __builtin_amdgcn_readfirstlane(u - __shfl(js.start, j));
    return j;
}

// The job of batch voxel v (uniform).
__device__ __forceinline__ int job_of_voxel(const JobTable& T, long v) {
    const int l = lane_id();
    const bool before = l < T.n_jobs && T.jobs[l < T.n_jobs ? l : 0].first <= v;
    return __builtin_amdgcn_readfirstlane(__popcll(ballot(before)) - 1);
}

// The FOD solve of a list of voxels: the signal (single-tissue) or the anisotropic signal
// max(w s - W_iso p, 0) of each voxel's isotropic fractions p (multi-tissue,
// multi_tissue.cpp:20-24), its correlations and the ridge NNLS (solver.cpp:14-21,
// nnls.cpp:294-350); the amplitudes go to x_out (n_vox, n).
struct FodArgs {
    const real* signals;    // (n_vox, m)
    int m;
    int multi;                // 1: anisotropic signal of iso; 0: the signal itself
    const real* weights;    // (m), multi only
    const real* w_iso;      // (m, ci), multi only
    const real* iso;        // (n_vox, ci), multi only
    int ci;
    RidgeDesignDev design;
    real tol;
    real* x_out;            // (n_vox, n)
    int* status;              // per voxel, OR-ed
    VoxelSet vs;
    SolveScratch scratch;
    const real* corr;       // (list position, n) precomputed correlations, or null
    unsigned long long* work; // profiled launches: the completed solves' FP64 operations
    unsigned long long* active_mask; // (n_vox) warm-start supports in and out (WARM kernels)
    JobTable jobs;            // BATCHED kernels: the batch's jobs (vs.list: the lists' base)
};

// The fast low-resolution instantiation is asked for 8 waves per SIMD, what its LDS factor
// (4.2 KB per wave, 20 columns) allows: the register budget that leaves costs a few spilled
// registers and still beats 7 waves with a 24-column factor.  The capacities a fit may be tuned
// to (LR_CAPACITIES) are built each for the waves its factor allows.
#ifndef LSD_FOD_LR_WAVES
#define LSD_FOD_LR_WAVES 8
#endif
constexpr int lr_waves_for(int kcap) { return kcap <= 20 ? LSD_FOD_LR_WAVES : (kcap <= 24 ? 7 : 6); }
// More than 8 row slots (above 512 measurements) hold the signal in 12 or 16 registers a lane,
// so those solves take two waves per SIMD and the registers that gives.
constexpr int fod_waves_per_eu(int ncpl, int kcap, bool global_l, int rpl) {
    return (!global_l && ncpl == 1) ? (rpl > 8 ? 2 : lr_waves_for(kcap)) : 1;
}

// WARM (the mixed-precision form's low-resolution solves, NCPL 1): each voxel's solve starts from
// the support its previous solve left in active_mask and leaves its own there.
// BATCHED: the voxels of a batch's jobs (A.jobs), each with its own signals, measurements and
// design; precomputed correlations are indexed by batch voxel number.
template <int NCPL, int RPL, int KCAP, bool GLOBAL_L, bool WARM = false, bool BATCHED = false>
__global__ void __launch_bounds__(WAVE * WAVES_PER_BLOCK)
    __attribute__((amdgpu_waves_per_eu(fod_waves_per_eu(NCPL, KCAP, GLOBAL_L, RPL))))
    fod_kernel(FodArgs A) {
    constexpr int RPLQ = (RPL * WAVE + KCAP + WAVE - 1) / WAVE;
    extern __shared__ real smem[];
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / WAVE);
    const long gw = static_cast<long>(blockIdx.x) * (blockDim.x / WAVE) + w;
    const long n_waves = static_cast<long>(gridDim.x) * (blockDim.x / WAVE);
    real* slab = A.scratch.base + gw * A.scratch.per_wave;
    real* qrs = slab;
    real* L = GLOBAL_L ? slab + static_cast<long>(RPLQ * WAVE) * (KCAP + 1)
                         : smem + w * lds_factor_doubles(KCAP, RPL);
    long count = 0;
    JobSpace js{};
    if constexpr (BATCHED) {
        if (A.jobs.by_voxel) {
            count = *A.vs.list_count;
        } else {
            js = job_space(A.jobs, 1);
            count = js.total;
        }
    } else {
        count = voxel_count(A.vs);
    }
    const int l = lane_id();
    unsigned long long wave_work = 0;
    // One voxel's solve with its design (the launch's, or its job's, read where it is: the
    // kernel arguments or the job table) and its signal row.
    auto solve_voxel = [&](long it, long v, long row, const RidgeDesignDev& design, real tol,
                           const real* signals, int m, const real* weights, const real* w_iso) {
        RowVec<RPL> s;
        real corr[NCPL];
        if (!GLOBAL_L && A.corr != nullptr) {
            // The correlation pass computed them (the fast kernels carry no QR fits, the only
            // other use of the signal).
            s.fill(real(0.0));
            const long at = BATCHED ? v : it;
#pragma unroll
            for (int t = 0; t < NCPL; ++t) {
                const int c = l + WAVE * t;
                corr[t] = c < design.n ? A.corr[at * design.n + c] : real(0.0);
            }
        } else {
            load_signal<RPL>(signals, row, m, s);
            if (A.multi) {
                real p[NQ];
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    p[q] = q < A.ci ? A.iso[v * A.ci + q] : real(0.0);
                }
                RowVec<RPL> a;
                anisotropic_signal<RPL>(s, weights, w_iso, m, A.ci, p, a);
                s = a;
            }
            correlations<NCPL, RPL>(design, s, real(0.0), corr);
        }
        RidgeNnlsWave<NCPL, KCAP, RPLQ, GLOBAL_L> solver(design, tol, 3 * design.n, real(0.0), L,
                                                         qrs);
        const unsigned long long warm = WARM ? A.active_mask[v] : 0ull;
        const int st = solver.template run<RPL>(corr, s, warm);
        if (st & ST_OVERFLOW) {
            record_overflow(A.vs, v);
            return;
        }
        if constexpr (WARM) {
            const unsigned long long support = solver.support();
            if (l == 0) {
                A.active_mask[v] = support;
            }
        }
        wave_work += solver.work;
#pragma unroll
        for (int t = 0; t < NCPL; ++t) {
            const int c = l + WAVE * t;
            if (c < design.n) {
                A.x_out[v * design.n + c] = solver.x[t];
            }
        }
        if (l == 0 && st != 0) {
            A.status[v] |= st;
        }
    };
    for (long it = gw; it < count; it += n_waves) {
        if constexpr (BATCHED) {
            long v = 0;
            int j = 0;
            if (A.jobs.by_voxel) {
                v = __builtin_amdgcn_readfirstlane(A.vs.list[it]);
                j = job_of_voxel(A.jobs, v);
            } else {
                int local = 0;
                j = job_of_unit(A.jobs, js, static_cast<int>(it), local);
                const long first = A.jobs.jobs[j].first;
                v = A.vs.list ? __builtin_amdgcn_readfirstlane(A.vs.list[first + local])
                              : first + local;
            }
            const JobStage& J = A.jobs.jobs[j];
            solve_voxel(it, v, v - J.first, J.design, J.tol, J.signals, J.m, J.weights, J.w_iso);
        } else {
            const long v = A.vs.list ? A.vs.list[it] : it;
            solve_voxel(it, v, v, A.design, A.tol, A.signals, A.m, A.weights, A.w_iso);
        }
    }
    if (A.work != nullptr && l == 0 && wave_work != 0) {
        atomicAdd(A.work, wave_work);
    }
}

// The correlations K^T a of a list of voxels, a the signal (single-tissue) or the anisotropic
// signal max(w s - W_iso p, 0) (multi_tissue.cpp:20-24), one voxel per lane: the design rows are
// wave-uniform, so a wave's 64 voxels share every instruction and each K entry is one LDS
// broadcast read.  Each correlation is the same in-order sum of K(r, c) a_r, rounded
// the same way, as the FOD kernels' own (correlations() above).  Columns go in passes of NCB.
struct CorrArgs {
    const real* signals;    // (n_vox, m)
    int m;
    int multi;
    const real* weights;    // (m), multi only
    const real* w_iso;      // (m, ci), multi only
    const real* iso;        // (n_vox, ci), multi only
    int ci;
    const real* kernel;     // (m, n), then at least NCB readable doubles
    int n;
    const int* list;          // voxels by list position, or null for all
    const int* list_count;
    long n_vox;
    real* corr;             // (list position, n); BATCHED: (batch voxel, n)
    JobTable jobs;            // BATCHED: the batch's jobs (list: the lists' base; n common)
};

constexpr int CORR_COLUMNS = 44; // columns per pass of the correlation kernel
constexpr int CORR_ROWS = 16;    // signal rows staged per LDS tile

// Each wave takes 64 consecutive list positions.  Their signals are staged through the wave's
// LDS tile CORR_ROWS rows at a time with coalesced loads (each load instruction covers 16 rows
// of 4 voxels) and read back one voxel per lane; the matching CORR_ROWS x NCB chunk of K is
// staged once per block, shared by its waves.  BATCHED: a block's 256 positions are those of one
// job (each job's list is cut into blocks of its own), whose signals, measurements and design it
// reads.
template <int NCB, bool BATCHED = false, int FQ = NQ>
__global__ void __launch_bounds__(WAVE * WAVES_PER_BLOCK) correlation_kernel(CorrArgs A) {
    __shared__ real tile[WAVES_PER_BLOCK][CORR_ROWS][WAVE + 1];
    __shared__ real ktile[CORR_ROWS][NCB];
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / WAVE);
    const int l = lane_id();
    constexpr long block_voxels = static_cast<long>(WAVES_PER_BLOCK) * WAVE;
    long all = 0;
    long n_units = 0;
    JobSpace js{};
    if constexpr (BATCHED) {
        js = job_space(A.jobs, static_cast<int>(block_voxels));
        n_units = js.total;
    } else {
        all = A.list ? static_cast<long>(*A.list_count) : A.n_vox;
        n_units = (all + block_voxels - 1) / block_voxels;
    }
    // The block's waves move through the column passes and row chunks together (the kernel
    // chunk in LDS is shared), so a wave past the list still takes part in every barrier.
    for (long u = blockIdx.x; u < n_units; u += gridDim.x) {
        long count = all;
        long block_base = u * block_voxels;
        const int* list = A.list;
        long first = 0;
        const real* signals = A.signals;
        const real* kernel = A.kernel;
        const real* weights = A.weights;
        const real* w_iso = A.w_iso;
        int m = A.m;
        if constexpr (BATCHED) {
            int local_block = 0;
            const int j = job_of_unit(A.jobs, js, static_cast<int>(u), local_block);
            const JobStage& J = A.jobs.jobs[j];
            first = J.first;
            count = A.jobs.counts != nullptr ? A.jobs.counts[j] : J.n_vox;
            block_base = local_block * block_voxels;
            list = A.list != nullptr ? A.list + first : nullptr;
            signals = J.signals;
            kernel = J.design.kernel;
            weights = J.weights;
            w_iso = J.w_iso;
            m = J.m;
        }
        const long base = block_base + static_cast<long>(w) * WAVE;
        const long it = base + l;
        const bool on = it < count;
        const long v = on ? (list ? list[it] : first + it) : first;
        const long row = v - first;
        real p[FQ];
#pragma unroll
        for (int q = 0; q < FQ; ++q) {
            p[q] = (on && A.multi && q < A.ci) ? A.iso[v * A.ci + q] : real(0.0);
        }
        for (int c0 = 0; c0 < A.n; c0 += NCB) {
            real acc[NCB];
#pragma unroll
            for (int j = 0; j < NCB; ++j) {
                acc[j] = real(0.0);
            }
            for (int r0 = 0; r0 < m; r0 += CORR_ROWS) {
                const int rows = min(CORR_ROWS, m - r0);
                // The kernel chunk K[r0 .. r0 + rows, c0 .. c0 + NCB), loaded once per block.
                // Columns past n read the next row (or the padding after the last) and are
                // not stored.
                for (int e = threadIdx.x; e < CORR_ROWS * NCB; e += blockDim.x) {
                    const int rr = e / NCB;
                    const int j = e % NCB;
                    ktile[rr][j] = rr < rows ? kernel[static_cast<long>(r0 + rr) * A.n + c0 + j]
                                             : real(0.0);
                }
#pragma unroll 4
                for (int k = 0; k < CORR_ROWS; ++k) {
                    const int e = k * WAVE + l;
                    const int j = e / CORR_ROWS;
                    const int rr = e % CORR_ROWS;
                    const long vj = __shfl(row, j);
                    const bool load = base + j < count && rr < rows;
                    tile[w][rr][j] = load ? signals[vj * m + r0 + rr] : real(0.0);
                }
                __syncthreads();
                for (int rr = 0; rr < rows; ++rr) {
                    const int r = r0 + rr;
                    const real sr = tile[w][rr][l];
                    real ar = sr;
                    if (A.multi) {
                        const real ws = sr * weights[r];
                        real wp = real(0.0);
#pragma unroll
                        for (int q = 0; q < FQ; ++q) {
                            if (q < A.ci) {
                                wp = wp + w_iso[static_cast<long>(r) * A.ci + q] * p[q];
                            }
                        }
                        ar = fmax(ws - wp, real(0.0));
                    }
                    // Every lane reads the same kernel entries: LDS broadcasts.
#pragma unroll
                    for (int j = 0; j < NCB; ++j) {
                        acc[j] = fma(ktile[rr][j], ar, acc[j]);
                    }
                }
                __syncthreads();
            }
            if (on) {
#pragma unroll
                for (int j = 0; j < NCB; ++j) {
                    if (c0 + j < A.n) {
                        A.corr[(BATCHED ? v : it) * A.n + c0 + j] = acc[j];
                    }
                }
            }
        }
    }
}

// The tissue fractions of a list of voxels from their FOD amplitudes (fractions.cpp:45-76).
// A round of the low-resolution loop (multi_tissue.cpp:59-78) fits the lumped fractions with the
// default tolerance, keeps the isotropic ones, counts the round and appends the voxel to the
// next round's list unless no fraction moved by fraction_tol; the final fit (multi_tissue.cpp:
// 92-99) fits the per-block fractions with final_tol and normalises them to sum to 1.
struct FractionArgs {
    const real* signals;    // (n_vox, m), unweighted
    int m;
    const real* x;          // (n_vox, ca) FOD amplitudes
    FractionKernelDev kernel;
    int final_fit;
    real final_tol;
    real fraction_tol;
    real* iso;              // (n_vox, ci)     loop rounds
    real* prev;             // (n_vox, 1 + ci) loop rounds
    int* rounds;              // (n_vox)         loop rounds
    int* next_list;           // loop rounds: the voxels of the next round
    int* next_count;
    real* fractions;        // (n_vox, n_blocks + ci), final fit
    int* status;
    const int* list;          // voxels to fit, or null for all
    const int* list_count;
    long n_vox;
    JobTable jobs;            // BATCHED: the batch's jobs (list, next_list: the lists' base;
                              // next_count: per-job counts; kernel: the common column counts)
};

// The low-resolution fraction kernel is asked for 5 waves per SIMD with 4 row slots.
// With more row slots the per-wave QR slab and the block's isotropic columns cap the resident
// waves at 3 per SIMD (6 slots) and 2 (8 slots), so asking for more would only spill registers.
// The normal-equation fits hold no slab; 4 waves per SIMD leave them registers without spills.
#ifndef LSD_FRACTION_LR_WAVES
#define LSD_FRACTION_LR_WAVES 5
#endif
constexpr int fraction_waves_per_eu(int ncpl, int rpl, bool normal) {
    return ncpl != 1 ? 1 : normal ? 4 : (rpl <= 4 ? LSD_FRACTION_LR_WAVES : (rpl <= 6 ? 3 : 2));
}

// NORMAL: the mixed-precision form's loop fits through the normal equations (no QR slab).
// BATCHED: the voxels of a batch's jobs, each wave keeping its own copy of its current job's
// isotropic columns and appending to the job's next list.
template <int NCPL, int RPL, bool FINAL, bool NORMAL = false, bool BATCHED = false, int FQ = NQ>
__global__ void __launch_bounds__(WAVE * WAVES_PER_BLOCK)
    __attribute__((amdgpu_waves_per_eu(fraction_waves_per_eu(NCPL, RPL, NORMAL)))) fraction_kernel(FractionArgs A) {
    extern __shared__ real smem[];
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / WAVE);
    const long gw = static_cast<long>(blockIdx.x) * (blockDim.x / WAVE) + w;
    const long n_waves = static_cast<long>(gridDim.x) * (blockDim.x / WAVE);
    // The fit's columns: one lumped group (the loop) or one per block (the final fit), then the
    // isotropic tissues.
    const int qcols = (FINAL ? A.kernel.n_blocks : 1) + A.kernel.ci;
    const int slab = NORMAL ? 0 : qcols * RPL * WAVE;
    real* work = smem + w * slab;
    // The block's copy of the isotropic columns, after the waves' QR slabs (BATCHED: the wave's).
    real* iso_lds = smem + static_cast<int>(blockDim.x / WAVE) * slab +
                    (BATCHED ? w * A.kernel.ci * RPL * WAVE : 0);
    long count = 0;
    JobSpace js{};
    if constexpr (BATCHED) {
        js = job_space(A.jobs, 1);
        count = js.total;
    } else {
        for (int i = threadIdx.x; i < A.kernel.ci * RPL * WAVE; i += blockDim.x) {
            const int c = i / (RPL * WAVE);
            const int r = i % (RPL * WAVE);
            iso_lds[i] = r < A.m ? A.kernel.iso_t[static_cast<long>(c) * A.m + r] : real(0.0);
        }
        __syncthreads();
        count = A.list ? static_cast<long>(*A.list_count) : A.n_vox;
    }
    const int l = lane_id();
    const int ca = A.kernel.ca;
    const int ci = A.kernel.ci;
    // The loop's voxels for the next round, buffered per wave (one per lane) and appended to the
    // next list with one atomic per buffer instead of one per voxel: every wave of the launch
    // appends to the same counter.  (The list's order is arbitrary either way; no voxel's result
    // depends on it.)  BATCHED: the buffer holds one job's voxels and goes to that job's list.
    int pending = 0;
    int pending_voxel = 0;
    int pending_job = 0;
    auto flush_pending = [&]() {
        if (pending > 0) {
            int* next_count = A.next_count;
            int* next_list = A.next_list;
            if constexpr (BATCHED) {
                next_count += pending_job;
                next_list += A.jobs.jobs[pending_job].first;
            }
            int base = 0;
            if (l == 0) {
                base = atomicAdd(next_count, pending);
            }
            base = __builtin_amdgcn_readfirstlane(base);
            if (l < pending) {
                next_list[base + l] = pending_voxel;
            }
            pending = 0;
        }
    };
    int staged_job = -1;
    for (long it = gw; it < count; it += n_waves) {
        long v = 0;
        long row = 0;
        FractionKernelDev K = A.kernel;
        const real* signals = A.signals;
        int m = A.m;
        if constexpr (BATCHED) {
            int local = 0;
            const int j = job_of_unit(A.jobs, js, static_cast<int>(it), local);
            const JobStage& J = A.jobs.jobs[j];
            v = A.list ? __builtin_amdgcn_readfirstlane(A.list[J.first + local]) : J.first + local;
            row = v - J.first;
            K = J.frac;
            signals = J.signals;
            m = J.m;
            if (j != staged_job) {
                if (!FINAL && pending > 0) {
                    flush_pending();
                }
#pragma unroll
                for (int t = 0; t < RPL; ++t) {
                    const int r = l + WAVE * t;
                    for (int c = 0; c < ci; ++c) {
                        iso_lds[c * (RPL * WAVE) + r] =
                            r < m ? K.iso_t[static_cast<long>(c) * m + r] : real(0.0);
                    }
                }
                wave_lds_sync();
                staged_job = j;
            }
            pending_job = j;
        } else {
            v = A.list ? A.list[it] : it;
            row = v;
        }
        RowVec<RPL> s;
        load_signal<RPL>(signals, row, m, s);
        // The loop's previous fractions (lane q holds fraction q) and round count, loaded before
        // the fit so that their latency overlaps it.
        const int nlf = 1 + ci;
        real prev = real(0.0);
        int rounds = 0;
        if (!FINAL) {
            prev = l < nlf ? A.prev[v * nlf + l] : real(0.0);
            rounds = l == 0 ? A.rounds[v] : 0;
        }
        real fod[NCPL];
#pragma unroll
        for (int t = 0; t < NCPL; ++t) {
            const int c = l + WAVE * t;
            fod[t] = c < ca ? A.x[v * ca + c] : real(0.0);
        }
        real fr[FQ];
        // The final fit's per-block groups, or the loop's one lumped group.
        const bool ok = FINAL
                            ? fractions_wave<NCPL, RPL, FQ, (FQ == NQ_SMALL ? GROUPS_SMALL : MAX_GROUPS)>(
                                  K, fod, s, false, true, A.final_tol, fr, work, iso_lds)
                            : fractions_wave<NCPL, RPL, FQ, 1, NORMAL>(K, fod, s, true, false,
                                                                       real(0.0), fr, work, iso_lds);
        if (!ok && l == 0) {
            A.status[v] |= ST_NOT_CONVERGED;
        }
        if (FINAL) {
            const int nf = A.kernel.n_blocks + ci;
            real total = real(0.0);
#pragma unroll
            for (int q = 0; q < FQ; ++q) {
                if (q < nf) {
                    total = total + fr[q];
                }
            }
            total = total + FRACTION_SUM_OFFSET;
#pragma unroll
            for (int q = 0; q < FQ; ++q) {
                if (l == q && q < nf) {
                    A.fractions[v * nf + q] = fr[q] / total;
                }
            }
            continue;
        }
        real mine = real(0.0);
#pragma unroll
        for (int q = 0; q < FQ; ++q) {
            mine = l == q ? fr[q] : mine;
        }
        const bool converged = ballot(l < nlf && !(fabs(mine - prev) < A.fraction_tol)) == 0ull;
        if (l >= 1 && l < nlf) {
            A.iso[v * ci + (l - 1)] = mine;
        }
        if (!converged && l < nlf) {
            A.prev[v * nlf + l] = mine;
        }
        if (l == 0) {
            A.rounds[v] = rounds + 1;
        }
        if (!converged) {
            // Buffered: lane `pending` holds the voxel; a full buffer goes out at once.
            pending_voxel = l == pending ? static_cast<int>(v) : pending_voxel;
            pending += 1;
            if (pending == WAVE) {
                flush_pending();
            }
        }
    }
    if (!FINAL) {
        flush_pending();
    }
}

// Kernel families, for the launch helpers.  TUNABLE: the fast low-resolution launch (one column
// slot) takes the factor capacity tuned for its row class (lr_capacity) instead of kcap_lds.
template <int NCPL, int RPL, int KCAP, bool G>
struct fod_family {
    static constexpr auto kernel = fod_kernel<NCPL, RPL, KCAP, G>;
    static constexpr bool TUNABLE = true;
};
template <int NCPL, int RPL, int KCAP, bool G>
struct fod_warm_family {
    static constexpr auto kernel = fod_kernel<NCPL, RPL, KCAP, G, true>;
    static constexpr bool TUNABLE = false;
};
template <int NCPL, int RPL, int KCAP, bool G>
struct fod_batched_family {
    static constexpr auto kernel = fod_kernel<NCPL, RPL, KCAP, G, false, true>;
    static constexpr bool TUNABLE = true;
};

// The fast launch (LDS factor of KC columns) of a family.
template <template <int, int, int, bool> class Family, int NCPL, int RPL, int KC, class Args>
int launch_fast_at(const char* name, const Args& fast, int waves, hipStream_t stream) {
    const size_t lds = sizeof(real) * lds_factor_doubles(KC, RPL) * WAVES_PER_BLOCK;
    return launch_fill(name, Family<NCPL, RPL, KC, false>::kernel, fast, waves, lds, stream);
}

// The fast launch at the family's capacity: for a tunable one-slot solve the capacity tuned for
// the row class (16, 20, 24 or 32 columns; a voxel whose active set outgrows it finishes in the
// rerun with the same arithmetic, so every capacity gives the same bits), else kcap_lds.
template <template <int, int, int, bool> class Family, int NCPL, int RPL, class Args>
int launch_fast(const char* name, const Args& fast, int waves, hipStream_t stream) {
#if !defined(LSD_FP32)
    if constexpr (NCPL == 1 && Family<1, RPL, 20, false>::TUNABLE) {
        switch (lsd::lr_capacity(RPL)) {
        case 16:
            return launch_fast_at<Family, NCPL, RPL, 16>(name, fast, waves, stream);
        case 24:
            return launch_fast_at<Family, NCPL, RPL, 24>(name, fast, waves, stream);
        case 32:
            return launch_fast_at<Family, NCPL, RPL, 32>(name, fast, waves, stream);
        default:
            break;
        }
    }
#endif
    return launch_fast_at<Family, NCPL, RPL, kcap_lds(NCPL)>(name, fast, waves, stream);
}

// ---------------------------------------------------------------------------------------------
// Host-side launch helpers (shared by the instantiation units)
// ---------------------------------------------------------------------------------------------

template <int RPL, int KCAP>
constexpr long slab_doubles(bool global_l) {
    constexpr int RPLQ = (RPL * WAVE + KCAP + WAVE - 1) / WAVE;
    return static_cast<long>(RPLQ * WAVE) * (KCAP + 1) +
           (global_l ? static_cast<long>(lds_factor_doubles(KCAP, RPL)) : 0);
}

// The same slab size for the rerun (global-memory) instantiation of (ncpl, rpl), at run time.
constexpr long rerun_slab_doubles(int ncpl, int rpl) {
    const int kcap = ncpl * WAVE;
    const int rplq = (rpl * WAVE + kcap + WAVE - 1) / WAVE;
    return static_cast<long>(rplq * WAVE) * (kcap + 1) + static_cast<long>(lds_factor_doubles(kcap, rpl));
}


inline bool attach_work(FodArgs& a, unsigned long long* slot) {
    a.work = slot;
    return slot != nullptr;
}
#ifndef LSD_FP32
struct DrlArgs;
inline bool attach_work(DrlArgs& a, unsigned long long* slot); // drl_wave.hpp
#endif

// The fast launch (LDS factor) of a kernel family, then its rerun of the overflowed voxels with
// the factor in global memory, sized for every column (KG = NCPL * 64 >= n).  The rerun runs on
// `want` waves of scratch slabs from rs: allocated into rs when it is empty (a caller that
// launches the same instantiation repeatedly, the low-resolution rounds, allocates once), or as
// many whole blocks of slabs as a fixed region holds.
template <template <int, int, int, bool> class Family, int NCPL, int RPL, class Args>
int rerun_waves_in(RerunScratch* rs, int want, hipStream_t stream, int* waves) {
    const size_t per = sizeof(real) * static_cast<size_t>(slab_doubles<RPL, NCPL * WAVE>(true));
    if (rs->fixed) {
        const size_t fit = rs->bytes / per / WAVES_PER_BLOCK * WAVES_PER_BLOCK;
        if (fit < static_cast<size_t>(WAVES_PER_BLOCK)) {
            return fail("the workspace left for the global-memory solves holds " +
                        std::to_string(rs->bytes) + " bytes, less than one block of " +
                        std::to_string(WAVES_PER_BLOCK) + " slabs of " + std::to_string(per) +
                        " bytes; pass fewer voxels per call");
        }
        *waves = static_cast<int>(std::min<size_t>(fit, static_cast<size_t>(want)));
        return 0;
    }
    if (rs->base == nullptr) {
        const int rc = device_alloc(&rs->base, per * static_cast<size_t>(want), stream);
        if (rc) return rc;
        rs->bytes = per * static_cast<size_t>(want);
    }
    *waves = static_cast<int>(std::min<size_t>(rs->bytes / per, static_cast<size_t>(want)) /
                              WAVES_PER_BLOCK * WAVES_PER_BLOCK);
    return 0;
}

template <template <int, int, int, bool> class Family, int NCPL, int RPL, class Args>
int run_pair(const char* name, const char* rerun_name, const Args& fast, Args rerun, int waves,
             int* d_overflow_list, int* d_overflow_count, hipStream_t stream, RerunScratch* rs) {
    constexpr int KG = NCPL * WAVE;
    const long per = slab_doubles<RPL, KG>(true);
    if (fast.design.normal_equations == 0) {
        // QR fits throughout: every voxel takes the global-memory kernel.
        int all_waves = 0;
        int rc = rerun_waves_in<Family, NCPL, RPL, Args>(rs, std::min(waves, RERUN_WAVES_ALL),
                                                          stream, &all_waves);
        if (rc) return rc;
        rerun.scratch = SolveScratch{static_cast<real*>(rs->base), per};
        rerun.vs.overflow_list = nullptr;
        rerun.vs.overflow_count = nullptr;
        rerun.corr = nullptr;
        return launch(rerun_name, Family<NCPL, RPL, KG, true>::kernel, rerun, all_waves, 0,
                      stream);
    }
    int rc = launch_fast<Family, NCPL, RPL>(name, fast, waves, stream);
    if (rc) return rc;
    int rerun_waves = 0;
    rc = rerun_waves_in<Family, NCPL, RPL, Args>(rs, RERUN_WAVES, stream, &rerun_waves);
    if (rc) return rc;
    rerun.scratch = SolveScratch{static_cast<real*>(rs->base), per};
    rerun.vs.list = d_overflow_list;
    rerun.vs.list_count = d_overflow_count;
    rerun.vs.overflow_list = nullptr;
    rerun.vs.overflow_count = nullptr;
    rerun.corr = nullptr; // indexed by the fast launch's list positions
    return launch(rerun_name, Family<NCPL, RPL, KG, true>::kernel, rerun, rerun_waves, 0, stream);
}

// A fraction launch: each wave has its QR slab (per_wave bytes) in the block's LDS beside the
// block's shared bytes; four waves to a block, or two or one where four waves' slabs would take
// more than 64 KB (many fraction columns, more than 512 measurements).  A single wave past the
// device's LDS per block is refused.
template <class Kern, class Args>
int launch_fraction_blocks(const char* name, Kern kern, const Args& args, int waves,
                           size_t per_wave, size_t shared, hipStream_t stream) {
    constexpr size_t BLOCK_LDS = 64 * 1024;
    int per_block = WAVES_PER_BLOCK;
    while (per_block > 1 && per_wave * per_block + shared > BLOCK_LDS) {
        per_block /= 2;
    }
    const size_t lds = per_wave * per_block + shared;
    int dev = 0;
    int max_lds = 0;
    LSD_HIP(hipGetDevice(&dev));
    LSD_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    if (lds > static_cast<size_t>(max_lds)) {
        return fail("the fraction fit needs " + std::to_string(lds) + " bytes of LDS per wave, more "
                    "than the device's " + std::to_string(max_lds) +
                    "; fewer tissues or measurements fit");
    }
    if (per_block == WAVES_PER_BLOCK) {
        return launch_fill(name, kern, args, waves, lds, stream);
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, LAUNCH_WAVE * per_block, lds) ==
            hipSuccess &&
        per_cu > 0) {
        int cus = 0;
        LSD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        const long resident = static_cast<long>(per_cu) * cus * per_block;
        if (resident > 0 && resident < waves) {
            waves = static_cast<int>(resident);
        }
    }
    return launch(name, kern, args, waves, lds, stream, per_block);
}

template <int NC, int RP, bool FINAL>
int launch_fraction_rows(const char* name, const FractionArgs& args, int waves, hipStream_t stream) {
    const int qcols = (args.final_fit ? args.kernel.n_blocks : 1) + args.kernel.ci;
    const size_t slab = sizeof(real) * static_cast<size_t>(RP) * WAVE;
    if (fraction_columns_small(args.kernel.n_blocks, args.kernel.ci)) {
        return launch_fraction_blocks(name, fraction_kernel<NC, RP, FINAL, false, false, NQ_SMALL>,
                                      args, waves, qcols * slab, args.kernel.ci * slab, stream);
    }
    return launch_fraction_blocks(name, fraction_kernel<NC, RP, FINAL>, args, waves, qcols * slab,
                                  args.kernel.ci * slab, stream);
}

// The correlation pass with the isotropic fractions of at most NQ_SMALL - 1 tissues, or NQ - 1.
template <int NCB, bool BATCHED = false>
int launch_correlations(const char* name, const CorrArgs& args, int waves, hipStream_t stream) {
    if (args.ci + 1 <= NQ_SMALL) {
        return launch_fill(name, correlation_kernel<NCB, BATCHED, NQ_SMALL>, args, waves, 0, stream);
    }
    return launch_fill(name, correlation_kernel<NCB, BATCHED>, args, waves, 0, stream);
}

// Launchers of one (NCPL, RPL) instantiation, defined in its own translation unit.
// Names label the launches in the timing records (string literals: they are kept).
#define LSD_LAUNCHER_DECL(NC, RP)                                                              \
    int launch_fod_##NC##_##RP(const char* name, const char* rerun_name, const FodArgs& fast,   \
                               const FodArgs& rerun, int waves, int* ol, int* oc,               \
                               hipStream_t stream, RerunScratch* scratch);                     \
    int launch_fraction_##NC##_##RP(const char* name, const FractionArgs& args, int waves,      \
                                    hipStream_t stream);
#define LSD_LAUNCHER_DEF(NC, RP)                                                               \
    int launch_fod_##NC##_##RP(const char* name, const char* rerun_name, const FodArgs& fast,   \
                               const FodArgs& rerun, int waves, int* ol, int* oc,               \
                               hipStream_t stream, RerunScratch* scratch) {                    \
        return run_pair<fod_family, NC, RP>(name, rerun_name, fast, rerun, waves, ol, oc,       \
                                            stream, scratch);                                  \
    }                                                                                          \
    int launch_fraction_##NC##_##RP(const char* name, const FractionArgs& args, int waves,      \
                                    hipStream_t stream) {                                      \
        return args.final_fit ? launch_fraction_rows<NC, RP, true>(name, args, waves, stream)  \
                              : launch_fraction_rows<NC, RP, false>(name, args, waves, stream); \
    }

// Every instantiated (NCPL, RPL): X(NC, RP).
#if defined(LSD_FP32)
// The plain FP32 form: the 44- and 300-direction solves of the multi-tissue fit.
#define LSD_FOR_EACH_LAUNCHER(X) X(1, 4) X(1, 6) X(1, 8) X(5, 4) X(5, 6) X(5, 8)
#elif defined(LSD_FAST_BUILD)
#define LSD_FOR_EACH_LAUNCHER(X) X(1, 4) X(5, 4)
#else
#define LSD_FOR_EACH_RPL(X, NC) X(NC, 4) X(NC, 6) X(NC, 8) X(NC, 12) X(NC, 16)
#define LSD_FOR_EACH_LAUNCHER(X)                                                               \
    LSD_FOR_EACH_RPL(X, 1) LSD_FOR_EACH_RPL(X, 3) LSD_FOR_EACH_RPL(X, 5) LSD_FOR_EACH_RPL(X, 10)
#endif

LSD_FOR_EACH_LAUNCHER(LSD_LAUNCHER_DECL)

} // namespace LSD_NS
