// MI355X (gfx950) spherical deconvolution: kernels and the C-ABI of liblarmor_sd.so.
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

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/larmor_sd.h"
#include "fractions_wave.hpp"
#include "ridge_wave.hpp"

namespace lsd {

constexpr int NQ = 4;             // most fraction columns (groups + isotropic tissues)
constexpr int KCAP_LDS = 32;      // LDS factor capacity (active columns)
constexpr int WAVES_PER_BLOCK = 4;
constexpr int RERUN_WAVES = 32;   // waves of a rerun launch
constexpr double FRACTION_SUM_OFFSET = 1e-16;

constexpr int lds_factor_doubles(int kcap) { return (kcap + 1) * kcap + WAVE; }

// ---------------------------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------------------------

// corr_c = sum over rows, in order, of K(r, c) a_r, minus l1 (nnls.cpp:302).
template <int NCPL, int RPL>
__device__ __forceinline__ void correlations(const RidgeDesignDev& D, const RowVec<RPL>& a, double l1,
                             double (&corr)[NCPL]) {
    const int l = lane_id();
#pragma unroll
    for (int t = 0; t < NCPL; ++t) {
        corr[t] = 0.0;
    }
#pragma unroll
    for (int ts = 0; ts < RPL; ++ts) {
        const int rows_here = D.m - WAVE * ts;
        const int count = rows_here < WAVE ? rows_here : WAVE;
        for (int ll = 0; ll < count; ++ll) {
            const int r = WAVE * ts + ll;
            const double ar = readlane(a.v[ts], ll);
            const double* krow = D.kernel + static_cast<long>(r) * D.n;
#pragma unroll
            for (int t = 0; t < NCPL; ++t) {
                const int c = l + WAVE * t;
                if (c < D.n) {
                    corr[t] = corr[t] + krow[c] * ar;
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
__device__ __forceinline__ void load_signal(const double* signals, long v, int m, RowVec<RPL>& s) {
    const int l = lane_id();
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
        const int r = l + WAVE * t;
        s.v[t] = r < m ? signals[v * m + r] : 0.0;
    }
}

// max(w s - W_iso p, 0) per row (multi_tissue.cpp:20-24), p the isotropic fractions.
template <int RPL>
__device__ __forceinline__ void anisotropic_signal(const RowVec<RPL>& s, const double* weights,
                                   const double* w_iso, int m, int ci, const double (&p)[NQ],
                                   RowVec<RPL>& a) {
    const int l = lane_id();
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
        const int r = l + WAVE * t;
        double out = 0.0;
        if (r < m) {
            const double ws = s.v[t] * weights[r];
            double acc = 0.0;
#pragma unroll
            for (int tt = 0; tt < NQ; ++tt) {
                if (tt < ci) {
                    acc = acc + w_iso[static_cast<long>(r) * ci + tt] * p[tt];
                }
            }
            out = fmax(ws - acc, 0.0);
        }
        a.v[t] = out;
    }
}

struct SolveScratch {
    double* base;     // this launch's slab array
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

struct SingleArgs {
    const double* signals;
    int m;
    RidgeDesignDev design;
    double tol;
    double* fods;
    int* status;
    VoxelSet vs;
    SolveScratch scratch;
};

// fit_single_tissue: one ridge NNLS per voxel (solver.cpp:14-21 through sd.cpp:fit_single_tissue).
template <int NCPL, int RPL, int KCAP, bool GLOBAL_L>
__global__ void __launch_bounds__(WAVE * WAVES_PER_BLOCK) single_kernel(SingleArgs A) {
    constexpr int RPLQ = (RPL * WAVE + KCAP + WAVE - 1) / WAVE;
    extern __shared__ double smem[];
    const int w = threadIdx.x / WAVE;
    const long gw = static_cast<long>(blockIdx.x) * (blockDim.x / WAVE) + w;
    const long n_waves = static_cast<long>(gridDim.x) * (blockDim.x / WAVE);
    double* slab = A.scratch.base + gw * A.scratch.per_wave;
    double* qrs = slab;
    double* L = GLOBAL_L ? slab + static_cast<long>(RPLQ * WAVE) * (KCAP + 1)
                         : smem + w * lds_factor_doubles(KCAP);
    const long count = voxel_count(A.vs);
    const int l = lane_id();
    for (long it = gw; it < count; it += n_waves) {
        const long v = A.vs.list ? A.vs.list[it] : it;
        RowVec<RPL> s;
        load_signal<RPL>(A.signals, v, A.m, s);
        double corr[NCPL];
        correlations<NCPL, RPL>(A.design, s, 0.0, corr);
        RidgeNnlsWave<NCPL, KCAP, RPLQ> solver(A.design, A.tol, 3 * A.design.n, 0.0, L, qrs);
        const int st = solver.template run<RPL>(corr, s);
        if (st & ST_OVERFLOW) {
            record_overflow(A.vs, v);
            continue;
        }
#pragma unroll
        for (int t = 0; t < NCPL; ++t) {
            const int c = l + WAVE * t;
            if (c < A.design.n) {
                A.fods[v * A.design.n + c] = solver.x[t];
            }
        }
        if (l == 0) {
            A.status[v] = st;
        }
    }
}

struct LoopArgs {
    const double* signals;
    int m;
    const double* weights;
    RidgeDesignDev design;    // low-resolution solver
    FractionKernelDev kernel; // low-resolution kernel (unweighted)
    const double* w_iso;      // (m, ci) weighted isotropic columns
    double tol;
    int max_rounds;
    double fraction_tol;
    double* iso_out;          // (n_vox, ci)
    int* rounds_out;
    int* status;
    VoxelSet vs;
    SolveScratch scratch;
};

// The low-resolution rounds of fit_multi_tissue_block (multi_tissue.cpp:47-80): per round the
// anisotropic signal, its low-resolution FOD and the lumped fractions, until no fraction moves by
// fraction_tol or max_rounds rounds ran; the voxel keeps its last isotropic fractions.
template <int NCPL, int RPL, int KCAP, bool GLOBAL_L>
__global__ void __launch_bounds__(WAVE * WAVES_PER_BLOCK) loop_kernel(LoopArgs A) {
    constexpr int RPLQ = (RPL * WAVE + KCAP + WAVE - 1) / WAVE;
    extern __shared__ double smem[];
    const int w = threadIdx.x / WAVE;
    const long gw = static_cast<long>(blockIdx.x) * (blockDim.x / WAVE) + w;
    const long n_waves = static_cast<long>(gridDim.x) * (blockDim.x / WAVE);
    double* slab = A.scratch.base + gw * A.scratch.per_wave;
    double* qrs = slab;
    double* L = GLOBAL_L ? slab + static_cast<long>(RPLQ * WAVE) * (KCAP + 1)
                         : smem + w * lds_factor_doubles(KCAP);
    const long count = voxel_count(A.vs);
    const int l = lane_id();
    const int ci = A.kernel.ci;
    const int nlf = 1 + ci;
    for (long it = gw; it < count; it += n_waves) {
        const long v = A.vs.list ? A.vs.list[it] : it;
        RowVec<RPL> s;
        load_signal<RPL>(A.signals, v, A.m, s);
        double p[NQ];
        double prev[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            p[q] = 0.0;
            prev[q] = __builtin_inf();
        }
        int status = 0;
        int rounds = 0;
        for (int round = 0; round < A.max_rounds; ++round) {
            RowVec<RPL> a;
            anisotropic_signal<RPL>(s, A.weights, A.w_iso, A.m, ci, p, a);
            double corr[NCPL];
            correlations<NCPL, RPL>(A.design, a, 0.0, corr);
            RidgeNnlsWave<NCPL, KCAP, RPLQ> solver(A.design, A.tol, 3 * A.design.n, 0.0, L, qrs);
            const int st = solver.template run<RPL>(corr, a);
            status |= st;
            if (st & ST_OVERFLOW) {
                break;
            }
            double fr[NQ];
            if (!fractions_wave<NCPL, RPL, NQ>(A.kernel, solver.x, s, true, false, 0.0, fr)) {
                status |= ST_NOT_CONVERGED;
            }
            rounds += 1;
            bool converged = true;
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                if (q >= 1 && q < nlf) {
                    p[q - 1] = fr[q];
                }
                if (q < nlf && !(fabs(fr[q] - prev[q]) < A.fraction_tol)) {
                    converged = false;
                }
            }
            if (converged) {
                break;
            }
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                prev[q] = q < nlf ? fr[q] : prev[q];
            }
        }
        if (status & ST_OVERFLOW) {
            record_overflow(A.vs, v);
            continue;
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            if (l == q && q < ci) {
                A.iso_out[v * ci + q] = p[q];
            }
        }
        if (l == 0) {
            A.rounds_out[v] = rounds;
            A.status[v] = status;
        }
    }
}

struct FinalArgs {
    const double* signals;
    int m;
    const double* weights;
    RidgeDesignDev design;    // high-resolution solver
    FractionKernelDev kernel; // high-resolution kernel (unweighted)
    const double* w_iso;
    double tol;
    double final_tol;
    const double* iso_in;     // (n_vox, ci)
    double* fods;             // (n_vox, ca)
    double* fractions;        // (n_vox, n_blocks + ci)
    int* status;
    VoxelSet vs;
    SolveScratch scratch;
};

// The high-resolution FOD and the final per-block fractions, normalised to sum to 1
// (multi_tissue.cpp:82-99).
template <int NCPL, int RPL, int KCAP, bool GLOBAL_L>
__global__ void __launch_bounds__(WAVE * WAVES_PER_BLOCK) final_kernel(FinalArgs A) {
    constexpr int RPLQ = (RPL * WAVE + KCAP + WAVE - 1) / WAVE;
    extern __shared__ double smem[];
    const int w = threadIdx.x / WAVE;
    const long gw = static_cast<long>(blockIdx.x) * (blockDim.x / WAVE) + w;
    const long n_waves = static_cast<long>(gridDim.x) * (blockDim.x / WAVE);
    double* slab = A.scratch.base + gw * A.scratch.per_wave;
    double* qrs = slab;
    double* L = GLOBAL_L ? slab + static_cast<long>(RPLQ * WAVE) * (KCAP + 1)
                         : smem + w * lds_factor_doubles(KCAP);
    const long count = voxel_count(A.vs);
    const int l = lane_id();
    const int ci = A.kernel.ci;
    const int nf = A.kernel.n_blocks + ci;
    for (long it = gw; it < count; it += n_waves) {
        const long v = A.vs.list ? A.vs.list[it] : it;
        RowVec<RPL> s;
        load_signal<RPL>(A.signals, v, A.m, s);
        double p[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            p[q] = q < ci ? A.iso_in[v * ci + q] : 0.0;
        }
        RowVec<RPL> a;
        anisotropic_signal<RPL>(s, A.weights, A.w_iso, A.m, ci, p, a);
        double corr[NCPL];
        correlations<NCPL, RPL>(A.design, a, 0.0, corr);
        RidgeNnlsWave<NCPL, KCAP, RPLQ> solver(A.design, A.tol, 3 * A.design.n, 0.0, L, qrs);
        int st = solver.template run<RPL>(corr, a);
        if (st & ST_OVERFLOW) {
            record_overflow(A.vs, v);
            continue;
        }
        if (A.fods) {
#pragma unroll
            for (int t = 0; t < NCPL; ++t) {
                const int c = l + WAVE * t;
                if (c < A.design.n) {
                    A.fods[v * A.design.n + c] = solver.x[t];
                }
            }
        }
        double fr[NQ];
        if (!fractions_wave<NCPL, RPL, NQ>(A.kernel, solver.x, s, false, true, A.final_tol, fr)) {
            st |= ST_NOT_CONVERGED;
        }
        double total = 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            if (q < nf) {
                total = total + fr[q];
            }
        }
        total = total + FRACTION_SUM_OFFSET;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            if (l == q && q < nf) {
                A.fractions[v * nf + q] = fr[q] / total;
            }
        }
        if (l == 0) {
            A.status[v] |= st;
        }
    }
}

__global__ void zero_ints(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        p[i] = 0;
    }
}

// ---------------------------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------------------------

thread_local std::string g_error;

int fail(const std::string& message) {
    g_error = message;
    return LSD_ERR_INVALID;
}

int hip_fail(hipError_t e, const char* what) {
    g_error = std::string(what) + ": " + hipGetErrorString(e);
    return LSD_ERR_HIP;
}

#define LSD_HIP(call)                                                                          \
    do {                                                                                       \
        const hipError_t lsd_e = (call);                                                       \
        if (lsd_e != hipSuccess) {                                                             \
            return hip_fail(lsd_e, #call);                                                     \
        }                                                                                      \
    } while (0)

std::string shape_tuple(int64_t r, int64_t c) {
    return "(" + std::to_string(r) + ", " + std::to_string(c) + ")";
}

// Which instantiation covers a problem size.
int ncpl_for(int64_t n) {
    if (n <= 64) return 1;
    if (n <= 192) return 3;
    if (n <= 320) return 5;
    if (n <= 640) return 10;
    return -1;
}
int rpl_for(int64_t m) {
    if (m <= 256) return 4;
    if (m <= 384) return 6;
    if (m <= 512) return 8;
    return -1;
}

int device_alloc(void** p, size_t bytes, hipStream_t stream) {
    if (bytes == 0) {
        bytes = 8;
    }
    LSD_HIP(hipMallocAsync(p, bytes, stream));
    return 0;
}

} // namespace lsd

using namespace lsd;

struct lsd_solver {
    int64_t m = 0;
    int64_t n = 0;
    double ridge = 0.0;
    double tol = 0.0;
    int normal_equations = 0;
    double* d_kernel = nullptr;
    double* d_gram = nullptr;
};

struct lsd_tissue_kernel {
    int64_t m = 0;
    int64_t ca = 0;
    int64_t ci = 0;
    std::vector<int> blocks;     // (n_blocks, 2)
    double* d_aniso_t = nullptr; // (ca, m)
    double* d_iso_t = nullptr;   // (ci, m)
    double* d_w_iso = nullptr;   // (m, ci)
    int* d_blocks = nullptr;
};

namespace lsd {

RidgeDesignDev design_of(const lsd_solver* s) {
    return RidgeDesignDev{static_cast<int>(s->m), static_cast<int>(s->n), s->d_kernel, s->d_gram,
                          s->ridge, s->normal_equations};
}

FractionKernelDev fractions_of(const lsd_tissue_kernel* k) {
    return FractionKernelDev{static_cast<int>(k->m), static_cast<int>(k->ca), static_cast<int>(k->ci),
                             k->d_aniso_t, k->d_iso_t, k->d_blocks,
                             static_cast<int>(k->blocks.size() / 2)};
}

// Grid of the main launch: enough waves to fill the chip, never more than voxels.
int main_waves(int64_t n_vox) {
    int dev = 0;
    hipGetDevice(&dev);
    int cus = 0;
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (cus <= 0) {
        cus = 256;
    }
    const int64_t target = static_cast<int64_t>(cus) * 16;
    int64_t waves = std::min<int64_t>(target, std::max<int64_t>(n_vox, 1));
    waves = (waves + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK * WAVES_PER_BLOCK;
    return static_cast<int>(waves);
}

template <int RPL, int KCAP>
constexpr long slab_doubles(bool global_l) {
    constexpr int RPLQ = (RPL * WAVE + KCAP + WAVE - 1) / WAVE;
    return static_cast<long>(RPLQ * WAVE) * (KCAP + 1) +
           (global_l ? static_cast<long>(lds_factor_doubles(KCAP)) : 0);
}

// Launches kern over `waves` waves with the factor in LDS (or not).
template <class Kern, class Args>
int launch(Kern kern, const Args& args, int waves, size_t lds_bytes, hipStream_t stream) {
    const int blocks = (waves + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(WAVE * WAVES_PER_BLOCK), lds_bytes, stream, args);
    LSD_HIP(hipGetLastError());
    return 0;
}

#define LSD_DISPATCH_BODY(KERNEL, ARGS_T)                                                      \
    template <int NCPL, int RPL>                                                               \
    static int run_pair(const ARGS_T& fast, ARGS_T rerun, int waves, int64_t n_items,           \
                        int* d_overflow_list, int* d_overflow_count, hipStream_t stream,         \
                        void** rerun_scratch) {                                                \
        constexpr int KG = NCPL * WAVE;                                                        \
        const size_t lds = sizeof(double) * lds_factor_doubles(KCAP_LDS) * WAVES_PER_BLOCK;     \
        int rc = launch(KERNEL<NCPL, RPL, KCAP_LDS, false>, fast, waves, lds, stream);          \
        if (rc) return rc;                                                                     \
        const long per = slab_doubles<RPL, KG>(true);                                          \
        rc = device_alloc(rerun_scratch, sizeof(double) * per * RERUN_WAVES, stream);           \
        if (rc) return rc;                                                                     \
        rerun.scratch = SolveScratch{static_cast<double*>(*rerun_scratch), per};               \
        rerun.vs.list = d_overflow_list;                                                       \
        rerun.vs.list_count = d_overflow_count;                                                \
        rerun.vs.overflow_list = nullptr;                                                      \
        rerun.vs.overflow_count = nullptr;                                                     \
        (void)n_items;                                                                         \
        return launch(KERNEL<NCPL, RPL, KG, true>, rerun, RERUN_WAVES, 0, stream);             \
    }

struct SingleDispatch {
    static constexpr int kMaxNcpl = 10;
    LSD_DISPATCH_BODY(single_kernel, SingleArgs)
};
struct LoopDispatch {
    static constexpr int kMaxNcpl = 3;
    LSD_DISPATCH_BODY(loop_kernel, LoopArgs)
};
struct FinalDispatch {
    static constexpr int kMaxNcpl = 10;
    LSD_DISPATCH_BODY(final_kernel, FinalArgs)
};

template <class D, class Args>
int dispatch(int ncpl, int rpl, const Args& fast, const Args& rerun, int waves, int64_t n_items,
             int* ol, int* oc, hipStream_t stream, void** rerun_scratch) {
#define LSD_CASE(NC, RP)                                                                       \
    if constexpr (NC <= D::kMaxNcpl) {                                                         \
        if (ncpl == NC && rpl == RP)                                                           \
            return D::template run_pair<NC, RP>(fast, rerun, waves, n_items, ol, oc, stream,    \
                                                rerun_scratch);                                \
    }
#ifdef LSD_FAST_BUILD
    // Development build: the 3-shell workload's sizes only.
    LSD_CASE(1, 4) LSD_CASE(5, 4)
#else
    LSD_CASE(1, 4) LSD_CASE(1, 6) LSD_CASE(1, 8)
    LSD_CASE(3, 4) LSD_CASE(3, 6) LSD_CASE(3, 8)
    LSD_CASE(5, 4) LSD_CASE(5, 6) LSD_CASE(5, 8)
    LSD_CASE(10, 4) LSD_CASE(10, 6) LSD_CASE(10, 8)
#endif
#undef LSD_CASE
    return fail("no kernel instantiation for " + std::to_string(ncpl) + " column slots and " +
                std::to_string(rpl) + " row slots");
}

int upload(double** dst, const std::vector<double>& src) {
    LSD_HIP(hipMalloc(reinterpret_cast<void**>(dst), std::max<size_t>(8, src.size() * sizeof(double))));
    if (!src.empty()) {
        LSD_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    return 0;
}

} // namespace lsd

extern "C" {

const char* lsd_last_error(void) { return lsd::g_error.c_str(); }

int lsd_abi_version(void) { return LSD_ABI_VERSION; }

int lsd_solver_create_l2nnls(const double* kernel, int64_t m, int64_t n, double ridge, double tol_x,
                             lsd_solver** out) {
    if (!(ridge >= 0.0)) {
        return fail("reg must be at least 0");
    }
    if (!(tol_x > 0.0)) {
        return fail("tol_x must be positive");
    }
    if (m < 0 || n < 0 || (m * n > 0 && kernel == nullptr)) {
        return fail("kernel must be (m, n) with m, n >= 0");
    }
    auto* s = new lsd_solver();
    s->m = m;
    s->n = n;
    s->ridge = ridge;
    s->tol = tol_x;
    // Gram K^T K + ridge^2 I with each entry a sequential sum over rows (bitwise symmetric).
    std::vector<double> gram(static_cast<size_t>(n * n));
    for (int64_t a = 0; a < n; ++a) {
        for (int64_t b = 0; b < n; ++b) {
            double acc = 0.0;
            for (int64_t i = 0; i < m; ++i) {
                acc += kernel[i * n + a] * kernel[i * n + b];
            }
            gram[a * n + b] = acc;
        }
    }
    const double r2 = ridge * ridge;
    for (int64_t a = 0; a < n; ++a) {
        gram[a * n + a] += r2;
    }
    double fro = 0.0;
    for (int64_t i = 0; i < m * n; ++i) {
        fro += kernel[i] * kernel[i];
    }
    // nnls.cpp:280-292: the normal equations with no ridge, or a well-conditioned one.
    s->normal_equations = (r2 == 0.0 || fro + r2 <= 1e8 * r2) ? 1 : 0;
    std::vector<double> k(kernel, kernel + m * n);
    int rc = upload(&s->d_kernel, k);
    if (!rc) rc = upload(&s->d_gram, gram);
    if (rc) {
        lsd_solver_destroy(s);
        return rc;
    }
    *out = s;
    return 0;
}

void lsd_solver_destroy(lsd_solver* s) {
    if (!s) return;
    if (s->d_kernel) (void)hipFree(s->d_kernel);
    if (s->d_gram) (void)hipFree(s->d_gram);
    delete s;
}

int64_t lsd_solver_n_dirs(const lsd_solver* s) { return s->n; }
int64_t lsd_solver_n_meas(const lsd_solver* s) { return s->m; }
int lsd_solver_normal_equations(const lsd_solver* s) { return s->normal_equations; }

int lsd_tissue_kernel_create(const double* k_aniso, const double* k_iso, const double* w_k_iso,
                             int64_t m, int64_t ca, int64_t ci, const int64_t* blocks,
                             int64_t n_blocks, lsd_tissue_kernel** out) {
    bool valid = n_blocks > 0;
    for (int64_t b = 0; b < n_blocks; ++b) {
        const int64_t start = blocks[2 * b];
        const int64_t stop = blocks[2 * b + 1];
        valid = valid && start >= 0 && start < stop && stop <= ca;
    }
    if (!valid) {
        std::string text = "[";
        for (int64_t b = 0; b < n_blocks; ++b) {
            text += (b ? ", " : "") + shape_tuple(blocks[2 * b], blocks[2 * b + 1]);
        }
        return fail("aniso_blocks must be non-empty ranges within 0.." + std::to_string(ca) +
                    ", got " + text + "]");
    }
    if (n_blocks + ci > NQ || 1 + ci > NQ) {
        return fail("at most " + std::to_string(NQ) +
                    " fraction columns (anisotropic blocks plus isotropic tissues) are supported");
    }
    auto* k = new lsd_tissue_kernel();
    k->m = m;
    k->ca = ca;
    k->ci = ci;
    for (int64_t b = 0; b < 2 * n_blocks; ++b) {
        k->blocks.push_back(static_cast<int>(blocks[b]));
    }
    std::vector<double> at(static_cast<size_t>(ca * m));
    for (int64_t i = 0; i < m; ++i) {
        for (int64_t c = 0; c < ca; ++c) {
            at[c * m + i] = k_aniso[i * ca + c];
        }
    }
    std::vector<double> it(static_cast<size_t>(ci * m));
    for (int64_t i = 0; i < m; ++i) {
        for (int64_t c = 0; c < ci; ++c) {
            it[c * m + i] = k_iso[i * ci + c];
        }
    }
    std::vector<double> wi(w_k_iso, w_k_iso + m * ci);
    int rc = upload(&k->d_aniso_t, at);
    if (!rc) rc = upload(&k->d_iso_t, it);
    if (!rc) rc = upload(&k->d_w_iso, wi);
    if (!rc) {
        const hipError_t e1 = hipMalloc(reinterpret_cast<void**>(&k->d_blocks),
                                        sizeof(int) * k->blocks.size());
        const hipError_t e2 =
            e1 == hipSuccess ? hipMemcpy(k->d_blocks, k->blocks.data(), sizeof(int) * k->blocks.size(),
                                         hipMemcpyHostToDevice)
                             : e1;
        if (e2 != hipSuccess) rc = hip_fail(e2, "kernel blocks upload");
    }
    if (rc) {
        lsd_tissue_kernel_destroy(k);
        return rc;
    }
    *out = k;
    return 0;
}

void lsd_tissue_kernel_destroy(lsd_tissue_kernel* k) {
    if (!k) return;
    if (k->d_aniso_t) (void)hipFree(k->d_aniso_t);
    if (k->d_iso_t) (void)hipFree(k->d_iso_t);
    if (k->d_w_iso) (void)hipFree(k->d_w_iso);
    if (k->d_blocks) (void)hipFree(k->d_blocks);
    delete k;
}

int lsd_fit_single_tissue(const double* signals, int64_t n_voxels, int64_t m,
                          const lsd_solver* solver, double* fods, int32_t* status, void* stream_) {
    auto stream = static_cast<hipStream_t>(stream_);
    if (m != solver->m) {
        return fail("signals " + shape_tuple(n_voxels, m) + " do not have the solver's " +
                    std::to_string(solver->m) + " measurements");
    }
    const int ncpl = ncpl_for(solver->n);
    const int rpl = rpl_for(m);
    if (ncpl < 0 || rpl < 0) {
        return fail("problem size " + shape_tuple(m, solver->n) + " exceeds 512 measurements or 640 directions");
    }
    if (n_voxels == 0) {
        return 0;
    }
    const int waves = main_waves(n_voxels);
    void* scratch = nullptr;
    void* rerun_scratch = nullptr;
    void* counters = nullptr;
    void* olist = nullptr;
    long per = 0;
#define SLAB_CASE(RP) if (rpl == RP) per = slab_doubles<RP, KCAP_LDS>(false);
    SLAB_CASE(4) SLAB_CASE(6) SLAB_CASE(8)
#undef SLAB_CASE
    int rc = device_alloc(&scratch, sizeof(double) * per * waves, stream);
    if (!rc) rc = device_alloc(&counters, sizeof(int) * 4, stream);
    if (!rc) rc = device_alloc(&olist, sizeof(int) * n_voxels, stream);
    if (rc) return rc;
    hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, stream, static_cast<int*>(counters), 4);
    SingleArgs fast{};
    fast.signals = signals;
    fast.m = static_cast<int>(m);
    fast.design = design_of(solver);
    fast.tol = solver->tol;
    fast.fods = fods;
    fast.status = status;
    fast.vs = VoxelSet{n_voxels, nullptr, nullptr, static_cast<int*>(olist), static_cast<int*>(counters)};
    fast.scratch = SolveScratch{static_cast<double*>(scratch), per};
    SingleArgs rerun = fast;
    rc = dispatch<SingleDispatch>(ncpl, rpl, fast, rerun, waves, n_voxels, static_cast<int*>(olist),
                                  static_cast<int*>(counters), stream, &rerun_scratch);
    (void)hipFreeAsync(scratch, stream);
    (void)hipFreeAsync(counters, stream);
    (void)hipFreeAsync(olist, stream);
    if (rerun_scratch) (void)hipFreeAsync(rerun_scratch, stream);
    return rc;
}

int lsd_fit_multi_tissue(const double* signals, int64_t n_voxels, int64_t m, const double* weights,
                         const lsd_tissue_kernel* lr, const lsd_solver* lr_solver,
                         const lsd_tissue_kernel* hr, const lsd_solver* hr_solver,
                         int64_t max_rounds, double fraction_tol, double final_tol, double* fods,
                         double* fractions, int32_t* rounds, int32_t* status, void* stream_) {
    auto stream = static_cast<hipStream_t>(stream_);
    if (m != hr->m || lr->m != m) {
        return fail("signals " + shape_tuple(n_voxels, m) + ", weights " + std::to_string(m) +
                    " and the kernels' " + std::to_string(hr->m) + " measurements do not match");
    }
    if (lr->ci != hr->ci) {
        return fail("the two kernels have different numbers of isotropic tissues");
    }
    auto matching = [&](const char* name, const lsd_tissue_kernel* k, const lsd_solver* s) {
        if (s->n != k->ca || s->m != m) {
            return fail(std::string("the ") + name + " solver (" + std::to_string(s->n) +
                        " directions, " + std::to_string(s->m) +
                        " measurements) does not match its kernel (" + std::to_string(k->ca) +
                        " directions, " + std::to_string(m) + " measurements)");
        }
        return 0;
    };
    int rc = matching("low-resolution", lr, lr_solver);
    if (!rc) rc = matching("high-resolution", hr, hr_solver);
    if (rc) return rc;
    if (max_rounds < 0) {
        return fail("max_outer must be at least 0");
    }
    const int ncpl_lr = ncpl_for(lr->ca);
    const int ncpl_hr = ncpl_for(hr->ca);
    const int rpl = rpl_for(m);
    if (ncpl_lr < 0 || ncpl_hr < 0 || rpl < 0) {
        return fail("problem size exceeds 512 measurements or 640 directions");
    }
    if (n_voxels == 0) {
        return 0;
    }
    const int ci = static_cast<int>(lr->ci);
    const int waves = main_waves(n_voxels);
    long per = 0;
#define SLAB_CASE(RP) if (rpl == RP) per = slab_doubles<RP, KCAP_LDS>(false);
    SLAB_CASE(4) SLAB_CASE(6) SLAB_CASE(8)
#undef SLAB_CASE
    void* scratch = nullptr;
    void* counters = nullptr;
    void* olist = nullptr;
    void* iso = nullptr;
    void* rerun1 = nullptr;
    void* rerun2 = nullptr;
    rc = device_alloc(&scratch, sizeof(double) * per * waves, stream);
    if (!rc) rc = device_alloc(&counters, sizeof(int) * 4, stream);
    if (!rc) rc = device_alloc(&olist, sizeof(int) * 2 * n_voxels, stream);
    if (!rc) rc = device_alloc(&iso, sizeof(double) * std::max(1, ci) * n_voxels, stream);
    if (rc) return rc;
    int* cnt = static_cast<int*>(counters);
    int* list1 = static_cast<int*>(olist);
    int* list2 = list1 + n_voxels;
    hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, stream, cnt, 4);

    LoopArgs la{};
    la.signals = signals;
    la.m = static_cast<int>(m);
    la.weights = weights;
    la.design = design_of(lr_solver);
    la.kernel = fractions_of(lr);
    la.w_iso = lr->d_w_iso;
    la.tol = lr_solver->tol;
    la.max_rounds = static_cast<int>(max_rounds);
    la.fraction_tol = fraction_tol;
    la.iso_out = static_cast<double*>(iso);
    la.rounds_out = rounds;
    la.status = status;
    la.vs = VoxelSet{n_voxels, nullptr, nullptr, list1, cnt};
    la.scratch = SolveScratch{static_cast<double*>(scratch), per};
    rc = dispatch<LoopDispatch>(ncpl_lr, rpl, la, la, waves, n_voxels, list1, cnt, stream, &rerun1);

    if (!rc) {
        FinalArgs fa{};
        fa.signals = signals;
        fa.m = static_cast<int>(m);
        fa.weights = weights;
        fa.design = design_of(hr_solver);
        fa.kernel = fractions_of(hr);
        fa.w_iso = hr->d_w_iso;
        fa.tol = hr_solver->tol;
        fa.final_tol = final_tol;
        fa.iso_in = static_cast<double*>(iso);
        fa.fods = fods;
        fa.fractions = fractions;
        fa.status = status;
        fa.vs = VoxelSet{n_voxels, nullptr, nullptr, list2, cnt + 1};
        fa.scratch = SolveScratch{static_cast<double*>(scratch), per};
        rc = dispatch<FinalDispatch>(ncpl_hr, rpl, fa, fa, waves, n_voxels, list2, cnt + 1, stream,
                                     &rerun2);
    }
    (void)hipFreeAsync(scratch, stream);
    (void)hipFreeAsync(counters, stream);
    (void)hipFreeAsync(olist, stream);
    (void)hipFreeAsync(iso, stream);
    if (rerun1) (void)hipFreeAsync(rerun1, stream);
    if (rerun2) (void)hipFreeAsync(rerun2, stream);
    return rc;
}

} // extern "C"
