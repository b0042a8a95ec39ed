// The log-signal fits of many voxels, one wavefront per voxel: lsd_fit_log_signal, larmor's
// fit_log_signal without constraints (cpp/src/python/dwi.cpp:41-82; cpp/src/dwi/log_fit.cpp:28-62,
// 159-185).  The host prepares the design once (unit-norm columns, their norms and the
// pseudo-inverse, make_log_signal_design log_fit.cpp:134-149); every voxel then costs a few
// hundred wave sums.
#include "kernels.hpp"
#include "log_fit_wave.hpp"

namespace lsd {

namespace {

struct LogFitArgs {
    LogFitDev F;
    const double* log_s; // (n_vox, n)
    long n_vox;
    double* params;      // (n_vox, p)
    int* refits;         // (n_vox), or null
};

constexpr int LOG_FIT_PMAX = 8;

template <int RPL>
__global__ void __launch_bounds__(WAVE * WAVES_PER_BLOCK) log_fit_kernel(LogFitArgs A) {
    const long gw = static_cast<long>(blockIdx.x) * (blockDim.x / WAVE) +
                    __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / WAVE);
    const long n_waves = static_cast<long>(gridDim.x) * (blockDim.x / WAVE);
    const int l = lane_id();
    for (long v = gw; v < A.n_vox; v += n_waves) {
        RowVec<RPL> y;
        load_signal<RPL>(A.log_s, v, A.F.n, y);
        double params[LOG_FIT_PMAX];
        const int refits = log_fit_voxel<RPL, LOG_FIT_PMAX>(A.F, y, params);
#pragma unroll
        for (int a = 0; a < LOG_FIT_PMAX; ++a) {
            if (a < A.F.p && l == a) {
                A.params[v * A.F.p + a] = params[a];
            }
        }
        if (l == 0 && A.refits != nullptr) {
            A.refits[v] = refits;
        }
    }
}

} // namespace

} // namespace lsd

using namespace lsd;

extern "C" {

int lsd_fit_log_signal(const double* xt, const double* pinv, const double* scales, int64_t n,
                       int64_t p, const double* log_s, int64_t n_voxels, int weighted,
                       int64_t max_reweightings, double tol, double* params, int32_t* refits,
                       void* stream_) {
    auto stream = static_cast<hipStream_t>(stream_);
    if (n < p || p < 1) {
        return fail("x must be (n, p) with n >= p >= 1: got (" + std::to_string(n) + ", " +
                    std::to_string(p) + ")");
    }
    if (p > LOG_FIT_PMAX || n > 8 * WAVE) {
        return fail("problem size " + shape_tuple(n, p) +
                    " exceeds 512 measurements or 8 parameters");
    }
    if (max_reweightings < 0) {
        return fail("max_iter must be at least 0");
    }
    if (n_voxels < 0) {
        return fail("n_voxels must be at least 0");
    }
    if (n_voxels == 0) {
        return 0;
    }
    LogFitArgs a{};
    a.F = LogFitDev{xt, pinv, scales, static_cast<int>(n), static_cast<int>(p), weighted != 0,
                    static_cast<int>(std::min<int64_t>(max_reweightings, 1 << 30)), tol};
    a.log_s = log_s;
    a.n_vox = n_voxels;
    a.params = params;
    a.refits = refits;
    const int waves = main_waves(n_voxels);
    return n <= 4 * WAVE ? launch_fill("log_fit", log_fit_kernel<4>, a, waves, 0, stream)
                         : launch_fill("log_fit", log_fit_kernel<8>, a, waves, 0, stream);
}

} // extern "C"
