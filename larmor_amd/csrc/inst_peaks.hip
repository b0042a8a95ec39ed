// Instantiations and launchers of the peak search, derivative and climb kernels: exact ones for
// symmetric series of even lmax 0-12 and full series of lmax 1-8, and the lmax-20 ones, which take
// every higher lmax (up to 20) at run time.
#include "kernels.hpp"
#include "peaks_wave.hpp"

namespace lsd {

namespace {

// Spheres up to 704 vertices go four waves to a block (at most 33.8 KB of LDS); larger ones one
// wave to a block, so that a block's LDS stays at one wave's share (up to 36 KB for 3072
// vertices) and the grid is sized by the occupancy of one-wave blocks.
constexpr int PEAKS_BLOCK_LDS_LIMIT = 11 * WAVE * (8 + 4) * WAVES_PER_BLOCK;

template <int LMAX, bool FULL>
int launch_peaks_t(const PeaksArgs& a, int waves, hipStream_t stream) {
    const size_t per_wave = static_cast<size_t>(peaks_lds_bytes_per_wave(a.nv));
    if (per_wave * WAVES_PER_BLOCK <= static_cast<size_t>(PEAKS_BLOCK_LDS_LIMIT)) {
        return launch_fill("peaks", peaks_kernel<LMAX, FULL>, a, waves, per_wave * WAVES_PER_BLOCK,
                           stream);
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, peaks_kernel<LMAX, FULL>, WAVE,
                                                     per_wave) == hipSuccess &&
        per_cu > 0) {
        int dev = 0;
        int cus = 0;
        LSD_HIP(hipGetDevice(&dev));
        LSD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        const long resident = static_cast<long>(per_cu) * cus;
        if (resident > 0 && resident < waves) {
            waves = static_cast<int>(resident);
        }
    }
    return launch("peaks", peaks_kernel<LMAX, FULL>, a, waves, per_wave, stream, 1);
}

template <int LMAX, bool FULL>
int launch_derivs_t(const double* coeffs, long n, int nc, const double* theta, const double* phi,
                    double* out, ShConstantsDev k, hipStream_t stream) {
    const int threads = WAVE * WAVES_PER_BLOCK;
    const long blocks = (n + threads - 1) / threads;
    hipLaunchKernelGGL((sh_derivatives_kernel<LMAX, FULL>), dim3(blocks), dim3(threads), 0, stream,
                       coeffs, n, nc, theta, phi, out, k);
    LSD_HIP(hipGetLastError());
    return 0;
}

template <int LMAX, bool FULL>
int launch_climb_t(const double* coeffs, long n, int nc, const double* starts, int max_iters,
                   double tol, double max_step, double* dirs, double* amps, ShConstantsDev k,
                   hipStream_t stream) {
    const int threads = WAVE * WAVES_PER_BLOCK;
    const long blocks = (n + threads - 1) / threads;
    hipLaunchKernelGGL((sh_climb_kernel<LMAX, FULL>), dim3(blocks), dim3(threads), 0, stream,
                       coeffs, n, nc, starts, max_iters, tol, max_step, dirs, amps, k);
    LSD_HIP(hipGetLastError());
    return 0;
}

// X(LMAX, FULL) over every instantiated layout.
#define LSD_FOR_EACH_SH(X)                                                                     \
    X(0, false) X(2, false) X(4, false) X(6, false) X(8, false) X(10, false) X(12, false)        \
    X(1, true) X(2, true) X(3, true) X(4, true) X(5, true) X(6, true) X(7, true) X(8, true)

} // namespace

bool sh_supported(int lmax, bool full) {
    return lmax >= 0 && lmax <= SH_WIDE_LMAX && (full || lmax % 2 == 0);
}

// Each launcher: the exact instantiation when there is one, else the lmax-20 one with the
// series' lmax in the constants.
int launch_peaks(int lmax, bool full, const PeaksArgs& a_in, int waves, hipStream_t stream) {
    PeaksArgs a = a_in;
    a.k.lmax = lmax;
#define LSD_TRY(L, F)                                                                          \
    if (lmax == L && full == F) return launch_peaks_t<L, F>(a, waves, stream);
    LSD_FOR_EACH_SH(LSD_TRY)
#undef LSD_TRY
    if (!sh_supported(lmax, full)) return fail("no peak kernel for lmax " + std::to_string(lmax));
    return full ? launch_peaks_t<SH_WIDE_LMAX, true>(a, waves, stream)
                : launch_peaks_t<SH_WIDE_LMAX, false>(a, waves, stream);
}

int launch_sh_derivatives(int lmax, bool full, const double* coeffs, long n, int nc,
                          const double* theta, const double* phi, double* out, ShConstantsDev k,
                          hipStream_t stream) {
    k.lmax = lmax;
#define LSD_TRY(L, F)                                                                          \
    if (lmax == L && full == F)                                                                \
        return launch_derivs_t<L, F>(coeffs, n, nc, theta, phi, out, k, stream);
    LSD_FOR_EACH_SH(LSD_TRY)
#undef LSD_TRY
    if (!sh_supported(lmax, full)) return fail("no derivative kernel for lmax " + std::to_string(lmax));
    return full ? launch_derivs_t<SH_WIDE_LMAX, true>(coeffs, n, nc, theta, phi, out, k, stream)
                : launch_derivs_t<SH_WIDE_LMAX, false>(coeffs, n, nc, theta, phi, out, k, stream);
}

int launch_sh_climb(int lmax, bool full, const double* coeffs, long n, int nc,
                    const double* starts, int max_iters, double tol, double max_step,
                    double* dirs, double* amps, ShConstantsDev k, hipStream_t stream) {
    k.lmax = lmax;
#define LSD_TRY(L, F)                                                                          \
    if (lmax == L && full == F)                                                                \
        return launch_climb_t<L, F>(coeffs, n, nc, starts, max_iters, tol, max_step, dirs,     \
                                    amps, k, stream);
    LSD_FOR_EACH_SH(LSD_TRY)
#undef LSD_TRY
    if (!sh_supported(lmax, full)) return fail("no climb kernel for lmax " + std::to_string(lmax));
    return full ? launch_climb_t<SH_WIDE_LMAX, true>(coeffs, n, nc, starts, max_iters, tol,
                                                     max_step, dirs, amps, k, stream)
                : launch_climb_t<SH_WIDE_LMAX, false>(coeffs, n, nc, starts, max_iters, tol,
                                                      max_step, dirs, amps, k, stream);
}

} // namespace lsd
