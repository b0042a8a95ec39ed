// Instantiations and launcher of the streamline tracking kernel: det and prob for symmetric
// series of even lmax 0-12 and full series of lmax 1-8, the lmax-20 instantiations that take
// every higher lmax at run time, and the fixed-rule test algorithm.
#include "launch.hpp"
#include "track_wave.hpp"

namespace lsd {

namespace {

template <int LMAX, bool FULL, int ALGO>
int launch_track_t(const TrackArgs& a, int lanes, hipStream_t stream) {
    auto kern = track_kernel<LMAX, FULL, ALGO>;
    // Resident waves of this instantiation on the whole device.
    int per_cu = 0;
    int dev = 0;
    int cus = 0;
    LSD_HIP(hipGetDevice(&dev));
    LSD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, LAUNCH_WAVE * WAVES_PER_BLOCK,
                                                     0) != hipSuccess ||
        per_cu < 1) {
        per_cu = 1;
    }
    const long resident = static_cast<long>(per_cu) * std::max(cus, 1) * WAVES_PER_BLOCK;
    // Few seeds are spread over the device: each wave then feeds only `lanes` of its lanes from
    // the queue, so that every SIMD has streamlines to step.
    if (lanes <= 0) {
        lanes = static_cast<int>(std::min<long>(64, std::max<long>(1, (a.n_seeds + resident - 1) / resident)));
    }
    TrackArgs args = a;
    args.lanes = lanes;
    const long waves = std::min<long>(resident, std::max<long>(1, (a.n_seeds + lanes - 1) / lanes));
    return launch("track", kern, args, static_cast<int>(waves), 0, stream);
}

} // namespace

bool track_supported(int lmax, bool full) {
    return full ? (lmax >= 1 && lmax <= SH_WIDE_LMAX)
                : (lmax >= 0 && lmax <= SH_WIDE_LMAX && lmax % 2 == 0);
}

int launch_track(int algo, int lmax, bool full, const TrackArgs& a_in, int lanes,
                 hipStream_t stream) {
    TrackArgs a = a_in;
    a.k.lmax = lmax;
    if (algo == TRACK_TEST) {
        return launch_track_t<0, false, TRACK_TEST>(a, lanes, stream);
    }
#define LSD_TRY(L, F)                                                                          \
    if (lmax == L && full == F) {                                                              \
        return algo == TRACK_DET ? launch_track_t<L, F, TRACK_DET>(a, lanes, stream)            \
                                 : launch_track_t<L, F, TRACK_PROB>(a, lanes, stream);         \
    }
    LSD_TRY(0, false) LSD_TRY(2, false) LSD_TRY(4, false) LSD_TRY(6, false) LSD_TRY(8, false)
    LSD_TRY(10, false) LSD_TRY(12, false)
    LSD_TRY(1, true) LSD_TRY(2, true) LSD_TRY(3, true) LSD_TRY(4, true) LSD_TRY(5, true)
    LSD_TRY(6, true) LSD_TRY(7, true) LSD_TRY(8, true)
#undef LSD_TRY
    if (!track_supported(lmax, full)) {
        return fail("no tracking kernel for lmax " + std::to_string(lmax));
    }
    if (full) {
        return algo == TRACK_DET ? launch_track_t<SH_WIDE_LMAX, true, TRACK_DET>(a, lanes, stream)
                                 : launch_track_t<SH_WIDE_LMAX, true, TRACK_PROB>(a, lanes, stream);
    }
    return algo == TRACK_DET ? launch_track_t<SH_WIDE_LMAX, false, TRACK_DET>(a, lanes, stream)
                             : launch_track_t<SH_WIDE_LMAX, false, TRACK_PROB>(a, lanes, stream);
}

} // namespace lsd
