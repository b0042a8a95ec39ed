// Instantiations and launcher of the CSD kernel: N coefficients (lmax 2, 4, 6, 8) by NK
// constraint-direction slots (k <= 128, 320, 512).
#include "csd_wave.hpp"
#include "kernels.hpp"

namespace lsd {

namespace {

template <int N, int NK>
int launch_csd_n(const CsdArgs& args, int waves, hipStream_t stream) {
    const size_t lds = sizeof(double) * csd_lds_doubles(N) * WAVES_PER_BLOCK;
    return launch_fill("csd", csd_kernel<N, NK>, args, waves, lds, stream);
}

template <int N>
int launch_csd_k(int nk, const CsdArgs& args, int waves, hipStream_t stream) {
    switch (nk) {
    case 2: return launch_csd_n<N, 2>(args, waves, stream);
    case 5: return launch_csd_n<N, 5>(args, waves, stream);
    case 8: return launch_csd_n<N, 8>(args, waves, stream);
    default: return fail("no CSD kernel for " + std::to_string(nk) + " direction slots");
    }
}

} // namespace

int csd_direction_slots(int64_t k) {
    if (k <= 128) return 2;
    if (k <= 320) return 5;
    if (k <= 512) return 8;
    return -1;
}

bool csd_supported_coefficients(int64_t c) { return c == 6 || c == 15 || c == 28 || c == 45; }

int launch_csd(int n, int nk, const CsdArgs& args, int waves, hipStream_t stream) {
    switch (n) {
    case 6: return launch_csd_k<6>(nk, args, waves, stream);
    case 15: return launch_csd_k<15>(nk, args, waves, stream);
    case 28: return launch_csd_k<28>(nk, args, waves, stream);
    case 45: return launch_csd_k<45>(nk, args, waves, stream);
    default: return fail("no CSD kernel for " + std::to_string(n) + " coefficients");
    }
}

} // namespace lsd
