// Instantiations and launchers of the CSD kernels: the register form for N coefficients (lmax 2,
// 4, 6, 8) by NK constraint-direction slots (k <= 128, 320, 512), and the wide form for any other
// N (lmax 10 and up) by the same slots, with the system in LDS or, beyond the workgroup's LDS, in
// global memory.
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

// The wide form's grid: a workgroup of CSD_WIDE_WAVES waves per voxel, at most what the device
// keeps resident (and, with the system in global memory, at most 4 per CU, each with its own
// slab).
template <int NK, bool GLOBAL_SYS>
int launch_csd_wide_g(const CsdArgs& args, hipStream_t stream) {
    const size_t lds = csd_wide_lds_bytes(args.n, args.k, GLOBAL_SYS);
    auto kern = csd_wide_kernel<NK, GLOBAL_SYS>;
    int dev = 0;
    int cus = 0;
    LSD_HIP(hipGetDevice(&dev));
    LSD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    int per_cu = 0;
    LSD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, LAUNCH_WAVE * CSD_WIDE_WAVES,
                                                         lds));
    if (per_cu <= 0) {
        return fail("the CSD system of " + std::to_string(args.n) + " coefficients does not fit a workgroup");
    }
    if (GLOBAL_SYS) {
        per_cu = std::min(per_cu, 4);
    }
    const long blocks = std::min<long>(args.n_vox, static_cast<long>(per_cu) * std::max(cus, 1));
    CsdArgs a = args;
    a.sys_slabs = nullptr;
    if (GLOBAL_SYS) {
        void* slabs = nullptr;
        const int rc = device_alloc(&slabs, sizeof(double) * static_cast<size_t>(blocks) *
                                                csd_lds_doubles(args.n), stream);
        if (rc) return rc;
        a.sys_slabs = static_cast<double*>(slabs);
    }
    int rc = launch("csd_wide", kern, a, static_cast<int>(blocks) * CSD_WIDE_WAVES, lds, stream,
                    CSD_WIDE_WAVES);
    if (GLOBAL_SYS) {
        const hipError_t e = hipFreeAsync(a.sys_slabs, stream);
        if (!rc && e != hipSuccess) rc = hip_fail(e, "hipFreeAsync");
    }
    return rc;
}

template <int NK>
int launch_csd_wide_k(const CsdArgs& args, hipStream_t stream) {
    int dev = 0;
    int max_lds = 0;
    LSD_HIP(hipGetDevice(&dev));
    LSD_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    if (csd_wide_lds_bytes(args.n, args.k, false) <= static_cast<size_t>(max_lds)) {
        return launch_csd_wide_g<NK, false>(args, stream);
    }
    return launch_csd_wide_g<NK, true>(args, stream);
}

} // namespace

int csd_direction_slots(int64_t k) {
    if (k <= 128) return 2;
    if (k <= 320) return 5;
    if (k <= 512) return 8;
    return -1;
}

bool csd_supported_coefficients(int64_t c) { return c >= 1 && c <= CSD_MAX_COEFFICIENTS; }

int launch_csd(int n, int nk, const CsdArgs& args, int waves, hipStream_t stream) {
    switch (n) {
    case 6: return launch_csd_k<6>(nk, args, waves, stream);
    case 15: return launch_csd_k<15>(nk, args, waves, stream);
    case 28: return launch_csd_k<28>(nk, args, waves, stream);
    case 45: return launch_csd_k<45>(nk, args, waves, stream);
    default: break;
    }
    CsdArgs a = args;
    a.n = n;
    switch (nk) {
    case 2: return launch_csd_wide_k<2>(a, stream);
    case 5: return launch_csd_wide_k<5>(a, stream);
    case 8: return launch_csd_wide_k<8>(a, stream);
    default: return fail("no CSD kernel for " + std::to_string(nk) + " direction slots");
    }
}

} // namespace lsd
