// Instantiations and launchers of the LSI kernels: NKS constraint slots (k <= 128, 320, 512) by
// PS parameter slots (n <= 63, 127, 255).
#include "kernels.hpp"
#include "lsi_wave.hpp"

namespace lsd {

int lsi_constraint_slots(int64_t k) {
    if (k <= 128) return 2;
    if (k <= 320) return 5;
    if (k <= 512) return 8;
    return -1;
}

namespace {

template <int NKS, int PS>
int launch_lsi_t(const LsiArgs& fast, const LsiArgs& qr, int waves, int qr_waves,
                 hipStream_t stream) {
    int rc = launch_fill("lsi", lsi_kernel<NKS, PS>, fast, waves,
                         sizeof(double) * lsi_diag_doubles(PS) * WAVES_PER_BLOCK, stream);
    if (rc) return rc;
    // One wave per block, its column slab in the block's LDS (PS = 1) or in the scratch.
    const size_t lds = PS == 1 ? sizeof(double) * lsi_qr_slab(1) : 0;
    return launch("lsi_qr", lsi_qr_kernel<NKS, PS>, qr, qr_waves, lds, stream, 1);
}

template <int NKS>
int launch_lsi_k(int ps, const LsiArgs& fast, const LsiArgs& qr, int waves, int qr_waves,
                 hipStream_t stream) {
    switch (ps) {
    case 1: return launch_lsi_t<NKS, 1>(fast, qr, waves, qr_waves, stream);
    case 2: return launch_lsi_t<NKS, 2>(fast, qr, waves, qr_waves, stream);
    case 4: return launch_lsi_t<NKS, 4>(fast, qr, waves, qr_waves, stream);
    default: return fail("no LSI kernel for " + std::to_string(ps) + " parameter slots");
    }
}

} // namespace

int launch_lsi(int nks, int ps, const LsiArgs& fast, const LsiArgs& qr, int waves, int qr_waves,
               hipStream_t stream) {
    switch (nks) {
    case 2: return launch_lsi_k<2>(ps, fast, qr, waves, qr_waves, stream);
    case 5: return launch_lsi_k<5>(ps, fast, qr, waves, qr_waves, stream);
    case 8: return launch_lsi_k<8>(ps, fast, qr, waves, qr_waves, stream);
    default: return fail("no LSI kernel for " + std::to_string(nks) + " constraint slots");
    }
}

} // namespace lsd
