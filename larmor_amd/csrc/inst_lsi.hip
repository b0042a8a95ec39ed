// Instantiations and launchers of the LSI kernels: NKS constraint slots (k <= 128, 320, 512).
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

template <int NKS>
int launch_lsi_t(const LsiArgs& fast, const LsiArgs& qr, int waves, int qr_waves,
                 hipStream_t stream) {
    int rc = launch_fill("lsi", lsi_kernel<NKS>, fast, waves,
                         sizeof(double) * LSI_DIAG_DOUBLES * WAVES_PER_BLOCK, stream);
    if (rc) return rc;
    // One wave per block, its column slab in the block's LDS.
    return launch("lsi_qr", lsi_qr_kernel<NKS>, qr, qr_waves, sizeof(double) * LSI_QR_SLAB, stream,
                  1);
}

} // namespace

int launch_lsi(int nks, const LsiArgs& fast, const LsiArgs& qr, int waves, int qr_waves,
               hipStream_t stream) {
    switch (nks) {
    case 2: return launch_lsi_t<2>(fast, qr, waves, qr_waves, stream);
    case 5: return launch_lsi_t<5>(fast, qr, waves, qr_waves, stream);
    case 8: return launch_lsi_t<8>(fast, qr, waves, qr_waves, stream);
    default: return fail("no LSI kernel for " + std::to_string(nks) + " constraint slots");
    }
}

} // namespace lsd
