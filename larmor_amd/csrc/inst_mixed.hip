// The mixed-precision form's kernels: the warm-started low-resolution FOD solves (1 column slot
// per lane, every row-slot count) and the BF16 matrix-core correlations (1 to 4 column tiles).
#include "corr_mfma.hpp"
#include "kernels.hpp"

namespace lsd {

template <int RP>
int launch_fod_warm(const char* name, const char* rerun_name, const FodArgs& fast,
                    const FodArgs& rerun, int waves, int* ol, int* oc, hipStream_t stream,
                    RerunScratch* scratch) {
    return run_pair<fod_warm_family, 1, RP>(name, rerun_name, fast, rerun, waves, ol, oc, stream,
                                            scratch);
}

int dispatch_fod_warm(const char* name, const char* rerun_name, int rpl, const FodArgs& fast,
                      const FodArgs& rerun, int waves, int* ol, int* oc, hipStream_t stream,
                      RerunScratch* rerun_scratch) {
    switch (rpl) {
    case 4:
        return launch_fod_warm<4>(name, rerun_name, fast, rerun, waves, ol, oc, stream, rerun_scratch);
    case 6:
        return launch_fod_warm<6>(name, rerun_name, fast, rerun, waves, ol, oc, stream, rerun_scratch);
    case 8:
        return launch_fod_warm<8>(name, rerun_name, fast, rerun, waves, ol, oc, stream, rerun_scratch);
    default:
        return fail("no warm-start kernel for " + std::to_string(rpl) + " row slots");
    }
}

int dispatch_fraction_normal(const char* name, int rpl, const FractionArgs& args, int waves,
                             hipStream_t stream) {
    // The loop fit only (one lumped group): no QR slab, the block's isotropic columns in LDS.
    const size_t lds_per_row_slot = sizeof(double) * args.kernel.ci * WAVE;
    const bool small = fraction_columns_small(args.kernel.n_blocks, args.kernel.ci);
    switch (rpl) {
    case 4:
        return small ? launch_fill(name, fraction_kernel<1, 4, false, true, false, NQ_SMALL>, args,
                                   waves, lds_per_row_slot * 4, stream)
                     : launch_fill(name, fraction_kernel<1, 4, false, true>, args, waves,
                                   lds_per_row_slot * 4, stream);
    case 6:
        return small ? launch_fill(name, fraction_kernel<1, 6, false, true, false, NQ_SMALL>, args,
                                   waves, lds_per_row_slot * 6, stream)
                     : launch_fill(name, fraction_kernel<1, 6, false, true>, args, waves,
                                   lds_per_row_slot * 6, stream);
    case 8:
        return small ? launch_fill(name, fraction_kernel<1, 8, false, true, false, NQ_SMALL>, args,
                                   waves, lds_per_row_slot * 8, stream)
                     : launch_fill(name, fraction_kernel<1, 8, false, true>, args, waves,
                                   lds_per_row_slot * 8, stream);
    default:
        return fail("no normal-equation fraction kernel for " + std::to_string(rpl) +
                    " row slots");
    }
}

int launch_corr_bf16(const CorrMfmaArgs& args, int tiles, hipStream_t stream) {
    const int waves = static_cast<int>(std::min<long>((args.n_vox + MFMA_TILE - 1) / MFMA_TILE,
                                                      1L << 30));
    switch (tiles) {
    case 1:
        return launch_fill("fod_lr_corr_bf16", corr_bf16_kernel<1>, args, waves, 0, stream);
    case 2:
        return launch_fill("fod_lr_corr_bf16", corr_bf16_kernel<2>, args, waves, 0, stream);
    case 3:
        return launch_fill("fod_lr_corr_bf16", corr_bf16_kernel<3>, args, waves, 0, stream);
    case 4:
        return launch_fill("fod_lr_corr_bf16", corr_bf16_kernel<4>, args, waves, 0, stream);
    default:
        return fail("the BF16 correlations take 1 to 4 column tiles, got " + std::to_string(tiles));
    }
}

} // namespace lsd
