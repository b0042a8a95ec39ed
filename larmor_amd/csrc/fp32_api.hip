// The plain FP32 form of the multi-tissue fit (lsd_fit_multi_tissue_fp32): the FP64 form's kernels
// and launch sequence with every value in single precision -- signals, the kernel and Gram
// copies, correlations, Cholesky factors, QR fits, fractions and outputs.
#define LSD_FP32 1
#include "fit_multi.hpp"

extern "C" {

int lsd_fit_multi_tissue_fp32(const float* signals, int64_t n_voxels, int64_t m,
                              const float* weights, const lsd_tissue_kernel* lr,
                              const lsd_solver* lr_solver, const lsd_tissue_kernel* hr,
                              const lsd_solver* hr_solver, int64_t max_rounds, double fraction_tol,
                              double final_tol, float* fods, float* fractions, int32_t* rounds,
                              int32_t* status, void* stream_) {
    auto stream = static_cast<hipStream_t>(stream_);
    lsd::Workspace mem(stream);
    return lsd32::fit_multi_tissue_in(mem, signals, n_voxels, m, weights, lr, lr_solver, hr,
                                      hr_solver, max_rounds, static_cast<float>(fraction_tol),
                                      static_cast<float>(final_tol), fods, fractions, rounds,
                                      status, stream);
}

} // extern "C"
