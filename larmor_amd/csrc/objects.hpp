// The device-resident objects behind the C-ABI's opaque handles (include/larmor_sd.h), shared by
// the translation units of both arithmetic forms.
#pragma once

#include <cstdint>
#include <vector>

struct lsd_solver {
    int64_t m = 0;
    int64_t n = 0;
    double ridge = 0.0;
    double tol = 0.0;
    int normal_equations = 0;
    double* d_kernel = nullptr; // K (m, n), then CORR_COLUMNS (>= DRL_COLUMNS) zeros
    double* d_gram = nullptr;   // (n, n); a dRL solver's then DRL_COLUMNS zeros
    // The plain FP32 form's copies of K (with the same zero padding) and the Gram matrix, each
    // entry the FP64 one rounded (ridge NNLS solvers).
    float* d_kernel_f32 = nullptr;
    float* d_gram_f32 = nullptr;
    // The mixed-precision form's BF16 halves of K packed for the matrix cores (ridge NNLS solvers
    // of at most 64 directions), and their number of 16-column tiles.
    unsigned short* d_kernel_bf16_hi = nullptr;
    unsigned short* d_kernel_bf16_lo = nullptr;
    int bf16_tiles = 0;
    // Damped Richardson-Lucy (FodSolver.drl) instead of ridge NNLS.
    bool drl = false;
    double nu = 0.0;
    int64_t n_iter = 0;
    double init = 0.0;
    // Elastic net (FodSolver.lasso) instead of ridge NNLS.
    bool lasso = false;
    double l1_threshold = 0.0;
    double shrink = 1.0;
    int64_t max_sweeps = 0;
    double* d_columns_t = nullptr;          // (n, m) the scaled columns
    double* d_scales = nullptr;             // (n)
    int* d_nonzero = nullptr;               // (n)
    unsigned long long* d_unconverged = nullptr;
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
    // The plain FP32 form's copies, each entry the FP64 one rounded.
    float* d_aniso_t_f32 = nullptr;
    float* d_iso_t_f32 = nullptr;
    float* d_w_iso_f32 = nullptr;
};

