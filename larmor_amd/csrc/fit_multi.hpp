// The fit orchestration shared by the FP64 form (namespace lsd, in sd_api.hip) and the plain FP32
// form (namespace lsd32, in inst_fp32.hip): state clearing, the launchers of the instantiated
// (column slots, row slots) pairs, and the multi-tissue fit's launch sequence.
#pragma once

#include "kernels.hpp"
#include "objects.hpp"
#ifndef LSD_FP32
#include "corr_mfma.hpp"
#endif

namespace LSD_NS {

__global__ void zero_ints(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        p[i] = 0;
    }
}

// Starts the loop state: isotropic fractions 0, previous fractions +inf, no rounds, no status.
// Clears the per-voxel state of a fit; rounds may be null (single-tissue fits have none).
__global__ void init_loop_state(real* iso, long n_iso, real* prev, long n_prev, int* rounds,
                                int* status, long n_vox) {
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n_iso) iso[i] = real(0);
    if (i < n_prev) prev[i] = __builtin_inff();
    if (i < n_vox) {
        if (rounds != nullptr) rounds[i] = 0;
        status[i] = 0;
    }
}

// The launchers of an (NCPL, RPL) instantiation.
int dispatch_fod(const char* name, const char* rerun_name, int ncpl, int rpl, const FodArgs& fast,
                 const FodArgs& rerun, int waves, int* ol, int* oc, hipStream_t stream,
                 RerunScratch* rerun_scratch) {
#define LSD_TRY(NC, RP)                                                                        \
    if (ncpl == NC && rpl == RP)                                                               \
        return launch_fod_##NC##_##RP(name, rerun_name, fast, rerun, waves, ol, oc, stream,     \
                                      rerun_scratch);
    LSD_FOR_EACH_LAUNCHER(LSD_TRY)
#undef LSD_TRY
    return fail("no kernel instantiation for " + std::to_string(ncpl) + " column slots and " +
                std::to_string(rpl) + " row slots");
}

int dispatch_fraction(const char* name, int ncpl, int rpl, const FractionArgs& args, int waves,
                      hipStream_t stream) {
#define LSD_TRY(NC, RP)                                                                        \
    if (ncpl == NC && rpl == RP) return launch_fraction_##NC##_##RP(name, args, waves, stream);
    LSD_FOR_EACH_LAUNCHER(LSD_TRY)
#undef LSD_TRY
    return fail("no kernel instantiation for " + std::to_string(ncpl) + " column slots and " +
                std::to_string(rpl) + " row slots");
}

// The device copies of a solver's and a tissue kernel's matrices in this form's arithmetic type.
#ifdef LSD_FP32
inline RidgeDesignDev design_of(const lsd_solver* s) {
    return RidgeDesignDev{static_cast<int>(s->m), static_cast<int>(s->n), s->d_kernel_f32,
                          s->d_gram_f32, static_cast<real>(s->ridge), s->normal_equations};
}

inline FractionKernelDev fractions_of(const lsd_tissue_kernel* k) {
    return FractionKernelDev{static_cast<int>(k->m), static_cast<int>(k->ca), static_cast<int>(k->ci),
                             k->d_aniso_t_f32, k->d_iso_t_f32, k->d_blocks,
                             static_cast<int>(k->blocks.size() / 2)};
}

inline const real* weighted_iso_of(const lsd_tissue_kernel* k) { return k->d_w_iso_f32; }
#else
inline RidgeDesignDev design_of(const lsd_solver* s) {
    return RidgeDesignDev{static_cast<int>(s->m), static_cast<int>(s->n), s->d_kernel, s->d_gram,
                          s->ridge, s->normal_equations};
}

inline FractionKernelDev fractions_of(const lsd_tissue_kernel* k) {
    return FractionKernelDev{static_cast<int>(k->m), static_cast<int>(k->ca), static_cast<int>(k->ci),
                             k->d_aniso_t, k->d_iso_t, k->d_blocks,
                             static_cast<int>(k->blocks.size() / 2)};
}

inline const real* weighted_iso_of(const lsd_tissue_kernel* k) { return k->d_w_iso; }
#endif

// How the multi-tissue fit runs its low-resolution rounds.  The FP64 form (mixed = false) is the
// reference's loop.  The mixed-precision form (mixed = true) runs the rounds in two phases: first
// with the correlations K^T a from the BF16 matrix cores (split halves, FP32 accumulation), then,
// from where the first phase left every voxel, with the exact FP64 correlations (the refinement)
// until the fractions settle again; both phases start each voxel's NNLS from the support of its
// previous solve, and the factor, gradients and KKT test stay FP64 throughout; the rounds'
// fraction fits go through the normal equations of their few columns.  The high-resolution solve
// and the final fractions are the FP64 form's.
struct FitForm {
    bool mixed = false;
};

// The fraction tolerance of the mixed form's BF16 phase: about what the split products' error
// (a few 2^-17 of sum |K a|) leaves the fractions able to settle to; tighter, a voxel would spend
// BF16 rounds on changes below that error.  The FP64 phase then settles to the fit's own
// tolerance.
constexpr double MIXED_BF16_FRACTION_TOL = 1e-5;

#ifndef LSD_FP32
int dispatch_fod_warm(const char* name, const char* rerun_name, int rpl, const FodArgs& fast,
                      const FodArgs& rerun, int waves, int* ol, int* oc, hipStream_t stream,
                      RerunScratch* rerun_scratch);
int launch_corr_bf16(const CorrMfmaArgs& args, int tiles, hipStream_t stream);
int dispatch_fraction_normal(const char* name, int rpl, const FractionArgs& args, int waves,
                             hipStream_t stream);
#endif

// The multi-tissue fit with its per-call buffers taken from mem.  With a fixed workspace the
// high-resolution correlations are computed inside the FOD kernel (the same in-order sums) rather
// than held per voxel, and one rerun region after the per-voxel state serves the low- and
// high-resolution solves in turn (they run one after the other on the stream).
inline int fit_multi_tissue_in(Workspace& mem, const real* signals, int64_t n_voxels, int64_t m,
                               const real* weights, const lsd_tissue_kernel* lr,
                               const lsd_solver* lr_solver, const lsd_tissue_kernel* hr,
                               const lsd_solver* hr_solver, int64_t max_rounds,
                               real fraction_tol, real final_tol, real* fods,
                               real* fractions, int32_t* rounds, int32_t* status,
                               hipStream_t stream, FitForm form = {}) {
    if (m != hr->m || lr->m != m) {
        return fail("signals " + shape_tuple(n_voxels, m) + ", weights " + std::to_string(m) +
                    " and the kernels' " + std::to_string(hr->m) + " measurements do not match");
    }
    if (lr->ci != hr->ci) {
        return fail("the two kernels have different numbers of isotropic tissues");
    }
    auto matching = [&](const char* name, const lsd_tissue_kernel* k, const lsd_solver* s) {
        if (s->n != k->ca || s->m != m) {
            return fail(std::string("the ") + name + " solver (" + std::to_string(s->n) +
                        " directions, " + std::to_string(s->m) +
                        " measurements) does not match its kernel (" + std::to_string(k->ca) +
                        " directions, " + std::to_string(m) + " measurements)");
        }
        return 0;
    };
    int rc = matching("low-resolution", lr, lr_solver);
    if (!rc) rc = matching("high-resolution", hr, hr_solver);
    if (rc) return rc;
    if (max_rounds < 0) {
        return fail("max_outer must be at least 0");
    }
    const int ncpl_lr = ncpl_for(lr->ca);
    const int ncpl_hr = ncpl_for(hr->ca);
#ifdef LSD_FP32
    const int rpl = rpl_for(m);
    const bool wide_rows = false;
#else
    // The FP64 form with ridge NNLS solvers also takes 513 to 1,024 measurements; the mixed form,
    // the allocation-free entry point and the other solvers stop at 512.
    const bool wide_rows = !form.mixed && !mem.fixed() && !lr_solver->drl && !lr_solver->lasso &&
                           !hr_solver->drl && !hr_solver->lasso;
    const int rpl = wide_rows ? rpl_for_wide(m) : rpl_for(m);
#endif
    if (ncpl_lr < 0 || ncpl_hr < 0 || rpl < 0) {
        return fail(std::string("problem size exceeds ") + (wide_rows ? "1024" : "512") +
                    " measurements or 640 directions");
    }
    if (n_voxels == 0) {
        return 0;
    }
    if (form.mixed && (lr_solver->drl || lr_solver->lasso || hr_solver->drl || hr_solver->lasso ||
                       lr_solver->d_kernel_bf16_hi == nullptr || ncpl_lr != 1)) {
        return fail("the mixed-precision form takes ridge NNLS (l2nnls) solvers and a "
                    "low-resolution design of at most 64 directions");
    }
    const long nv = static_cast<long>(n_voxels);
    const int ci = static_cast<int>(lr->ci);
    const int nlf = 1 + ci;
    const int waves = main_waves(n_voxels);
    // The mixed form's two phases of rounds.
    const int64_t phases = form.mixed ? 2 : 1;
    // counters: per round its overflow and next-list counts, then the final solve's overflow.
    const long n_counters = 2 * (phases * max_rounds + 1);
    int* counters = nullptr;
    int* olist = nullptr;
    int* lists = nullptr;
    real* scratch = nullptr;
    real* iso = nullptr;
    real* prev = nullptr;
    real* x_lr = nullptr;
    real* corr_lr = nullptr;
    rc = mem.get(&counters, n_counters);
    if (!rc) rc = mem.get(&olist, nv);
    if (!rc) rc = mem.get(&lists, 2 * nv);
    if (!rc) rc = mem.get(&scratch, 1);
    if (!rc) rc = mem.get(&iso, std::max(1L, nv * ci));
    if (!rc) rc = mem.get(&prev, nv * nlf);
    if (!rc) rc = mem.get(&x_lr, nv * lr->ca);
    if (!rc && !lr_solver->drl && !lr_solver->lasso) rc = mem.get(&corr_lr, std::max(1L, nv * lr->ca));
    if (!rc && fods == nullptr) rc = mem.get(&fods, nv * hr->ca);
    if (!rc && rounds == nullptr) rc = mem.get(&rounds, nv);
    if (!rc && status == nullptr) rc = mem.get(&status, nv);
    unsigned long long* masks = nullptr;
    if (!rc && form.mixed) rc = mem.get(&masks, nv);
    if (rc) return rc;
    if (mem.fixed()) {
        // Refuse before any launch when the rest of the workspace holds less than one block of
        // rerun slabs for either resolution's solve.
        const size_t need = sizeof(real) * WAVES_PER_BLOCK *
                            static_cast<size_t>(std::max(rerun_slab_doubles(ncpl_lr, rpl),
                                                         rerun_slab_doubles(ncpl_hr, rpl)));
        if (mem.rest().bytes < need) {
            return fail("the fit needs more than the " + std::to_string(mem.rest().bytes) +
                        " bytes left of the allocation-free workspace for its rerun slabs (" +
                        std::to_string(need) + "); pass fewer voxels per call");
        }
    }
    hipLaunchKernelGGL(zero_ints, dim3((n_counters + 255) / 256), dim3(256), 0, stream, counters,
                       static_cast<int>(n_counters));
    hipLaunchKernelGGL(init_loop_state, dim3((nv * nlf + 255) / 256), dim3(256), 0, stream, iso,
                       nv * ci, prev, nv * nlf, rounds, status, nv);
    if (masks != nullptr) {
        hipLaunchKernelGGL(zero_ints, dim3((2 * nv + 255) / 256), dim3(256), 0, stream,
                           reinterpret_cast<int*>(masks), static_cast<int>(2 * nv));
    }

    FodArgs fod{};
    fod.signals = signals;
    fod.m = static_cast<int>(m);
    fod.multi = 1;
    fod.weights = weights;
    fod.iso = iso;
    fod.ci = ci;
    fod.status = status;
    fod.scratch = SolveScratch{scratch, 0};
    FractionArgs fr{};
    fr.signals = signals;
    fr.m = static_cast<int>(m);
    fr.fraction_tol = fraction_tol;
    fr.final_tol = final_tol;
    fr.iso = iso;
    fr.prev = prev;
    fr.rounds = rounds;
    fr.status = status;
    fr.n_vox = nv;

#ifdef LSD_FP32
    if (lr_solver->drl || lr_solver->lasso || hr_solver->drl || hr_solver->lasso) {
        return fail("the FP32 form takes ridge NNLS (l2nnls) solvers only");
    }
#else
    DrlBuffers drl_lr;
    if (lr_solver->drl) {
        rc = drl_buffers(lr_solver, nv, mem, drl_lr);
        if (rc) return rc;
    }
#endif

    // The low-resolution rounds: round r fits the voxels of its list (all of them in round 0)
    // and lists those that go on to round r + 1; lists and counts stay on the device.
    const int* list = nullptr;
    const int* count = nullptr;
    // The rerun scratch: allocated by the first round's rerun and reused after (or the fixed
    // workspace's rest, shared with the high-resolution solve).
    RerunScratch lr_rerun_scratch;
    for (int64_t rr = 0; rr < phases * max_rounds && !rc; ++rr) {
        const int64_t r = rr % max_rounds;
        // The mixed form's refinement phase starts again from every voxel.
        const bool bf16_phase = form.mixed && rr < max_rounds;
        (void)bf16_phase; // the FP32 form has no mixed phases
        if (form.mixed && rr == max_rounds) {
            list = nullptr;
            count = nullptr;
        }
        int* ovf_count = counters + 2 * rr;
        int* next_count = counters + 2 * rr + 1;
        int* next_list = lists + (r % 2) * nv;
        fod.design = design_of(lr_solver);
        fod.tol = lr_solver->tol;
        fod.w_iso = weighted_iso_of(lr);
        fod.x_out = x_lr;
        fod.vs = VoxelSet{nv, list, count, olist, ovf_count};
#ifndef LSD_FP32
        if (lr_solver->drl) {
            DrlArgs da{};
            da.signals = signals;
            da.m = static_cast<int>(m);
            da.multi = 1;
            da.weights = weights;
            da.w_iso = lr->d_w_iso;
            da.iso = iso;
            da.ci = ci;
            da.list = list;
            da.list_count = count;
            da.n_vox = nv;
            da.x_out = x_lr;
            da.status = status;
            rc = run_drl("drl_lr", lr_solver, da, drl_lr, stream);
        } else if (lr_solver->lasso) {
            LassoArgs la{};
            la.signals = signals;
            la.m = static_cast<int>(m);
            la.multi = 1;
            la.weights = weights;
            la.w_iso = lr->d_w_iso;
            la.iso = iso;
            la.ci = ci;
            la.list = list;
            la.list_count = count;
            la.n_vox = nv;
            la.x_out = x_lr;
            la.status = status;
            rc = run_lasso("lasso_lr", lr_solver, la, nv, stream);
        } else
#endif
        {
            CorrArgs ca{};
            ca.signals = signals;
            ca.m = static_cast<int>(m);
            ca.multi = 1;
            ca.weights = weights;
            ca.w_iso = weighted_iso_of(lr);
            ca.iso = iso;
            ca.ci = ci;
            ca.kernel = fod.design.kernel;
            ca.n = fod.design.n;
            ca.list = list;
            ca.list_count = count;
            ca.n_vox = nv;
            ca.corr = corr_lr;
#ifndef LSD_FP32
            if (bf16_phase) {
                CorrMfmaArgs ma{};
                ma.signals = signals;
                ma.m = static_cast<int>(m);
                ma.weights = weights;
                ma.w_iso = weighted_iso_of(lr);
                ma.iso = iso;
                ma.ci = ci;
                ma.k_hi = lr_solver->d_kernel_bf16_hi;
                ma.k_lo = lr_solver->d_kernel_bf16_lo;
                ma.n = fod.design.n;
                ma.chunks = static_cast<int>((m + MFMA_ROWS - 1) / MFMA_ROWS);
                ma.list = list;
                ma.list_count = count;
                ma.n_vox = nv;
                ma.corr = corr_lr;
                rc = launch_corr_bf16(ma, lr_solver->bf16_tiles, stream);
            } else
#endif
            {
                rc = launch_correlations<CORR_COLUMNS>("fod_lr_corr", ca,
                                                       static_cast<int>((nv + WAVE - 1) / WAVE), stream);
            }
            if (rc) break;
            fod.corr = corr_lr;
            fod.active_mask = masks;
            if (lr_rerun_scratch.base == nullptr && !lr_rerun_scratch.fixed) {
                lr_rerun_scratch = mem.rest();
            }
            const bool first = lr_rerun_scratch.base == nullptr;
#ifndef LSD_FP32
            if (form.mixed) {
                rc = dispatch_fod_warm(bf16_phase ? "fod_lr_warm_bf16" : "fod_lr_warm", "fod_lr_rerun_warm",
                                       rpl, fod, fod, waves, olist, ovf_count, stream,
                                       &lr_rerun_scratch);
            } else
#endif
            {
                rc = dispatch_fod("fod_lr", "fod_lr_rerun", ncpl_lr, rpl, fod, fod, waves, olist,
                                  ovf_count, stream, &lr_rerun_scratch);
            }
            fod.corr = nullptr;
            fod.active_mask = nullptr;
            if (first) {
                mem.adopt(lr_rerun_scratch.base);
            }
        }
        if (rc) break;
        fr.x = x_lr;
        fr.kernel = fractions_of(lr);
        fr.final_fit = 0;
        fr.fraction_tol = bf16_phase ? std::max(fraction_tol, static_cast<real>(MIXED_BF16_FRACTION_TOL))
                                     : fraction_tol;
        fr.list = list;
        fr.list_count = count;
        fr.next_list = next_list;
        fr.next_count = next_count;
#ifndef LSD_FP32
        if (form.mixed) {
            rc = dispatch_fraction_normal("fraction_loop_normal", rpl, fr, waves, stream);
        } else
#endif
        {
            rc = dispatch_fraction("fraction_loop", ncpl_lr, rpl, fr, waves, stream);
        }
        list = next_list;
        count = next_count;
    }
    if (!rc) {
        int* ovf_count_hr = counters + 2 * phases * max_rounds;
        // The high-resolution FOD of every voxel, then its final fractions.
#ifndef LSD_FP32
        if (hr_solver->drl) {
            DrlBuffers drl_hr;
            rc = drl_buffers(hr_solver, nv, mem, drl_hr);
            DrlArgs da{};
            da.signals = signals;
            da.m = static_cast<int>(m);
            da.multi = 1;
            da.weights = weights;
            da.w_iso = hr->d_w_iso;
            da.iso = iso;
            da.ci = ci;
            da.n_vox = nv;
            da.x_out = fods;
            da.status = status;
            if (!rc) rc = run_drl("drl_hr", hr_solver, da, drl_hr, stream);
        } else if (hr_solver->lasso) {
            LassoArgs la{};
            la.signals = signals;
            la.m = static_cast<int>(m);
            la.multi = 1;
            la.weights = weights;
            la.w_iso = hr->d_w_iso;
            la.iso = iso;
            la.ci = ci;
            la.n_vox = nv;
            la.x_out = fods;
            la.status = status;
            rc = run_lasso("lasso_hr", hr_solver, la, nv, stream);
        } else
#endif
        {
            int* ovf_count = ovf_count_hr;
            fod.design = design_of(hr_solver);
            fod.tol = hr_solver->tol;
            fod.w_iso = weighted_iso_of(hr);
            fod.x_out = fods;
            fod.vs = VoxelSet{nv, nullptr, nullptr, olist, ovf_count};
            // The correlations of every voxel by the lane-per-voxel pass, as in the rounds (a fixed
            // workspace leaves them to the FOD kernel).
            real* corr_hr = nullptr;
            if (!mem.fixed()) rc = mem.get(&corr_hr, std::max(1L, nv * fod.design.n));
            if (!rc && corr_hr != nullptr) {
                CorrArgs ca{};
                ca.signals = signals;
                ca.m = static_cast<int>(m);
                ca.multi = 1;
                ca.weights = weights;
                ca.w_iso = weighted_iso_of(hr);
                ca.iso = iso;
                ca.ci = ci;
                ca.kernel = fod.design.kernel;
                ca.n = fod.design.n;
                ca.n_vox = nv;
                ca.corr = corr_hr;
                rc = launch_correlations<CORR_COLUMNS>("fod_hr_corr", ca,
                                                       static_cast<int>((nv + WAVE - 1) / WAVE), stream);
            }
            if (!rc) {
                fod.corr = corr_hr;
                RerunScratch rerun_scratch = mem.rest();
                rc = dispatch_fod("fod_hr", "fod_hr_rerun", ncpl_hr, rpl, fod, fod, waves, olist,
                                  ovf_count, stream, &rerun_scratch);
                fod.corr = nullptr;
                mem.adopt(rerun_scratch.base);
            }
        }
    }
    if (!rc) {
        fr.x = fods;
        fr.kernel = fractions_of(hr);
        fr.final_fit = 1;
        fr.fractions = fractions;
        fr.list = nullptr;
        fr.list_count = nullptr;
        rc = dispatch_fraction("fraction_final", ncpl_hr, rpl, fr, waves, stream);
    }
    return rc;
}

} // namespace LSD_NS
