// The batched multi-tissue fit: the voxels of up to 64 volumes, each with its own gradient
// table (measurements, weights, kernels and their Gram matrices), fitted by one launch sequence
// -- every low-resolution round is one correlation, one FOD and one fraction launch over all the
// volumes' voxels still in the loop, the high-resolution solve and the final fractions one launch
// each.  Each volume's results are bit-identical to its own lsd_fit_multi_tissue: every kernel
// computes a voxel from that voxel's row and its own volume's design, in the same operations.
//
// This is the multi-volume analogue of the reference's block-dynamic scheduling of one volume's
// voxels over threads (cpp/src/common/parallel.hpp:25-88; the per-volume loop
// src/larmor/sd/model.py:323-391, multi_tissue.cpp:28-100 per voxel): a small volume no longer
// leaves the device idle between its own launches, and a batch costs the launches of one fit.
#include "kernels.hpp"
#include "objects.hpp"

namespace lsd {

namespace {

constexpr int JOB_CHUNK = 16; // job-table entries per upload launch (kernel arguments)

struct JobChunk {
    JobStage jobs[JOB_CHUNK];
    int count;
    JobStage* dst;
};

// Writes a chunk of the job table to device memory: stream-ordered, no host synchronisation, and
// the host copy may go as soon as the launch is enqueued (kernel arguments are copied).
__global__ void write_jobs(JobChunk c) {
    if (static_cast<int>(threadIdx.x) < c.count) {
        c.dst[threadIdx.x] = c.jobs[threadIdx.x];
    }
}

int upload_jobs(const std::vector<JobStage>& host, JobStage* dst, hipStream_t stream) {
    for (size_t at = 0; at < host.size(); at += JOB_CHUNK) {
        JobChunk c{};
        c.count = static_cast<int>(std::min<size_t>(JOB_CHUNK, host.size() - at));
        for (int i = 0; i < c.count; ++i) {
            c.jobs[i] = host[at + i];
        }
        c.dst = dst + at;
        hipLaunchKernelGGL(write_jobs, dim3(1), dim3(WAVE), 0, stream, c);
        LSD_HIP(hipGetLastError());
    }
    return 0;
}

// The fast FOD launch (LDS factor) over the batch's work items, then the rerun of the voxels it
// could not finish (a flat list of batch voxel numbers) with the factor in global memory.
template <int NCPL, int RPL>
int fod_pair_batched(const char* name, const char* rerun_name, const FodArgs& fast, int waves,
                     int* olist, int* ocount, hipStream_t stream, RerunScratch* rs) {
    constexpr int KG = NCPL * WAVE;
    FodArgs a = fast;
    a.vs.overflow_list = olist;
    a.vs.overflow_count = ocount;
    int rc = launch_fast<fod_batched_family, NCPL, RPL>(name, a, waves, stream);
    if (rc) return rc;
    int rerun_waves = 0;
    rc = rerun_waves_in<fod_batched_family, NCPL, RPL, FodArgs>(rs, RERUN_WAVES, stream,
                                                                 &rerun_waves);
    if (rc) return rc;
    FodArgs r = fast;
    r.scratch = SolveScratch{static_cast<real*>(rs->base), slab_doubles<RPL, KG>(true)};
    r.vs.list = olist;
    r.vs.list_count = ocount;
    r.vs.overflow_list = nullptr;
    r.vs.overflow_count = nullptr;
    r.corr = nullptr;
    r.jobs.by_voxel = 1;
    r.jobs.counts = nullptr;
    return launch(rerun_name, fod_batched_family<NCPL, RPL, KG, true>::kernel, r, rerun_waves, 0,
                  stream);
}

template <int NCPL, int RPL, bool FINAL>
int fraction_batched(const char* name, const FractionArgs& a, int waves, hipStream_t stream) {
    const int qcols = (FINAL ? a.kernel.n_blocks : 1) + a.kernel.ci;
    // Each wave keeps its QR slab and its own copy of its job's isotropic columns.
    const size_t slab = sizeof(real) * static_cast<size_t>(RPL) * WAVE;
    if (fraction_columns_small(a.kernel.n_blocks, a.kernel.ci)) {
        return launch_fraction_blocks(name, fraction_kernel<NCPL, RPL, FINAL, false, true, NQ_SMALL>,
                                      a, waves, (qcols + a.kernel.ci) * slab, 0, stream);
    }
    return launch_fraction_blocks(name, fraction_kernel<NCPL, RPL, FINAL, false, true>, a, waves,
                                  (qcols + a.kernel.ci) * slab, 0, stream);
}

#define LSD_BATCHED_RPL(X, NC) X(NC, 4) X(NC, 6) X(NC, 8)

int dispatch_fod_batched(const char* name, const char* rerun_name, int ncpl, int rpl,
                         const FodArgs& a, int waves, int* olist, int* ocount, hipStream_t stream,
                         RerunScratch* rs) {
#define LSD_TRY(NC, RP)                                                                        \
    if (ncpl == NC && rpl == RP)                                                               \
        return fod_pair_batched<NC, RP>(name, rerun_name, a, waves, olist, ocount, stream, rs);
    LSD_BATCHED_RPL(LSD_TRY, 1)
    LSD_BATCHED_RPL(LSD_TRY, 5)
#undef LSD_TRY
    return fail("the batched fit takes designs of at most 64 low-resolution and 320 "
                "high-resolution directions (got " + std::to_string(ncpl) + " column slots)");
}

// The rounds' fits (low resolution, one column slot) and the final fits (high resolution, five).
int dispatch_fraction_batched(const char* name, int ncpl, int rpl, bool final_fit,
                              const FractionArgs& a, int waves, hipStream_t stream) {
#define LSD_TRY_LOOP(NC, RP)                                                                   \
    if (!final_fit && ncpl == NC && rpl == RP) return fraction_batched<NC, RP, false>(name, a, waves, stream);
#define LSD_TRY_FINAL(NC, RP)                                                                  \
    if (final_fit && ncpl == NC && rpl == RP) return fraction_batched<NC, RP, true>(name, a, waves, stream);
    LSD_BATCHED_RPL(LSD_TRY_LOOP, 1)
    LSD_BATCHED_RPL(LSD_TRY_FINAL, 5)
#undef LSD_TRY_LOOP
#undef LSD_TRY_FINAL
    return fail("no batched fraction kernel for " + std::to_string(ncpl) + " column slots");
}

__global__ void zero_batch_ints(int* p, long n) {
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        p[i] = 0;
    }
}

__global__ void init_batch_state(real* iso, long n_iso, real* prev, long n_prev, int* rounds,
                                 int* status, long n_vox) {
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n_iso) iso[i] = real(0);
    if (i < n_prev) prev[i] = __builtin_inff();
    if (i < n_vox) {
        rounds[i] = 0;
        status[i] = 0;
    }
}

RidgeDesignDev ridge_design_of(const lsd_solver* s) {
    return RidgeDesignDev{static_cast<int>(s->m), static_cast<int>(s->n), s->d_kernel, s->d_gram,
                          s->ridge, s->normal_equations};
}

FractionKernelDev fraction_design_of(const lsd_tissue_kernel* k) {
    return FractionKernelDev{static_cast<int>(k->m), static_cast<int>(k->ca),
                             static_cast<int>(k->ci), k->d_aniso_t, k->d_iso_t, k->d_blocks,
                             static_cast<int>(k->blocks.size() / 2)};
}

int check_jobs(const lsd_batch_job* jobs, int64_t n_jobs, int* rpl_out) {
    if (n_jobs < 1 || n_jobs > MAX_JOBS) {
        return fail("a batch holds 1 to " + std::to_string(MAX_JOBS) + " jobs, got " +
                    std::to_string(n_jobs));
    }
    const lsd_batch_job& j0 = jobs[0];
    int rpl = -1;
    for (int64_t i = 0; i < n_jobs; ++i) {
        const lsd_batch_job& j = jobs[i];
        const std::string at = "job " + std::to_string(i) + ": ";
        if (j.lr == nullptr || j.hr == nullptr || j.lr_solver == nullptr || j.hr_solver == nullptr) {
            return fail(at + "the kernels and solvers must be given");
        }
        if (j.n_voxels < 0 || (j.n_voxels > 0 && (j.signals == nullptr || j.weights == nullptr))) {
            return fail(at + "signals and weights must be device arrays of its voxels");
        }
        if (j.lr->m != j.m || j.hr->m != j.m || j.lr_solver->m != j.m || j.hr_solver->m != j.m) {
            return fail(at + "its " + std::to_string(j.m) +
                        " measurements do not match its kernels and solvers");
        }
        if (j.lr_solver->drl || j.lr_solver->lasso || j.hr_solver->drl || j.hr_solver->lasso) {
            return fail(at + "the batched fit takes ridge NNLS (l2nnls) solvers");
        }
        if (j.lr_solver->n != j.lr->ca || j.hr_solver->n != j.hr->ca) {
            return fail(at + "a solver does not match its kernel's directions");
        }
        if (j.lr->ca != j0.lr->ca || j.hr->ca != j0.hr->ca || j.lr->ci != j0.lr->ci ||
            j.hr->ci != j0.lr->ci || j.lr->blocks.size() != j0.lr->blocks.size() ||
            j.hr->blocks.size() != j0.hr->blocks.size()) {
            return fail(at + "the jobs of a batch share their direction, isotropic tissue and "
                             "block counts");
        }
        const int r = rpl_for(j.m);
        if (r < 0) {
            return fail(at + "more than 512 measurements");
        }
        if (rpl >= 0 && r != rpl) {
            return fail(at + "the jobs of a batch share their row-slot class (measurements "
                             "<= 256, <= 384 or <= 512)");
        }
        rpl = r;
    }
    *rpl_out = rpl;
    return 0;
}

} // namespace

} // namespace lsd

using namespace lsd;

extern "C" {

int lsd_fit_multi_tissue_batched(const lsd_batch_job* jobs, int64_t n_jobs, int64_t max_rounds,
                                 double fraction_tol, double final_tol, double* fods,
                                 double* fractions, int32_t* rounds, int32_t* status,
                                 void* stream_) {
    auto stream = static_cast<hipStream_t>(stream_);
    int rpl = -1;
    int rc = check_jobs(jobs, n_jobs, &rpl);
    if (rc) return rc;
    if (max_rounds < 0) {
        return fail("max_outer must be at least 0");
    }
    const lsd_batch_job& j0 = jobs[0];
    const int ncpl_lr = ncpl_for(j0.lr->ca);
    const int ncpl_hr = ncpl_for(j0.hr->ca);
    const int J = static_cast<int>(n_jobs);
    std::vector<JobStage> lr_host(J);
    std::vector<JobStage> hr_host(J);
    long nv = 0;
    for (int i = 0; i < J; ++i) {
        const lsd_batch_job& j = jobs[i];
        JobStage s{};
        s.first = nv;
        s.n_vox = j.n_voxels;
        s.signals = j.signals;
        s.m = static_cast<int>(j.m);
        s.weights = j.weights;
        s.w_iso = j.lr->d_w_iso;
        s.design = ridge_design_of(j.lr_solver);
        s.tol = j.lr_solver->tol;
        s.frac = fraction_design_of(j.lr);
        lr_host[i] = s;
        s.w_iso = j.hr->d_w_iso;
        s.design = ridge_design_of(j.hr_solver);
        s.tol = j.hr_solver->tol;
        s.frac = fraction_design_of(j.hr);
        hr_host[i] = s;
        nv += j.n_voxels;
    }
    if (nv == 0) {
        return 0;
    }
    if ((fods == nullptr || fractions == nullptr)) {
        return fail("fods and fractions must be device arrays of the batch's voxels");
    }
    Workspace mem(stream);
    const int ci = static_cast<int>(j0.lr->ci);
    const int nlf = 1 + ci;
    const long n_lr = j0.lr->ca;
    const long n_hr = j0.hr->ca;
    // counters: per round the jobs' next-list counts and the round's overflow count, then the
    // high-resolution solve's overflow count.
    const long per_round = J + 1;
    const long n_counters = per_round * max_rounds + 1;
    JobStage* d_lr = nullptr;
    JobStage* d_hr = nullptr;
    int* counters = nullptr;
    int* olist = nullptr;
    int* lists = nullptr;
    real* iso = nullptr;
    real* prev = nullptr;
    real* x_lr = nullptr;
    real* corr_lr = nullptr;
    real* corr_hr = nullptr;
    rc = mem.get(&d_lr, J);
    if (!rc) rc = mem.get(&d_hr, J);
    if (!rc) rc = mem.get(&counters, n_counters);
    if (!rc) rc = mem.get(&olist, nv);
    if (!rc) rc = mem.get(&lists, 2 * nv);
    if (!rc) rc = mem.get(&iso, std::max(1L, nv * ci));
    if (!rc) rc = mem.get(&prev, nv * nlf);
    if (!rc) rc = mem.get(&x_lr, nv * n_lr);
    if (!rc) rc = mem.get(&corr_lr, nv * n_lr);
    if (!rc) rc = mem.get(&corr_hr, nv * n_hr);
    if (!rc && rounds == nullptr) rc = mem.get(&rounds, nv);
    if (!rc && status == nullptr) rc = mem.get(&status, nv);
    if (!rc) rc = upload_jobs(lr_host, d_lr, stream);
    if (!rc) rc = upload_jobs(hr_host, d_hr, stream);
    if (rc) return rc;
    hipLaunchKernelGGL(zero_batch_ints, dim3((n_counters + 255) / 256), dim3(256), 0, stream,
                       counters, n_counters);
    hipLaunchKernelGGL(init_batch_state, dim3((nv * nlf + 255) / 256), dim3(256), 0, stream, iso,
                       nv * ci, prev, nv * nlf, rounds, status, nv);
    const int waves = main_waves(nv);
    // Correlation launches: one block per 256 items of a job, at most one partial block each.
    const int corr_waves =
        static_cast<int>(std::min<long>((nv / (WAVES_PER_BLOCK * WAVE) + J) * WAVES_PER_BLOCK,
                                        1L << 30));

    FodArgs fod{};
    fod.m = 0;
    fod.multi = 1;
    fod.iso = iso;
    fod.ci = ci;
    fod.status = status;
    fod.jobs = JobTable{d_lr, J, nullptr, 0};
    FractionArgs fr{};
    fr.fraction_tol = static_cast<real>(fraction_tol);
    fr.final_tol = static_cast<real>(final_tol);
    fr.iso = iso;
    fr.prev = prev;
    fr.rounds = rounds;
    fr.status = status;
    fr.n_vox = nv;
    fr.kernel = fraction_design_of(j0.lr);
    CorrArgs ca{};
    ca.multi = 1;
    ca.iso = iso;
    ca.ci = ci;
    ca.n_vox = nv;

    RerunScratch lr_rerun;
    const int* list = nullptr;
    const int* counts = nullptr;
    for (int64_t r = 0; r < max_rounds && !rc; ++r) {
        int* next_counts = counters + per_round * r;
        int* ovf = next_counts + J;
        int* next_list = lists + (r % 2) * nv;
        ca.n = static_cast<int>(n_lr);
        ca.list = list;
        ca.corr = corr_lr;
        ca.jobs = JobTable{d_lr, J, counts, 0};
        rc = launch_correlations<CORR_COLUMNS, true>("batch_lr_corr", ca, corr_waves, stream);
        if (rc) break;
        fod.x_out = x_lr;
        fod.corr = corr_lr;
        fod.vs = VoxelSet{nv, list, nullptr, nullptr, nullptr};
        fod.jobs = JobTable{d_lr, J, counts, 0};
        const bool first = lr_rerun.base == nullptr;
        rc = dispatch_fod_batched("batch_fod_lr", "batch_fod_lr_rerun", ncpl_lr, rpl, fod, waves,
                                  olist, ovf, stream, &lr_rerun);
        if (first) mem.adopt(lr_rerun.base);
        if (rc) break;
        fr.x = x_lr;
        fr.kernel = fraction_design_of(j0.lr);
        fr.final_fit = 0;
        fr.list = list;
        fr.next_list = next_list;
        fr.next_count = next_counts;
        fr.jobs = JobTable{d_lr, J, counts, 0};
        rc = dispatch_fraction_batched("batch_fraction_loop", ncpl_lr, rpl, false, fr, waves,
                                       stream);
        list = next_list;
        counts = next_counts;
    }
    if (!rc) {
        int* ovf = counters + per_round * max_rounds;
        ca.n = static_cast<int>(n_hr);
        ca.list = nullptr;
        ca.corr = corr_hr;
        ca.jobs = JobTable{d_hr, J, nullptr, 0};
        rc = launch_correlations<CORR_COLUMNS, true>("batch_hr_corr", ca, corr_waves, stream);
        if (!rc) {
            fod.x_out = fods;
            fod.corr = corr_hr;
            fod.vs = VoxelSet{nv, nullptr, nullptr, nullptr, nullptr};
            fod.jobs = JobTable{d_hr, J, nullptr, 0};
            RerunScratch hr_rerun;
            rc = dispatch_fod_batched("batch_fod_hr", "batch_fod_hr_rerun", ncpl_hr, rpl, fod,
                                      waves, olist, ovf, stream, &hr_rerun);
            mem.adopt(hr_rerun.base);
        }
    }
    if (!rc) {
        fr.x = fods;
        fr.kernel = fraction_design_of(j0.hr);
        fr.final_fit = 1;
        fr.fractions = fractions;
        fr.list = nullptr;
        fr.next_list = nullptr;
        fr.next_count = nullptr;
        fr.jobs = JobTable{d_hr, J, nullptr, 0};
        rc = dispatch_fraction_batched("batch_fraction_final", ncpl_hr, rpl, true, fr, waves,
                                       stream);
    }
    return rc;
}

} // extern "C"
