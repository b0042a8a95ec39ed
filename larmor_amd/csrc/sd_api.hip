// Host side of liblarmor_sd.so: the device-resident solver and kernel objects and the C-ABI
// entry points declared in include/larmor_sd.h, which replace the spherical-deconvolution
// bindings of larmor's compiled core (cpp/src/python/sd.cpp).

#include <type_traits>

#include "kernels.hpp"

namespace lsd {

thread_local std::string g_error;
bool g_profiling = false;
std::vector<ProfiledLaunch> g_launches;

__global__ void zero_ints(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        p[i] = 0;
    }
}

std::string shape_tuple(int64_t r, int64_t c) {
    return "(" + std::to_string(r) + ", " + std::to_string(c) + ")";
}

// Which instantiation covers a problem size.
int ncpl_for(int64_t n) {
    if (n <= 64) return 1;
    if (n <= 192) return 3;
    if (n <= 320) return 5;
    if (n <= 640) return 10;
    return -1;
}
int rpl_for(int64_t m) {
    if (m <= 256) return 4;
    if (m <= 384) return 6;
    if (m <= 512) return 8;
    return -1;
}

// The launcher of a family's (NCPL, RPL) instantiation.
template <class Args>
int dispatch(int ncpl, int rpl, const Args& fast, const Args& rerun, int waves, int* ol, int* oc,
             hipStream_t stream, void** rerun_scratch) {
#define LSD_TRY(FAMILY, ARGS, NC, RP)                                                          \
    if constexpr (std::is_same_v<Args, ARGS>) {                                                \
        if (ncpl == NC && rpl == RP)                                                           \
            return launch_##FAMILY##_##NC##_##RP(fast, rerun, waves, ol, oc, stream,            \
                                                 rerun_scratch);                               \
    }
    LSD_FOR_EACH_LAUNCHER(LSD_TRY)
#undef LSD_TRY
    return fail("no kernel instantiation for " + std::to_string(ncpl) + " column slots and " +
                std::to_string(rpl) + " row slots");
}

} // namespace lsd

using namespace lsd;

struct lsd_solver {
    int64_t m = 0;
    int64_t n = 0;
    double ridge = 0.0;
    double tol = 0.0;
    int normal_equations = 0;
    double* d_kernel = nullptr;
    double* d_gram = nullptr;
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
};


namespace lsd {

RidgeDesignDev design_of(const lsd_solver* s) {
    return RidgeDesignDev{static_cast<int>(s->m), static_cast<int>(s->n), s->d_kernel, s->d_gram,
                          s->ridge, s->normal_equations};
}

FractionKernelDev fractions_of(const lsd_tissue_kernel* k) {
    return FractionKernelDev{static_cast<int>(k->m), static_cast<int>(k->ca), static_cast<int>(k->ci),
                             k->d_aniso_t, k->d_iso_t, k->d_blocks,
                             static_cast<int>(k->blocks.size() / 2)};
}

// Grid of the main launch: enough waves to fill the chip, never more than voxels.
int main_waves(int64_t n_vox) {
    int dev = 0;
    hipGetDevice(&dev);
    int cus = 0;
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (cus <= 0) {
        cus = 256;
    }
    const int64_t target = static_cast<int64_t>(cus) * 16;
    int64_t waves = std::min<int64_t>(target, std::max<int64_t>(n_vox, 1));
    waves = (waves + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK * WAVES_PER_BLOCK;
    return static_cast<int>(waves);
}


int upload(double** dst, const std::vector<double>& src) {
    LSD_HIP(hipMalloc(reinterpret_cast<void**>(dst), std::max<size_t>(8, src.size() * sizeof(double))));
    if (!src.empty()) {
        LSD_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    return 0;
}

} // namespace lsd

extern "C" {

const char* lsd_last_error(void) { return lsd::g_error.c_str(); }

int lsd_abi_version(void) { return LSD_ABI_VERSION; }

int lsd_profile_enable(int on) {
    lsd::g_profiling = on != 0;
    return 0;
}

int lsd_profile_collect(char* names, int64_t names_bytes, double* ms, int64_t* counts,
                        int64_t max_entries, int64_t* n_entries) {
    // Sum the elapsed time of each kernel name over the recorded launches, in first-seen order.
    std::vector<std::string> order;
    std::vector<double> total;
    std::vector<int64_t> count;
    for (auto& rec : lsd::g_launches) {
        LSD_HIP(hipEventSynchronize(rec.stop));
        float elapsed = 0.0f;
        LSD_HIP(hipEventElapsedTime(&elapsed, rec.start, rec.stop));
        size_t at = 0;
        while (at < order.size() && order[at] != rec.name) {
            ++at;
        }
        if (at == order.size()) {
            order.push_back(rec.name);
            total.push_back(0.0);
            count.push_back(0);
        }
        total[at] += elapsed;
        count[at] += 1;
        (void)hipEventDestroy(rec.start);
        (void)hipEventDestroy(rec.stop);
    }
    lsd::g_launches.clear();
    std::string joined;
    const int64_t n = std::min<int64_t>(max_entries, static_cast<int64_t>(order.size()));
    for (int64_t i = 0; i < n; ++i) {
        ms[i] = total[i];
        counts[i] = count[i];
        joined += order[i] + "\n";
    }
    *n_entries = n;
    if (names_bytes > 0) {
        std::strncpy(names, joined.c_str(), static_cast<size_t>(names_bytes - 1));
        names[names_bytes - 1] = '\0';
    }
    return 0;
}

int lsd_solver_create_l2nnls(const double* kernel, int64_t m, int64_t n, double ridge, double tol_x,
                             lsd_solver** out) {
    if (!(ridge >= 0.0)) {
        return fail("reg must be at least 0");
    }
    if (!(tol_x > 0.0)) {
        return fail("tol_x must be positive");
    }
    if (m < 0 || n < 0 || (m * n > 0 && kernel == nullptr)) {
        return fail("kernel must be (m, n) with m, n >= 0");
    }
    auto* s = new lsd_solver();
    s->m = m;
    s->n = n;
    s->ridge = ridge;
    s->tol = tol_x;
    // Gram K^T K + ridge^2 I with each entry a sequential sum over rows (bitwise symmetric).
    std::vector<double> gram(static_cast<size_t>(n * n));
    for (int64_t a = 0; a < n; ++a) {
        for (int64_t b = 0; b < n; ++b) {
            double acc = 0.0;
            for (int64_t i = 0; i < m; ++i) {
                acc += kernel[i * n + a] * kernel[i * n + b];
            }
            gram[a * n + b] = acc;
        }
    }
    const double r2 = ridge * ridge;
    for (int64_t a = 0; a < n; ++a) {
        gram[a * n + a] += r2;
    }
    double fro = 0.0;
    for (int64_t i = 0; i < m * n; ++i) {
        fro += kernel[i] * kernel[i];
    }
    // nnls.cpp:280-292: the normal equations with no ridge, or a well-conditioned one.
    s->normal_equations = (r2 == 0.0 || fro + r2 <= 1e8 * r2) ? 1 : 0;
    std::vector<double> k(kernel, kernel + m * n);
    int rc = upload(&s->d_kernel, k);
    if (!rc) rc = upload(&s->d_gram, gram);
    if (rc) {
        lsd_solver_destroy(s);
        return rc;
    }
    *out = s;
    return 0;
}

void lsd_solver_destroy(lsd_solver* s) {
    if (!s) return;
    if (s->d_kernel) (void)hipFree(s->d_kernel);
    if (s->d_gram) (void)hipFree(s->d_gram);
    delete s;
}

int64_t lsd_solver_n_dirs(const lsd_solver* s) { return s->n; }
int64_t lsd_solver_n_meas(const lsd_solver* s) { return s->m; }
int lsd_solver_normal_equations(const lsd_solver* s) { return s->normal_equations; }

int lsd_tissue_kernel_create(const double* k_aniso, const double* k_iso, const double* w_k_iso,
                             int64_t m, int64_t ca, int64_t ci, const int64_t* blocks,
                             int64_t n_blocks, lsd_tissue_kernel** out) {
    bool valid = n_blocks > 0;
    for (int64_t b = 0; b < n_blocks; ++b) {
        const int64_t start = blocks[2 * b];
        const int64_t stop = blocks[2 * b + 1];
        valid = valid && start >= 0 && start < stop && stop <= ca;
    }
    if (!valid) {
        std::string text = "[";
        for (int64_t b = 0; b < n_blocks; ++b) {
            text += (b ? ", " : "") + shape_tuple(blocks[2 * b], blocks[2 * b + 1]);
        }
        return fail("aniso_blocks must be non-empty ranges within 0.." + std::to_string(ca) +
                    ", got " + text + "]");
    }
    if (n_blocks + ci > NQ || 1 + ci > NQ || n_blocks > MAX_GROUPS) {
        return fail("at most " + std::to_string(MAX_GROUPS) + " anisotropic blocks and " +
                    std::to_string(NQ) +
                    " fraction columns (anisotropic blocks plus isotropic tissues) are supported");
    }
    auto* k = new lsd_tissue_kernel();
    k->m = m;
    k->ca = ca;
    k->ci = ci;
    for (int64_t b = 0; b < 2 * n_blocks; ++b) {
        k->blocks.push_back(static_cast<int>(blocks[b]));
    }
    std::vector<double> at(static_cast<size_t>(ca * m));
    for (int64_t i = 0; i < m; ++i) {
        for (int64_t c = 0; c < ca; ++c) {
            at[c * m + i] = k_aniso[i * ca + c];
        }
    }
    std::vector<double> it(static_cast<size_t>(ci * m));
    for (int64_t i = 0; i < m; ++i) {
        for (int64_t c = 0; c < ci; ++c) {
            it[c * m + i] = k_iso[i * ci + c];
        }
    }
    std::vector<double> wi(w_k_iso, w_k_iso + m * ci);
    int rc = upload(&k->d_aniso_t, at);
    if (!rc) rc = upload(&k->d_iso_t, it);
    if (!rc) rc = upload(&k->d_w_iso, wi);
    if (!rc) {
        const hipError_t e1 = hipMalloc(reinterpret_cast<void**>(&k->d_blocks),
                                        sizeof(int) * k->blocks.size());
        const hipError_t e2 =
            e1 == hipSuccess ? hipMemcpy(k->d_blocks, k->blocks.data(), sizeof(int) * k->blocks.size(),
                                         hipMemcpyHostToDevice)
                             : e1;
        if (e2 != hipSuccess) rc = hip_fail(e2, "kernel blocks upload");
    }
    if (rc) {
        lsd_tissue_kernel_destroy(k);
        return rc;
    }
    *out = k;
    return 0;
}

void lsd_tissue_kernel_destroy(lsd_tissue_kernel* k) {
    if (!k) return;
    if (k->d_aniso_t) (void)hipFree(k->d_aniso_t);
    if (k->d_iso_t) (void)hipFree(k->d_iso_t);
    if (k->d_w_iso) (void)hipFree(k->d_w_iso);
    if (k->d_blocks) (void)hipFree(k->d_blocks);
    delete k;
}

int lsd_fit_single_tissue(const double* signals, int64_t n_voxels, int64_t m,
                          const lsd_solver* solver, double* fods, int32_t* status, void* stream_) {
    auto stream = static_cast<hipStream_t>(stream_);
    if (m != solver->m) {
        return fail("signals " + shape_tuple(n_voxels, m) + " do not have the solver's " +
                    std::to_string(solver->m) + " measurements");
    }
    const int ncpl = ncpl_for(solver->n);
    const int rpl = rpl_for(m);
    if (ncpl < 0 || rpl < 0) {
        return fail("problem size " + shape_tuple(m, solver->n) + " exceeds 512 measurements or 640 directions");
    }
    if (n_voxels == 0) {
        return 0;
    }
    const int waves = main_waves(n_voxels);
    void* scratch = nullptr;
    void* rerun_scratch = nullptr;
    void* counters = nullptr;
    void* olist = nullptr;
    long per = 0;
#define SLAB_CASE(RP) if (rpl == RP) per = slab_doubles<RP, KCAP_LDS>(false);
    SLAB_CASE(4) SLAB_CASE(6) SLAB_CASE(8)
#undef SLAB_CASE
    int rc = device_alloc(&scratch, sizeof(double) * per * waves, stream);
    if (!rc) rc = device_alloc(&counters, sizeof(int) * 4, stream);
    if (!rc) rc = device_alloc(&olist, sizeof(int) * n_voxels, stream);
    if (rc) return rc;
    hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, stream, static_cast<int*>(counters), 4);
    SingleArgs fast{};
    fast.signals = signals;
    fast.m = static_cast<int>(m);
    fast.design = design_of(solver);
    fast.tol = solver->tol;
    fast.fods = fods;
    fast.status = status;
    fast.vs = VoxelSet{n_voxels, nullptr, nullptr, static_cast<int*>(olist), static_cast<int*>(counters)};
    fast.scratch = SolveScratch{static_cast<double*>(scratch), per};
    SingleArgs rerun = fast;
    rc = dispatch(ncpl, rpl, fast, rerun, waves, static_cast<int*>(olist),
                  static_cast<int*>(counters), stream, &rerun_scratch);
    (void)hipFreeAsync(scratch, stream);
    (void)hipFreeAsync(counters, stream);
    (void)hipFreeAsync(olist, stream);
    if (rerun_scratch) (void)hipFreeAsync(rerun_scratch, stream);
    return rc;
}

int lsd_fit_multi_tissue(const double* signals, int64_t n_voxels, int64_t m, const double* weights,
                         const lsd_tissue_kernel* lr, const lsd_solver* lr_solver,
                         const lsd_tissue_kernel* hr, const lsd_solver* hr_solver,
                         int64_t max_rounds, double fraction_tol, double final_tol, double* fods,
                         double* fractions, int32_t* rounds, int32_t* status, void* stream_) {
    auto stream = static_cast<hipStream_t>(stream_);
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
    const int rpl = rpl_for(m);
    if (ncpl_lr < 0 || ncpl_hr < 0 || rpl < 0) {
        return fail("problem size exceeds 512 measurements or 640 directions");
    }
    if (n_voxels == 0) {
        return 0;
    }
    const int ci = static_cast<int>(lr->ci);
    const int waves = main_waves(n_voxels);
    long per = 0;
#define SLAB_CASE(RP) if (rpl == RP) per = slab_doubles<RP, KCAP_LDS>(false);
    SLAB_CASE(4) SLAB_CASE(6) SLAB_CASE(8)
#undef SLAB_CASE
    void* scratch = nullptr;
    void* counters = nullptr;
    void* olist = nullptr;
    void* iso = nullptr;
    void* rerun1 = nullptr;
    void* rerun2 = nullptr;
    rc = device_alloc(&scratch, sizeof(double) * per * waves, stream);
    if (!rc) rc = device_alloc(&counters, sizeof(int) * 4, stream);
    if (!rc) rc = device_alloc(&olist, sizeof(int) * 2 * n_voxels, stream);
    if (!rc) rc = device_alloc(&iso, sizeof(double) * std::max(1, ci) * n_voxels, stream);
    if (rc) return rc;
    int* cnt = static_cast<int*>(counters);
    int* list1 = static_cast<int*>(olist);
    int* list2 = list1 + n_voxels;
    hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, stream, cnt, 4);

    LoopArgs la{};
    la.signals = signals;
    la.m = static_cast<int>(m);
    la.weights = weights;
    la.design = design_of(lr_solver);
    la.kernel = fractions_of(lr);
    la.w_iso = lr->d_w_iso;
    la.tol = lr_solver->tol;
    la.max_rounds = static_cast<int>(max_rounds);
    la.fraction_tol = fraction_tol;
    la.iso_out = static_cast<double*>(iso);
    la.rounds_out = rounds;
    la.status = status;
    la.vs = VoxelSet{n_voxels, nullptr, nullptr, list1, cnt};
    la.scratch = SolveScratch{static_cast<double*>(scratch), per};
    rc = dispatch(ncpl_lr, rpl, la, la, waves, list1, cnt, stream, &rerun1);

    if (!rc) {
        FinalArgs fa{};
        fa.signals = signals;
        fa.m = static_cast<int>(m);
        fa.weights = weights;
        fa.design = design_of(hr_solver);
        fa.kernel = fractions_of(hr);
        fa.w_iso = hr->d_w_iso;
        fa.tol = hr_solver->tol;
        fa.final_tol = final_tol;
        fa.iso_in = static_cast<double*>(iso);
        fa.fods = fods;
        fa.fractions = fractions;
        fa.status = status;
        fa.vs = VoxelSet{n_voxels, nullptr, nullptr, list2, cnt + 1};
        fa.scratch = SolveScratch{static_cast<double*>(scratch), per};
        rc = dispatch(ncpl_hr, rpl, fa, fa, waves, list2, cnt + 1, stream, &rerun2);
    }
    (void)hipFreeAsync(scratch, stream);
    (void)hipFreeAsync(counters, stream);
    (void)hipFreeAsync(olist, stream);
    (void)hipFreeAsync(iso, stream);
    if (rerun1) (void)hipFreeAsync(rerun1, stream);
    if (rerun2) (void)hipFreeAsync(rerun2, stream);
    return rc;
}

} // extern "C"
