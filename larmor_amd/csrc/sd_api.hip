// Host side of liblarmor_sd.so: the device-resident solver and kernel objects and the C-ABI
// entry points declared in include/larmor_sd.h, which replace the spherical-deconvolution
// bindings of larmor's compiled core (cpp/src/python/sd.cpp).


#include "csd_wave.hpp"
#include "kernels.hpp"
#include "objects.hpp"
#include "drl_wave.hpp"
#include "lasso_wave.hpp"
#include "lsi_host.hpp"
#include "lsi_wave.hpp"
#include "peaks_wave.hpp"

#include <atomic>
#include <cmath>
#include <map>
#include <mutex>

namespace lsd {

thread_local std::string g_error;
bool g_profiling = false;

// The tuned low-resolution factor capacity per row class (RPL 4, 6, 8, 12, 16); 0 means the
// default.
static std::atomic<int> g_lr_capacity[5] = {0, 0, 0, 0, 0};

static int rpl_slot(int rpl) {
    switch (rpl) {
    case 4: return 0;
    case 6: return 1;
    case 8: return 2;
    case 12: return 3;
    case 16: return 4;
    default: return -1;
    }
}

int lr_capacity(int rpl) {
    const int slot = rpl_slot(rpl);
    const int v = slot < 0 ? 0 : g_lr_capacity[slot].load(std::memory_order_relaxed);
    return v == 0 ? 20 : v;
}
std::vector<ProfiledLaunch> g_launches;
unsigned long long* g_work_slots = nullptr;

} // namespace lsd

using namespace lsd;


namespace lsd {

template <class T>
int upload(T** dst, const std::vector<T>& src) {
    LSD_HIP(hipMalloc(reinterpret_cast<void**>(dst), std::max<size_t>(8, src.size() * sizeof(T))));
    if (!src.empty()) {
        LSD_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    }
    return 0;
}

// The FP32 copy of an FP64 host array, uploaded.
int upload_f32(float** dst, const std::vector<double>& src) {
    return upload(dst, std::vector<float>(src.begin(), src.end()));
}

// CSD launcher (inst_csd.hip).
int csd_direction_slots(int64_t k);
bool csd_supported_coefficients(int64_t c);
int launch_csd(int n, int nk, const CsdArgs& args, int waves, hipStream_t stream);

// SH projection (inst_sh.hip).
int launch_sh_project(const double* amps, long n, int k, const double* p, int c,
                      int exact_integral, double* out, hipStream_t stream);

// LSI launchers (inst_lsi.hip).
int lsi_constraint_slots(int64_t k);
int launch_lsi(int nks, int ps, const LsiArgs& fast, const LsiArgs& qr, int waves, int qr_waves,
               hipStream_t stream);
// Waves (one per block) of the LSI QR path; with up to 63 parameters each keeps its column slab
// in LDS, with more in a global slab (128 KB or 512 KB), so fewer run.
constexpr int LSI_QR_WAVES = 2048;
constexpr int LSI_QR_WAVES_WIDE = 512;
// Waves of the fast LSI path with more than 63 parameters (each with a factor slab of 132 KB or
// 526 KB).
constexpr int LSI_WAVES_WIDE = 2048;

// Peak, derivative and climb launchers (inst_peaks.hip).
bool sh_supported(int lmax, bool full);
int launch_peaks(int lmax, bool full, const PeaksArgs& a, int waves, hipStream_t stream);
int launch_sh_derivatives(int lmax, bool full, const double* coeffs, long n, int nc,
                          const double* theta, const double* phi, double* out, ShConstantsDev k,
                          hipStream_t stream);
int launch_sh_climb(int lmax, bool full, const double* coeffs, long n, int nc,
                    const double* starts, int max_iters, double tol, double max_step,
                    double* dirs, double* amps, ShConstantsDev k, hipStream_t stream);

constexpr int SH_MAX_LMAX = 20;

// The Legendre recurrence and ladder constants up to lmax 20 on the current device, made once per
// device with the expressions of harmonics.cpp (RecurrenceConstants, up, down) and kept for the
// life of the process.
int sh_constants(ShConstantsDev* out) {
    static std::mutex mu;
    static std::map<int, ShConstantsDev> per_device;
    int dev = 0;
    LSD_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    auto it = per_device.find(dev);
    if (it != per_device.end()) {
        *out = it->second;
        return 0;
    }
    constexpr int T = SH_TABLE;
    std::vector<double> table(static_cast<size_t>(T + T * T + T * T + T * T + T), 0.0);
    double* diagonal = table.data();
    double* factor = diagonal + T;
    double* up = factor + T * T;
    double* down = up + T * T;
    double* offdiag = down + T * T;
    double ratio = 1.0;
    for (int m = 0; m <= SH_MAX_LMAX; ++m) {
        const double mf = static_cast<double>(m);
        if (m > 0) {
            ratio = ratio * ((2.0 * mf - 1.0) / (2.0 * mf));
        }
        const double sign = m % 2 == 1 ? -1.0 : 1.0;
        diagonal[m] = sign * std::sqrt((2.0 * mf + 1.0) / (4.0 * M_PI) * ratio);
        offdiag[m] = std::sqrt(2.0 * mf + 3.0);
        for (int l = 0; l <= SH_MAX_LMAX; ++l) {
            const double lf = static_cast<double>(l);
            factor[l * T + m] = l > m ? std::sqrt((4.0 * lf * lf - 1.0) / (lf * lf - mf * mf)) : 0.0;
            if (m <= l) {
                up[l * T + m] = std::sqrt(static_cast<double>((l - m) * (l + m + 1)));
                down[l * T + m] = std::sqrt(static_cast<double>((l + m) * (l - m + 1)));
            }
        }
    }
    double* d = nullptr;
    const int rc = upload(&d, table);
    if (rc) return rc;
    ShConstantsDev k{d, d + T, d + T + T * T, d + T + 2 * T * T, d + T + 3 * T * T};
    per_device[dev] = k;
    *out = k;
    return 0;
}

int n_sh_coeffs(int64_t lmax, bool full) {
    return static_cast<int>(full ? (lmax + 1) * (lmax + 1) : (lmax + 1) * (lmax + 2) / 2);
}

// check_lmax (cpp/src/python/sh.cpp:41-54), then the GPU's own limits.
int check_sh_layout(int64_t lmax, bool full, int64_t n_coeffs) {
    if (lmax < 0 || lmax > SH_MAX_LMAX || (!full && lmax % 2 != 0)) {
        return fail("lmax must be at most " + std::to_string(SH_MAX_LMAX) +
                    ", and even unless full, got " + std::to_string(lmax));
    }
    const int expected = n_sh_coeffs(lmax, full);
    if (expected != n_coeffs) {
        return fail("lmax " + std::to_string(lmax) + " has " + std::to_string(expected) +
                    " coefficients, got " + std::to_string(n_coeffs));
    }
    return 0;
}

// The position buffers of damped Richardson-Lucy solves over lists of at most ld voxels.
struct DrlBuffers {
    long ld = 0;
    double* a = nullptr;
    double* ks = nullptr;
    double* f = nullptr;
    double* prev = nullptr;
    double* pred = nullptr;
};

int drl_buffers(const lsd_solver* solver, long ld, Workspace& mem, DrlBuffers& b) {
    b.ld = ld;
    const long cl = static_cast<long>(solver->n) * ld;
    int rc = mem.get(&b.a, static_cast<long>(solver->m) * ld);
    if (!rc) rc = mem.get(&b.ks, cl);
    if (!rc) rc = mem.get(&b.f, cl);
    if (!rc) rc = mem.get(&b.prev, cl);
    if (!rc) rc = mem.get(&b.pred, cl);
    return rc;
}

// A damped Richardson-Lucy solve: base carries the signal inputs, the list and the outputs; the
// solver supplies the matrices and settings.
int run_drl(const char* name, const lsd_solver* solver, DrlArgs base, const DrlBuffers& b,
            hipStream_t stream) {
    DrlArgs a = base;
    a.kernel = solver->d_kernel;
    a.gram = solver->d_gram;
    a.c = static_cast<int>(solver->n);
    const double nu2 = solver->nu * solver->nu;
    const double nu4 = nu2 * nu2;
    a.nu8 = nu4 * nu4;
    a.n_iter = static_cast<long>(solver->n_iter);
    a.init = solver->init;
    a.ld = b.ld;
    a.a = b.a;
    a.ks = b.ks;
    a.f = b.f;
    a.prev = b.prev;
    a.pred = b.pred;
    return launch_fill(name, drl_kernel<DRL_COLUMNS>, a, static_cast<int>((b.ld + WAVE - 1) / WAVE),
                       0, stream);
}

// An elastic-net solve: base carries the signal inputs, the list and the outputs; the solver
// supplies the standardised design and settings.
int run_lasso(const char* name, const lsd_solver* solver, LassoArgs base, long n_items,
              hipStream_t stream) {
    const int ncpl = ncpl_for(solver->n);
    const int rpl = rpl_for(solver->m);
    if (ncpl < 0 || rpl < 0) {
        return fail("problem size " + shape_tuple(solver->m, solver->n) +
                    " exceeds 512 measurements or 640 directions");
    }
    LassoArgs a = base;
    a.columns_t = solver->d_columns_t;
    a.scales = solver->d_scales;
    a.nonzero = solver->d_nonzero;
    a.c = static_cast<int>(solver->n);
    a.l1_threshold = solver->l1_threshold;
    a.shrink = solver->shrink;
    a.tol = solver->tol;
    a.max_sweeps = static_cast<long>(solver->max_sweeps);
    a.unconverged = solver->d_unconverged;
    return launch_lasso(name, ncpl, rpl, a, static_cast<int>(std::min<long>(n_items, 1L << 30)),
                        stream);
}

} // namespace lsd

#include "fit_multi.hpp"

struct lsd_sphere {
    int64_t nv = 0;
    int64_t n_coeffs = 0;          // columns of the basis, 0 without one
    double* d_vertices = nullptr;  // (nv, 3)
    int* d_neighbours = nullptr;   // (nv, 6)
    double* d_basis_t = nullptr;   // (n_coeffs, nv)
};

struct lsd_lsi {
    int64_t m = 0, n = 0, k = 0;
    double* d_q_t = nullptr;
    double* d_back = nullptr;
    double* d_e = nullptr;
    double* d_e_t = nullptr;
    double* d_gram = nullptr;
    double* d_abs_sums = nullptr;
};

struct lsd_csd {
    int64_t m = 0;
    int64_t c = 0;
    int64_t c0 = 0;
    int64_t k = 0;
    double threshold = 0.0;
    int64_t max_iter = 0;
    double* d_initial_t = nullptr;    // (m, c0)
    double* d_forward = nullptr;      // (m, c)
    double* d_constraint = nullptr;   // (k, c)
    double* d_constraint_t = nullptr; // (c, k)
    double* d_gram = nullptr;         // (c, c)
};

// The workspace of lsd_fit_multi_tissue_noalloc: a static device array, so that the entry
// point allocates nothing.  Every device the library runs on holds its own copy for the life of
// the process; host code reaches it only through the per-device address hipGetSymbolAddress
// returns (the array's name on the host is a stand-in, not a device address).
constexpr size_t NOALLOC_WORKSPACE_BYTES = size_t{64} << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_WORKSPACE_BYTES];

// The device address of `scratch` on the current device, looked up once per device.
static int noalloc_workspace(void** out) {
    static std::mutex mu;
    static std::map<int, void*> per_device;
    int dev = 0;
    LSD_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    auto it = per_device.find(dev);
    if (it == per_device.end()) {
        void* ptr = nullptr;
        LSD_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        it = per_device.emplace(dev, ptr).first;
    }
    *out = it->second;
    return 0;
}

extern "C" {

const char* lsd_last_error(void) { return lsd::g_error.c_str(); }

int lsd_abi_version(void) { return LSD_ABI_VERSION; }

int lsd_profile_enable(int on) {
    if (on != 0 && lsd::g_work_slots == nullptr) {
        // The work counters of the profiled launches (allocated once; zero between collects).
        void* p = nullptr;
        LSD_HIP(hipMalloc(&p, sizeof(unsigned long long) * lsd::WORK_SLOTS));
        LSD_HIP(hipMemset(p, 0, sizeof(unsigned long long) * lsd::WORK_SLOTS));
        lsd::g_work_slots = static_cast<unsigned long long*>(p);
    }
    lsd::g_profiling = on != 0;
    return 0;
}

int lsd_profile_collect(char* names, int64_t names_bytes, double* ms, int64_t* counts,
                        double* work, int64_t max_entries, int64_t* n_entries) {
    // Sum the elapsed time and counted work of each kernel name over the recorded launches, in
    // first-seen order.
    std::vector<std::string> order;
    std::vector<double> total;
    std::vector<int64_t> count;
    std::vector<double> flops;
    std::vector<unsigned long long> slots;
    if (lsd::g_work_slots != nullptr && !lsd::g_launches.empty()) {
        for (auto& rec : lsd::g_launches) {
            LSD_HIP(hipEventSynchronize(rec.stop));
        }
        const size_t n_slots =
            std::min<size_t>(lsd::g_launches.size(), static_cast<size_t>(lsd::WORK_SLOTS));
        slots.resize(n_slots);
        LSD_HIP(hipMemcpy(slots.data(), lsd::g_work_slots, sizeof(unsigned long long) * n_slots,
                          hipMemcpyDeviceToHost));
        LSD_HIP(hipMemset(lsd::g_work_slots, 0, sizeof(unsigned long long) * n_slots));
    }
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
            flops.push_back(0.0);
        }
        total[at] += elapsed;
        count[at] += 1;
        if (rec.work != nullptr) {
            const size_t slot = static_cast<size_t>(rec.work - lsd::g_work_slots);
            if (slot < slots.size()) {
                flops[at] += static_cast<double>(slots[slot]);
            }
        }
        (void)hipEventDestroy(rec.start);
        (void)hipEventDestroy(rec.stop);
    }
    lsd::g_launches.clear();
    std::string joined;
    const int64_t n = std::min<int64_t>(max_entries, static_cast<int64_t>(order.size()));
    for (int64_t i = 0; i < n; ++i) {
        ms[i] = total[i];
        counts[i] = count[i];
        if (work != nullptr) {
            work[i] = flops[i];
        }
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
    // The device copy carries CORR_COLUMNS zeros after the last row: the correlation pass reads
    // whole column passes from each row.
    std::vector<double> k(static_cast<size_t>(m * n + CORR_COLUMNS), 0.0);
    std::copy(kernel, kernel + m * n, k.begin());
    int rc = upload(&s->d_kernel, k);
    if (!rc) rc = upload(&s->d_gram, gram);
    if (!rc) rc = upload_f32(&s->d_kernel_f32, k);
    if (!rc) rc = upload_f32(&s->d_gram_f32, gram);
    if (!rc && n >= 1 && n <= MFMA_TILE * MFMA_MAX_TILES && m >= 1) {
        std::vector<unsigned short> hi;
        std::vector<unsigned short> lo;
        s->bf16_tiles = pack_bf16_halves(kernel, static_cast<int>(m), static_cast<int>(n), hi, lo);
        rc = upload(&s->d_kernel_bf16_hi, hi);
        if (!rc) rc = upload(&s->d_kernel_bf16_lo, lo);
    }
    if (rc) {
        lsd_solver_destroy(s);
        return rc;
    }
    *out = s;
    return 0;
}

int lsd_solver_create_drl(const double* kernel_t, int64_t c, int64_t m, const double* gram,
                          int64_t gram_rows, int64_t gram_cols, double nu, int64_t n_iter,
                          double init, lsd_solver** out) {
    if (gram_rows != c || gram_cols != c) {
        return fail("shapes do not match: k_t " + shape_tuple(c, m) + ", k_t_k " +
                    shape_tuple(gram_rows, gram_cols));
    }
    if (n_iter < 0) {
        return fail("n_iter must be at least 0");
    }
    if (c < 0 || m < 0 || (c * m > 0 && kernel_t == nullptr) || (c > 0 && gram == nullptr)) {
        return fail("k_t must be (c, m) and k_t_k (c, c) with c, m >= 0");
    }
    auto* s = new lsd_solver();
    s->m = m;
    s->n = c;
    s->drl = true;
    s->nu = nu;
    s->n_iter = n_iter;
    s->init = init;
    // K (m, c) = k_t^T for the lane-per-voxel products, and K^T K as given, each followed by one
    // pass of zero padding.
    std::vector<double> k(static_cast<size_t>(m * c + CORR_COLUMNS), 0.0);
    for (int64_t r = 0; r < m; ++r)
        for (int64_t j = 0; j < c; ++j) k[r * c + j] = kernel_t[j * m + r];
    std::vector<double> g(static_cast<size_t>(c * c + DRL_COLUMNS), 0.0);
    std::copy(gram, gram + c * c, g.begin());
    int rc = upload(&s->d_kernel, k);
    if (!rc) rc = upload(&s->d_gram, g);
    if (rc) {
        lsd_solver_destroy(s);
        return rc;
    }
    *out = s;
    return 0;
}

int lsd_solver_is_drl(const lsd_solver* s) { return s->drl ? 1 : 0; }

int lsd_solver_create_lasso(const double* kernel, int64_t m, int64_t n, double reg, double alpha,
                            double tol, int64_t max_iter, lsd_solver** out) {
    if (!(reg >= 0.0)) {
        return fail("reg must be at least 0");
    }
    if (!(alpha > 0.0) || !(alpha <= 1.0)) {
        return fail("alpha must be in (0, 1]");
    }
    if (!(tol > 0.0)) {
        return fail("tol must be positive");
    }
    if (max_iter < 1) {
        return fail("max_iter must be at least 1");
    }
    if (m < 1 || n < 0 || (m * n > 0 && kernel == nullptr)) {
        return fail("kernel must be (m, n) with m >= 1, n >= 0");
    }
    auto* s = new lsd_solver();
    s->m = m;
    s->n = n;
    s->tol = tol;
    s->lasso = true;
    s->l1_threshold = reg * alpha;
    s->shrink = 1.0 + reg * (1.0 - alpha);
    s->max_sweeps = max_iter;
    // standardize_columns (elastic_net.cpp:25-37): each column over its root mean square, the
    // squares summed over rows in order; a column of zeros keeps scale 1 and stays out.
    std::vector<double> cols(static_cast<size_t>(std::max<int64_t>(1, n * m)), 0.0);
    std::vector<double> scales(static_cast<size_t>(std::max<int64_t>(1, n)), 1.0);
    std::vector<int> nonzero(static_cast<size_t>(std::max<int64_t>(1, n)), 0);
    for (int64_t j = 0; j < n; ++j) {
        double squares = 0.0;
        for (int64_t r = 0; r < m; ++r) {
            squares += kernel[r * n + j] * kernel[r * n + j];
        }
        const double scale = std::sqrt(squares / static_cast<double>(m));
        nonzero[j] = scale > 0.0 ? 1 : 0;
        scales[j] = nonzero[j] ? scale : 1.0;
        for (int64_t r = 0; r < m; ++r) {
            cols[j * m + r] = kernel[r * n + j] / scales[j];
        }
    }
    int rc = upload(&s->d_columns_t, cols);
    if (!rc) rc = upload(&s->d_scales, scales);
    if (!rc) {
        const hipError_t e1 = hipMalloc(reinterpret_cast<void**>(&s->d_nonzero),
                                        sizeof(int) * nonzero.size());
        const hipError_t e2 =
            e1 == hipSuccess ? hipMemcpy(s->d_nonzero, nonzero.data(), sizeof(int) * nonzero.size(),
                                         hipMemcpyHostToDevice)
                             : e1;
        const hipError_t e3 =
            e2 == hipSuccess ? hipMalloc(reinterpret_cast<void**>(&s->d_unconverged),
                                         sizeof(unsigned long long))
                             : e2;
        const hipError_t e4 =
            e3 == hipSuccess ? hipMemset(s->d_unconverged, 0, sizeof(unsigned long long)) : e3;
        if (e4 != hipSuccess) rc = hip_fail(e4, "lsd_solver_create_lasso");
    }
    if (rc) {
        lsd_solver_destroy(s);
        return rc;
    }
    *out = s;
    return 0;
}

int lsd_solver_is_lasso(const lsd_solver* s) { return s->lasso ? 1 : 0; }

int lsd_solver_n_unconverged(const lsd_solver* s, int64_t* out) {
    *out = 0;
    if (!s->lasso) {
        return 0;
    }
    unsigned long long n = 0;
    LSD_HIP(hipDeviceSynchronize());
    LSD_HIP(hipMemcpy(&n, s->d_unconverged, sizeof(n), hipMemcpyDeviceToHost));
    *out = static_cast<int64_t>(n);
    return 0;
}

void lsd_solver_destroy(lsd_solver* s) {
    if (!s) return;
    if (s->d_kernel) (void)hipFree(s->d_kernel);
    if (s->d_gram) (void)hipFree(s->d_gram);
    if (s->d_kernel_f32) (void)hipFree(s->d_kernel_f32);
    if (s->d_kernel_bf16_hi) (void)hipFree(s->d_kernel_bf16_hi);
    if (s->d_kernel_bf16_lo) (void)hipFree(s->d_kernel_bf16_lo);
    if (s->d_gram_f32) (void)hipFree(s->d_gram_f32);
    if (s->d_columns_t) (void)hipFree(s->d_columns_t);
    if (s->d_scales) (void)hipFree(s->d_scales);
    if (s->d_nonzero) (void)hipFree(s->d_nonzero);
    if (s->d_unconverged) (void)hipFree(s->d_unconverged);
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
    if (!rc) rc = upload_f32(&k->d_aniso_t_f32, at);
    if (!rc) rc = upload_f32(&k->d_iso_t_f32, it);
    if (!rc) rc = upload_f32(&k->d_w_iso_f32, wi);
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
    if (k->d_aniso_t_f32) (void)hipFree(k->d_aniso_t_f32);
    if (k->d_iso_t_f32) (void)hipFree(k->d_iso_t_f32);
    if (k->d_w_iso_f32) (void)hipFree(k->d_w_iso_f32);
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
    if (solver->drl) {
        if (n_voxels == 0) {
            return 0;
        }
        Workspace mem(stream);
        int rc = 0;
        if (status == nullptr) rc = mem.get(&status, n_voxels);
        if (rc) return rc;
        hipLaunchKernelGGL(init_loop_state, dim3((n_voxels + 255) / 256), dim3(256), 0, stream,
                           nullptr, 0L, nullptr, 0L, nullptr, status, static_cast<long>(n_voxels));
        DrlBuffers b;
        rc = drl_buffers(solver, n_voxels, mem, b);
        if (rc) return rc;
        DrlArgs da{};
        da.signals = signals;
        da.m = static_cast<int>(m);
        da.multi = 0;
        da.n_vox = n_voxels;
        da.x_out = fods;
        da.status = status;
        return run_drl("drl_single", solver, da, b, stream);
    }
    if (solver->lasso) {
        if (n_voxels == 0) {
            return 0;
        }
        Workspace mem(stream);
        int rc = 0;
        if (status == nullptr) rc = mem.get(&status, n_voxels);
        if (rc) return rc;
        hipLaunchKernelGGL(init_loop_state, dim3((n_voxels + 255) / 256), dim3(256), 0, stream,
                           nullptr, 0L, nullptr, 0L, nullptr, status, static_cast<long>(n_voxels));
        LassoArgs la{};
        la.signals = signals;
        la.m = static_cast<int>(m);
        la.multi = 0;
        la.n_vox = n_voxels;
        la.x_out = fods;
        la.status = status;
        return run_lasso("lasso_single", solver, la, n_voxels, stream);
    }
    const int ncpl = ncpl_for(solver->n);
    const int rpl = rpl_for_wide(m);
    if (ncpl < 0 || rpl < 0) {
        return fail("problem size " + shape_tuple(m, solver->n) +
                    " exceeds 1024 measurements or 640 directions");
    }
    if (n_voxels == 0) {
        return 0;
    }
    Workspace mem(stream);
    int* counters = nullptr;
    int* olist = nullptr;
    double* scratch = nullptr;
    int rc = mem.get(&counters, 2);
    if (!rc) rc = mem.get(&olist, n_voxels);
    if (!rc) rc = mem.get(&scratch, 1);
    if (!rc && status == nullptr) rc = mem.get(&status, n_voxels);
    if (rc) return rc;
    hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, stream, counters, 2);
    hipLaunchKernelGGL(init_loop_state, dim3((n_voxels + 255) / 256), dim3(256), 0, stream,
                       nullptr, 0L, nullptr, 0L, nullptr, status, static_cast<long>(n_voxels));
    FodArgs fa{};
    fa.signals = signals;
    fa.m = static_cast<int>(m);
    fa.multi = 0;
    fa.design = design_of(solver);
    fa.tol = solver->tol;
    fa.x_out = fods;
    fa.status = status;
    fa.vs = VoxelSet{n_voxels, nullptr, nullptr, olist, counters};
    fa.scratch = SolveScratch{scratch, 0};
    // The correlations K^T s of every voxel by the lane-per-voxel pass (the design's QR-only
    // solves do not use them).
    if (fa.design.normal_equations != 0) {
        double* corr = nullptr;
        rc = mem.get(&corr, std::max(1L, static_cast<long>(n_voxels) * fa.design.n));
        if (rc) return rc;
        CorrArgs ca{};
        ca.signals = signals;
        ca.m = static_cast<int>(m);
        ca.multi = 0;
        ca.kernel = fa.design.kernel;
        ca.n = fa.design.n;
        ca.n_vox = n_voxels;
        ca.corr = corr;
        rc = launch_correlations<CORR_COLUMNS>("fod_single_corr", ca,
                                               static_cast<int>((n_voxels + WAVE - 1) / WAVE), stream);
        if (rc) return rc;
        fa.corr = corr;
    }
    FodArgs rerun = fa;
    rerun.corr = nullptr;
    RerunScratch rerun_scratch = mem.rest();
    rc = dispatch_fod("fod_single", "fod_single_rerun", ncpl, rpl, fa, rerun, main_waves(n_voxels),
                      olist, counters, stream, &rerun_scratch);
    mem.adopt(rerun_scratch.base);
    return rc;
}

int lsd_fit_multi_tissue(const double* signals, int64_t n_voxels, int64_t m, const double* weights,
                         const lsd_tissue_kernel* lr, const lsd_solver* lr_solver,
                         const lsd_tissue_kernel* hr, const lsd_solver* hr_solver,
                         int64_t max_rounds, double fraction_tol, double final_tol, double* fods,
                         double* fractions, int32_t* rounds, int32_t* status, void* stream_) {
    auto stream = static_cast<hipStream_t>(stream_);
    Workspace mem(stream);
    return fit_multi_tissue_in(mem, signals, n_voxels, m, weights, lr, lr_solver, hr, hr_solver,
                               max_rounds, fraction_tol, final_tol, fods, fractions, rounds,
                               status, stream);
}

int lsd_correlations(const double* signals, int64_t n_voxels, int64_t m, const double* weights,
                     const lsd_tissue_kernel* kernel, const lsd_solver* solver, const double* iso,
                     int precision, double* corr, void* stream_) {
    auto stream = static_cast<hipStream_t>(stream_);
    if (m != kernel->m || solver->m != m || solver->n != kernel->ca || solver->drl ||
        solver->lasso) {
        return fail("signals " + shape_tuple(n_voxels, m) + ", the kernel (" +
                    std::to_string(kernel->m) + " measurements, " + std::to_string(kernel->ca) +
                    " directions) and the ridge NNLS solver (" + std::to_string(solver->m) +
                    ", " + std::to_string(solver->n) + ") do not match");
    }
    if (precision != 0 && precision != 1) {
        return fail("precision must be 0 (FP64) or 1 (BF16 split), got " +
                    std::to_string(precision));
    }
    if (precision == 1 && solver->d_kernel_bf16_hi == nullptr) {
        return fail("the BF16 correlations take at most 64 directions");
    }
    if (n_voxels == 0) {
        return 0;
    }
    if (precision == 1) {
        CorrMfmaArgs ma{};
        ma.signals = signals;
        ma.m = static_cast<int>(m);
        ma.weights = weights;
        ma.w_iso = kernel->d_w_iso;
        ma.iso = iso;
        ma.ci = static_cast<int>(kernel->ci);
        ma.k_hi = solver->d_kernel_bf16_hi;
        ma.k_lo = solver->d_kernel_bf16_lo;
        ma.n = static_cast<int>(solver->n);
        ma.chunks = static_cast<int>((m + MFMA_ROWS - 1) / MFMA_ROWS);
        ma.n_vox = n_voxels;
        ma.corr = corr;
        return launch_corr_bf16(ma, solver->bf16_tiles, stream);
    }
    CorrArgs ca{};
    ca.signals = signals;
    ca.m = static_cast<int>(m);
    ca.multi = 1;
    ca.weights = weights;
    ca.w_iso = kernel->d_w_iso;
    ca.iso = iso;
    ca.ci = static_cast<int>(kernel->ci);
    ca.kernel = solver->d_kernel;
    ca.n = static_cast<int>(solver->n);
    ca.n_vox = n_voxels;
    ca.corr = corr;
    return launch_correlations<CORR_COLUMNS>("correlations", ca,
                                             static_cast<int>((n_voxels + WAVE - 1) / WAVE), stream);
}

int lsd_fit_multi_tissue_mixed(const double* signals, int64_t n_voxels, int64_t m,
                               const double* weights, const lsd_tissue_kernel* lr,
                               const lsd_solver* lr_solver, const lsd_tissue_kernel* hr,
                               const lsd_solver* hr_solver, int64_t max_rounds,
                               double fraction_tol, double final_tol, double* fods,
                               double* fractions, int32_t* rounds, int32_t* status,
                               void* stream_) {
    auto stream = static_cast<hipStream_t>(stream_);
    Workspace mem(stream);
    FitForm form;
    form.mixed = true;
    return fit_multi_tissue_in(mem, signals, n_voxels, m, weights, lr, lr_solver, hr, hr_solver,
                               max_rounds, fraction_tol, final_tol, fods, fractions, rounds,
                               status, stream, form);
}

int lsd_fit_multi_tissue_noalloc(const double* signals, int64_t n_voxels, int64_t m,
                                 const double* weights, const lsd_tissue_kernel* lr,
                                 const lsd_solver* lr_solver, const lsd_tissue_kernel* hr,
                                 const lsd_solver* hr_solver, int64_t max_rounds,
                                 double fraction_tol, double final_tol, double* fods,
                                 double* fractions, int32_t* rounds, int32_t* status,
                                 void* stream_) {
    auto stream = static_cast<hipStream_t>(stream_);
    const auto plain = [](const lsd_solver* s) { return !s->drl && !s->lasso; };
    if (!plain(lr_solver) || !plain(hr_solver)) {
        return fail("the allocation-free fit takes ridge NNLS (l2nnls) solvers only");
    }
    if (n_voxels > 0 && (fods == nullptr || fractions == nullptr)) {
        return fail("the allocation-free fit needs the fods and fractions outputs");
    }
    void* base = nullptr;
    const int rc = noalloc_workspace(&base);
    if (rc) return rc;
    Workspace mem(base, NOALLOC_WORKSPACE_BYTES, stream);
    return fit_multi_tissue_in(mem, signals, n_voxels, m, weights, lr, lr_solver, hr, hr_solver,
                               max_rounds, fraction_tol, final_tol, fods, fractions, rounds,
                               status, stream);
}

int64_t lsd_noalloc_workspace_bytes(void) { return static_cast<int64_t>(NOALLOC_WORKSPACE_BYTES); }

int lsd_tune_lr_capacity(int64_t m, int64_t capacity) {
    const int rpl = rpl_for_wide(m);
    if (rpl < 0) {
        return fail("problem size exceeds 1024 measurements");
    }
    if (capacity != 0 && capacity != 16 && capacity != 20 && capacity != 24 && capacity != 32) {
        return fail("the low-resolution factor capacity is 16, 20, 24 or 32 columns (0: the "
                    "default), got " + std::to_string(capacity));
    }
    g_lr_capacity[rpl_slot(rpl)].store(static_cast<int>(capacity), std::memory_order_relaxed);
    return 0;
}

int64_t lsd_tuned_lr_capacity(int64_t m) {
    const int rpl = rpl_for_wide(m);
    return rpl < 0 ? -1 : lr_capacity(rpl);
}

int64_t lsd_fit_scratch_bytes(int64_t m, int64_t n_lr, int64_t n_hr) {
    const int rpl = rpl_for_wide(m);
    const int64_t sizes[2] = {n_lr, n_hr};
    int64_t total = 0;
    for (const int64_t n : sizes) {
        if (n <= 0) continue;
        const int ncpl = ncpl_for(n);
        if (ncpl < 0 || rpl < 0) return -1;
        total += static_cast<int64_t>(sizeof(double)) * rerun_slab_doubles(ncpl, rpl) *
                 std::max(RERUN_WAVES, RERUN_WAVES_ALL);
    }
    return total;
}

int lsd_csd_create(const double* initial, int64_t c0, int64_t m, const double* forward,
                   int64_t c, const double* constraint, int64_t k, double norm_weight,
                   double threshold, int64_t max_iter, lsd_csd** out) {
    if (c0 < 0 || m < 0 || c < 0 || k < 0 || c0 > c) {
        return fail("shapes do not match: initial " + shape_tuple(c0, m) + ", forward " +
                    shape_tuple(m, c) + ", constraint " + shape_tuple(k, c));
    }
    if (max_iter < 0) {
        return fail("max_iter must be at least 0");
    }
    if (!csd_supported_coefficients(c)) {
        return fail("CSD on the GPU takes 1 to " + std::to_string(CSD_MAX_COEFFICIENTS) +
                    " coefficients (lmax up to 30), got " + std::to_string(c));
    }
    if (csd_direction_slots(k) < 0) {
        return fail("CSD on the GPU takes at most 512 constraint directions, got " +
                    std::to_string(k));
    }
    auto* p = new lsd_csd();
    p->m = m;
    p->c = c;
    p->c0 = c0;
    p->k = k;
    p->threshold = threshold;
    p->max_iter = max_iter;
    // make_csd_problem (csd.cpp:11-21): F^T F summed over the measurements in order, then the
    // norm weight on the diagonal.
    std::vector<double> gram(static_cast<size_t>(c * c));
    for (int64_t a = 0; a < c; ++a) {
        for (int64_t b = 0; b < c; ++b) {
            double acc = 0.0;
            for (int64_t r = 0; r < m; ++r) {
                acc = acc + forward[r * c + a] * forward[r * c + b];
            }
            gram[a * c + b] = acc;
        }
    }
    for (int64_t a = 0; a < c; ++a) {
        gram[a * c + a] = gram[a * c + a] + norm_weight;
    }
    std::vector<double> init_t(static_cast<size_t>(m * c0));
    for (int64_t i = 0; i < c0; ++i) {
        for (int64_t r = 0; r < m; ++r) {
            init_t[r * c0 + i] = initial[i * m + r];
        }
    }
    std::vector<double> con_t(static_cast<size_t>(c * k));
    for (int64_t d = 0; d < k; ++d) {
        for (int64_t j = 0; j < c; ++j) {
            con_t[j * k + d] = constraint[d * c + j];
        }
    }
    int rc = upload(&p->d_initial_t, init_t);
    if (!rc) rc = upload(&p->d_forward, std::vector<double>(forward, forward + m * c));
    if (!rc) rc = upload(&p->d_constraint, std::vector<double>(constraint, constraint + k * c));
    if (!rc) rc = upload(&p->d_constraint_t, con_t);
    if (!rc) rc = upload(&p->d_gram, gram);
    if (rc) {
        lsd_csd_destroy(p);
        return rc;
    }
    *out = p;
    return 0;
}

void lsd_csd_destroy(lsd_csd* p) {
    if (!p) return;
    for (double* d : {p->d_initial_t, p->d_forward, p->d_constraint, p->d_constraint_t, p->d_gram}) {
        if (d) (void)hipFree(d);
    }
    delete p;
}

int lsd_fit_csd(const double* signals, int64_t n_voxels, int64_t m, const lsd_csd* p,
                double* fods, uint8_t* converged, int32_t* solves, void* stream_) {
    auto stream = static_cast<hipStream_t>(stream_);
    if (m != p->m) {
        return fail("shapes do not match: signals " + shape_tuple(n_voxels, m) + ", initial " +
                    shape_tuple(p->c0, p->m) + ", forward " + shape_tuple(p->m, p->c) +
                    ", constraint " + shape_tuple(p->k, p->c));
    }
    if (n_voxels == 0) {
        return 0;
    }
    CsdArgs a{};
    a.signals = signals;
    a.n_vox = n_voxels;
    a.m = static_cast<int>(m);
    a.c0 = static_cast<int>(p->c0);
    a.initial_t = p->d_initial_t;
    a.forward = p->d_forward;
    a.constraint = p->d_constraint;
    a.constraint_t = p->d_constraint_t;
    a.gram = p->d_gram;
    a.k = static_cast<int>(p->k);
    a.threshold = p->threshold;
    a.max_iter = static_cast<int>(p->max_iter);
    a.fods = fods;
    a.converged = converged;
    a.solves = solves;
    return launch_csd(static_cast<int>(p->c), csd_direction_slots(p->k), a,
                      main_waves(n_voxels), stream);
}

int lsd_sphere_create(const double* vertices, int64_t nv, const int64_t* neighbours,
                      const double* basis, int64_t n_coeffs, lsd_sphere** out) {
    if (nv < 0 || nv > PEAK_VERTEX_SLOTS * WAVE) {
        return fail("the GPU peak search takes spheres of at most " +
                    std::to_string(PEAK_VERTEX_SLOTS * WAVE) + " vertices, got " +
                    std::to_string(nv));
    }
    std::vector<int> nb(static_cast<size_t>(nv * 6));
    for (int64_t i = 0; i < nv * 6; ++i) {
        if (neighbours[i] < 0 || neighbours[i] >= nv) {
            return fail("a neighbour index is outside the sphere");
        }
        nb[i] = static_cast<int>(neighbours[i]);
    }
    auto* sp = new lsd_sphere();
    sp->nv = nv;
    sp->n_coeffs = basis ? n_coeffs : 0;
    int rc = upload(&sp->d_vertices, std::vector<double>(vertices, vertices + nv * 3));
    if (!rc) {
        const hipError_t e = hipMalloc(reinterpret_cast<void**>(&sp->d_neighbours),
                                       std::max<size_t>(4, nb.size() * sizeof(int)));
        rc = e == hipSuccess
                 ? 0
                 : hip_fail(e, "hipMalloc(neighbours)");
        if (!rc && !nb.empty()) {
            const hipError_t e2 = hipMemcpy(sp->d_neighbours, nb.data(), nb.size() * sizeof(int),
                                            hipMemcpyHostToDevice);
            rc = e2 == hipSuccess ? 0 : hip_fail(e2, "hipMemcpy(neighbours)");
        }
    }
    if (!rc && basis) {
        std::vector<double> bt(static_cast<size_t>(nv * n_coeffs));
        for (int64_t d = 0; d < nv; ++d) {
            for (int64_t q = 0; q < n_coeffs; ++q) {
                bt[q * nv + d] = basis[d * n_coeffs + q];
            }
        }
        rc = upload(&sp->d_basis_t, bt);
    }
    if (rc) {
        lsd_sphere_destroy(sp);
        return rc;
    }
    *out = sp;
    return 0;
}

void lsd_sphere_destroy(lsd_sphere* sp) {
    if (!sp) return;
    if (sp->d_vertices) (void)hipFree(sp->d_vertices);
    if (sp->d_neighbours) (void)hipFree(sp->d_neighbours);
    if (sp->d_basis_t) (void)hipFree(sp->d_basis_t);
    delete sp;
}

int lsd_find_peaks(const double* coeffs, int64_t n_voxels, int64_t n_coeffs, int64_t lmax,
                   int full, const double* on_sphere, const lsd_sphere* sphere, int64_t n_peaks,
                   double relative_threshold, double threshold, double* directions,
                   double* amplitudes, int32_t* overflow, void* stream_) {
    auto stream = static_cast<hipStream_t>(stream_);
    int rc = check_sh_layout(lmax, full != 0, n_coeffs);
    if (rc) return rc;
    if (n_peaks < 1) {
        return fail("n_peaks must be at least 1");
    }
    if (n_peaks > WAVE) {
        return fail("the GPU peak search keeps at most 64 peaks per voxel, got " +
                    std::to_string(n_peaks));
    }
    if (on_sphere == nullptr && sphere->n_coeffs != n_coeffs) {
        return fail("the sphere has no basis for " + std::to_string(n_coeffs) +
                    " coefficients; pass on_sphere or create it with one");
    }
    if (n_voxels == 0) {
        return 0;
    }
    ShConstantsDev k{};
    rc = sh_constants(&k);
    if (rc) return rc;
    Workspace mem(stream);
    int* counter = overflow;
    if (!counter) {
        rc = mem.get(&counter, 1);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, stream, counter, 1);
    PeaksArgs a{};
    a.coeffs = coeffs;
    a.n_vox = n_voxels;
    a.n_coeffs = static_cast<int>(n_coeffs);
    a.on_sphere = on_sphere;
    a.basis_t = sphere->d_basis_t;
    a.vertices = sphere->d_vertices;
    a.neighbours = sphere->d_neighbours;
    a.nv = static_cast<int>(sphere->nv);
    a.n_peaks = static_cast<int>(n_peaks);
    a.relative_threshold = relative_threshold;
    a.threshold = threshold;
    a.k = k;
    a.directions = directions;
    a.amplitudes = amplitudes;
    a.overflow = counter;
    return launch_peaks(static_cast<int>(lmax), full != 0, a, main_waves(n_voxels), stream);
}

int lsd_sh_derivatives(const double* coeffs, int64_t n, int64_t n_coeffs, int64_t lmax, int full,
                       const double* theta, const double* phi, double* out, void* stream_) {
    int rc = check_sh_layout(lmax, full != 0, n_coeffs);
    if (rc) return rc;
    if (n == 0) return 0;
    ShConstantsDev k{};
    rc = sh_constants(&k);
    if (rc) return rc;
    return launch_sh_derivatives(static_cast<int>(lmax), full != 0, coeffs, n,
                                 static_cast<int>(n_coeffs), theta, phi, out, k,
                                 static_cast<hipStream_t>(stream_));
}

int lsd_sh_climb(const double* coeffs, int64_t n, int64_t n_coeffs, int64_t lmax, int full,
                 const double* starts, int64_t max_iters, double tol, double max_step,
                 double* directions, double* amplitudes, void* stream_) {
    int rc = check_sh_layout(lmax, full != 0, n_coeffs);
    if (rc) return rc;
    if (max_iters < 0) {
        return fail("max_iters must be at least 0");
    }
    if (n == 0) return 0;
    ShConstantsDev k{};
    rc = sh_constants(&k);
    if (rc) return rc;
    return launch_sh_climb(static_cast<int>(lmax), full != 0, coeffs, n,
                           static_cast<int>(n_coeffs), starts, static_cast<int>(max_iters), tol,
                           max_step, directions, amplitudes, k, static_cast<hipStream_t>(stream_));
}

int lsd_lsi_create(const double* a, int64_t m, int64_t n, const double* g, int64_t k,
                   int64_t g_cols, double ridge, lsd_lsi** out) {
    host::LsiHostDesign d;
    const std::string err = host::make_lsi_design(a, m, n, g, k, g_cols, ridge, d);
    if (!err.empty()) {
        return fail(err);
    }
    if (n > LSI_MAX_PARAMS) {
        return fail("the GPU LSI solve takes at most " + std::to_string(LSI_MAX_PARAMS) +
                    " parameters, got " + std::to_string(n));
    }
    if (lsi_constraint_slots(k) < 0) {
        return fail("the GPU LSI solve takes at most 512 constraints, got " + std::to_string(k));
    }
    auto* p = new lsd_lsi();
    p->m = m;
    p->n = n;
    p->k = k;
    int rc = upload(&p->d_q_t, d.q_t);
    if (!rc) rc = upload(&p->d_back, d.back);
    if (!rc) rc = upload(&p->d_e, d.e);
    if (!rc) rc = upload(&p->d_e_t, d.e_t);
    if (!rc) rc = upload(&p->d_gram, d.gram);
    if (!rc) rc = upload(&p->d_abs_sums, d.abs_sums);
    if (rc) {
        lsd_lsi_destroy(p);
        return rc;
    }
    *out = p;
    return 0;
}

void lsd_lsi_destroy(lsd_lsi* p) {
    if (!p) return;
    for (double* d : {p->d_q_t, p->d_back, p->d_e, p->d_e_t, p->d_gram, p->d_abs_sums}) {
        if (d) (void)hipFree(d);
    }
    delete p;
}

int64_t lsd_lsi_n_meas(const lsd_lsi* p) { return p->m; }
int64_t lsd_lsi_n_params(const lsd_lsi* p) { return p->n; }
int64_t lsd_lsi_n_constraints(const lsd_lsi* p) { return p->k; }

static int fit_lsi_impl(const double* signals, int64_t n_voxels, int64_t m, const lsd_lsi* p,
                        double* x, uint8_t* converged, int32_t* status, bool only_qr,
                        void* stream_);

int lsd_fit_lsi(const double* signals, int64_t n_voxels, int64_t m, const lsd_lsi* p, double* x,
                uint8_t* converged, int32_t* status, void* stream_) {
    return fit_lsi_impl(signals, n_voxels, m, p, x, converged, status, false, stream_);
}

int lsd_fit_lsi_by_qr(const double* signals, int64_t n_voxels, int64_t m, const lsd_lsi* p,
                      double* x, uint8_t* converged, int32_t* status, void* stream_) {
    return fit_lsi_impl(signals, n_voxels, m, p, x, converged, status, true, stream_);
}

static int fit_lsi_impl(const double* signals, int64_t n_voxels, int64_t m, const lsd_lsi* p,
                        double* x, uint8_t* converged, int32_t* status, bool only_qr,
                        void* stream_) {
    auto stream = static_cast<hipStream_t>(stream_);
    if (m != p->m) {
        return fail("signals " + shape_tuple(n_voxels, m) + " do not have the design's " +
                    std::to_string(p->m) + " measurements");
    }
    if (n_voxels == 0) {
        return 0;
    }
    const int ps = lsi_position_slots(p->n);
    const int waves = ps == 1 ? main_waves(n_voxels) : std::min(main_waves(n_voxels), LSI_WAVES_WIDE);
    const int qr_waves = std::min<int>(ps == 1 ? LSI_QR_WAVES : LSI_QR_WAVES_WIDE, waves);
    Workspace mem(stream);
    int* counter = nullptr;
    int* list = nullptr;
    double* factor = nullptr;
    double* qr_slabs = nullptr;
    int rc = mem.get(&counter, 1);
    if (!rc) rc = mem.get(&list, n_voxels);
    if (!rc) rc = mem.get(&factor, lsi_factor_slab(ps) * waves);
    if (!rc && ps > 1) rc = mem.get(&qr_slabs, lsi_qr_slab(ps) * qr_waves);

    if (!rc && status == nullptr) rc = mem.get(&status, n_voxels);
    if (rc) return rc;
    hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, stream, counter, 1);
    LsiArgs a{};
    a.signals = signals;
    a.n_vox = n_voxels;
    a.d = LsiDev{static_cast<int>(p->m), static_cast<int>(p->n), static_cast<int>(p->k),
                 p->d_q_t, p->d_back, p->d_e, p->d_e_t, p->d_gram, p->d_abs_sums};
    a.x_out = x;
    a.converged = converged;
    a.status = status;
    a.fallback_list = list;
    a.fallback_count = counter;
    a.scratch = factor;
    a.per_wave = lsi_factor_slab(ps);
    LsiArgs q = a;
    q.list = only_qr ? nullptr : list;
    q.list_count = only_qr ? nullptr : counter;
    q.fallback_list = nullptr;
    q.fallback_count = nullptr;
    q.scratch = qr_slabs; // null for one parameter slot: the QR path's slabs are in LDS
    q.per_wave = ps > 1 ? lsi_qr_slab(ps) : 0;
    // QR only: the fast kernel gets no voxels, the QR kernel all of them.
    if (only_qr) a.n_vox = 0;
    return launch_lsi(lsi_constraint_slots(p->k), ps, a, q, waves, qr_waves, stream);
}

int lsd_sh_project(const double* amplitudes, int64_t n, int64_t k, const double* projection,
                   int64_t c, int exact_integral, double* coeffs, void* stream) {
    if (n < 0 || k < 0 || c < 0) {
        return fail("shapes must be non-negative");
    }
    if (n == 0) {
        return 0;
    }
    return launch_sh_project(amplitudes, n, static_cast<int>(k), projection, static_cast<int>(c),
                             exact_integral, coeffs, static_cast<hipStream_t>(stream));
}

} // extern "C"
