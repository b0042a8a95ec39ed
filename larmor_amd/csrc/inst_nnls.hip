// One non-negative least-squares problem on one wavefront: lsd_lsqnonneg (the generic NNLS) and
// lsd_ridge_l1_nnls (ridge and L1 penalties), the reference's single-problem solvers
// (cpp/src/python/sd.cpp:209-258; nnls.cpp:258-350).
//
// Both run the column-spread Lawson-Hanson solver of the FOD fits (RidgeNnlsWave) with its
// factor and QR slab in global memory, sized for every column (KCAP = 64 NCPL >= n), so an
// active set never outgrows it:
//  - the generic NNLS takes its gradient C^T (d - C x) from the residual, each column a wave sum,
//    and fits every active set by pivoted QR (solve_nnls_with_limit, nnls.cpp:264-278);
//  - the ridge NNLS takes it from the correlations K^T d - l1 and the Gram matrix
//    K^T K + ridge^2 I, with the Cholesky factor and its pivoted-QR fallback on [K_A; ridge I]
//    (solve_ridge_nnls, nnls.cpp:294-350).  The Gram matrix, its Frobenius norm and the choice
//    of the normal equations (make_ridge_design, nnls.cpp:280-292) are made on the device: each
//    entry a sequential sum over rows, so the matrix is bitwise symmetric.
// Without a tolerance, the solve takes 10 eps max_c sum_r |C(r, c)| max(m, n), the column sums
// being wave sums (default_gradient_tolerance, nnls.cpp:109-113).
#include "kernels.hpp"

#include <cmath>

namespace lsd {

namespace {

struct SingleNnlsArgs {
    RidgeDesignDev design; // kernel c (m, n); ridge only: gram, ridge
    const double* d;       // (m)
    const double* fro;     // ridge only: sum of c's squared entries (device)
    double l1;
    double tol;
    int has_tol;
    double* slab; // factor and QR slab, slab_doubles<RPL, 64 NCPL>(true)
    double* x;    // (n)
    int* status;  // status bits, or null
};

// K^T K + ridge^2 I, one entry per thread, and (thread 0 of block 0) the sum of the squared
// entries of K, each in row (resp. storage) order.
__global__ void __launch_bounds__(256) ridge_gram_kernel(const double* k, int m, int n,
                                                         double ridge, double* gram,
                                                         double* fro) {
    const long at = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long entries = static_cast<long>(n) * n;
    if (at < entries) {
        const int a = static_cast<int>(at / n);
        const int b = static_cast<int>(at % n);
        double acc = 0.0;
        for (int i = 0; i < m; ++i) {
            acc += k[static_cast<long>(i) * n + a] * k[static_cast<long>(i) * n + b];
        }
        gram[at] = a == b ? acc + ridge * ridge : acc;
    }
    if (at == 0) {
        double s = 0.0;
        for (long i = 0; i < static_cast<long>(m) * n; ++i) {
            s += k[i] * k[i];
        }
        *fro = s;
    }
}

// The solve (one wave).  RESID: the generic NNLS; otherwise the ridge NNLS.
template <int NCPL, int RPL, bool RESID>
__global__ void __launch_bounds__(WAVE) single_nnls_kernel(SingleNnlsArgs A) {
    constexpr int KCAP = NCPL * WAVE;
    constexpr int RPLQ = (RPL * WAVE + KCAP + WAVE - 1) / WAVE;
    const int l = lane_id();
    RidgeDesignDev D = A.design;
    if constexpr (!RESID) {
        const double r2 = D.ridge * D.ridge;
        D.normal_equations = (r2 == 0.0 || *A.fro + r2 <= 1e8 * r2) ? 1 : 0;
    }
    RowVec<RPL> s;
    load_signal<RPL>(A.d, 0, D.m, s);
    auto entry = [&](int c, int t) {
        const int r = l + WAVE * t;
        return r < D.m ? D.kernel[static_cast<long>(r) * D.n + c] : 0.0;
    };
    double tol = A.tol;
    if (!A.has_tol) {
        double norm1 = 0.0;
        for (int c = 0; c < D.n; ++c) {
            const double sc = wave_sum_f<RPL>(0, D.m, [&](int t) { return fabs(entry(c, t)); });
            if (c == 0 || sc > norm1) {
                norm1 = sc;
            }
        }
        tol = 10.0 * REAL_EPS * norm1 * static_cast<double>(D.m > D.n ? D.m : D.n);
    }
    double corr[NCPL];
    if constexpr (RESID) {
        // The gradient at x = 0: C^T d, each column a wave sum (the residual is d itself).
#pragma unroll
        for (int t = 0; t < NCPL; ++t) {
            corr[t] = 0.0;
        }
        for (int c = 0; c < D.n; ++c) {
            const double gc = wave_sum_f<RPL>(0, D.m, [&](int t) { return entry(c, t) * s.v[t]; });
#pragma unroll
            for (int t = 0; t < NCPL; ++t) {
                corr[t] = (t == c / WAVE && l == c % WAVE) ? gc : corr[t];
            }
        }
    } else {
        correlations<NCPL, RPL>(D, s, A.l1, corr);
    }
    double* qrs = A.slab;
    double* L = A.slab + static_cast<long>(RPLQ * WAVE) * (KCAP + 1);
    RidgeNnlsWave<NCPL, KCAP, RPLQ, true, RESID> solver(D, tol, 3 * D.n, A.l1, L, qrs);
    const int st = solver.template run<RPL>(corr, s);
#pragma unroll
    for (int t = 0; t < NCPL; ++t) {
        const int c = l + WAVE * t;
        if (c < D.n) {
            A.x[c] = solver.x[t];
        }
    }
    if (l == 0 && A.status != nullptr) {
        *A.status = st;
    }
}

template <int NCPL, int RPL, bool RESID>
int launch_single(const SingleNnlsArgs& args, hipStream_t stream) {
    SingleNnlsArgs a = args;
    const size_t bytes = sizeof(double) * static_cast<size_t>(slab_doubles<RPL, NCPL * WAVE>(true));
    int rc = device_alloc(reinterpret_cast<void**>(&a.slab), bytes, stream);
    if (rc) return rc;
    rc = launch(RESID ? "lsqnonneg" : "ridge_l1_nnls", single_nnls_kernel<NCPL, RPL, RESID>, a,
                1, 0, stream, 1);
    const hipError_t e = hipFreeAsync(a.slab, stream);
    if (!rc && e != hipSuccess) rc = hip_fail(e, "hipFreeAsync");
    return rc;
}

// The instantiation for (n, m): columns in 1, 5 or 10 slots per lane, rows in 4, 8 or 16.
template <int NCPL, bool RESID>
int dispatch_rows(const SingleNnlsArgs& args, hipStream_t stream) {
    const int m = args.design.m;
    if (m <= 4 * WAVE) return launch_single<NCPL, 4, RESID>(args, stream);
    if (m <= 8 * WAVE) return launch_single<NCPL, 8, RESID>(args, stream);
    return launch_single<NCPL, 16, RESID>(args, stream);
}

template <bool RESID>
int dispatch_single(const SingleNnlsArgs& args, hipStream_t stream) {
    const int n = args.design.n;
    if (n <= WAVE) return dispatch_rows<1, RESID>(args, stream);
    if (n <= 5 * WAVE) return dispatch_rows<5, RESID>(args, stream);
    return dispatch_rows<10, RESID>(args, stream);
}

int check_problem(const double* c, int64_t m, int64_t n, int64_t d_len, const double* d,
                  double* x) {
    if (m < 0 || n < 0 || d_len != m) {
        return fail("c must be (m, n) and d (m,): got " + shape_tuple(m, n) + " and " +
                    std::to_string(d_len));
    }
    if (m > 16 * WAVE || n > 10 * WAVE) {
        return fail("problem size " + shape_tuple(m, n) +
                    " exceeds 1024 measurements or 640 columns");
    }
    if ((m * n > 0 && c == nullptr) || (m > 0 && d == nullptr) || (n > 0 && x == nullptr)) {
        return fail("c, d and x must be device arrays of the given sizes");
    }
    return 0;
}

} // namespace

} // namespace lsd

using namespace lsd;

extern "C" {

int lsd_lsqnonneg(const double* c, int64_t m, int64_t n, const double* d, int64_t d_len,
                  double tol_x, int has_tol, double* x, int32_t* status, void* stream_) {
    auto stream = static_cast<hipStream_t>(stream_);
    const int rc = check_problem(c, m, n, d_len, d, x);
    if (rc || n == 0) {
        return rc;
    }
    SingleNnlsArgs a{};
    a.design = RidgeDesignDev{static_cast<int>(m), static_cast<int>(n), c, nullptr, 0.0, 0};
    a.d = d;
    a.tol = tol_x;
    a.has_tol = has_tol != 0;
    a.x = x;
    a.status = status;
    return dispatch_single<true>(a, stream);
}

int lsd_ridge_l1_nnls(const double* c, int64_t m, int64_t n, const double* d, int64_t d_len,
                      double ridge, double l1, double tol_x, int has_tol, double* x,
                      int32_t* status, void* stream_) {
    auto stream = static_cast<hipStream_t>(stream_);
    int rc = check_problem(c, m, n, d_len, d, x);
    if (rc) {
        return rc;
    }
    // Written so that NaN fails too.
    if (!std::isfinite(ridge) || ridge < 0.0 || !std::isfinite(l1) || l1 < 0.0) {
        return fail("ridge and l1 must be finite and >= 0.");
    }
    if (l1 > 0.0 && ridge == 0.0) {
        return fail("l1 > 0 needs ridge > 0.");
    }
    if (n == 0) {
        return 0;
    }
    double* gram = nullptr;
    rc = device_alloc(reinterpret_cast<void**>(&gram), sizeof(double) * (n * n + 1), stream);
    if (rc) return rc;
    const long entries = n * n;
    hipLaunchKernelGGL(ridge_gram_kernel, dim3(static_cast<unsigned>((entries + 255) / 256)),
                       dim3(256), 0, stream, c, static_cast<int>(m), static_cast<int>(n), ridge,
                       gram, gram + entries);
    SingleNnlsArgs a{};
    a.design = RidgeDesignDev{static_cast<int>(m), static_cast<int>(n), c, gram, ridge, 1};
    a.d = d;
    a.fro = gram + entries;
    a.l1 = l1;
    a.tol = tol_x;
    a.has_tol = has_tol != 0;
    a.x = x;
    a.status = status;
    rc = dispatch_single<false>(a, stream);
    const hipError_t e = hipFreeAsync(gram, stream);
    if (!rc && e != hipSuccess) rc = hip_fail(e, "hipFreeAsync");
    return rc;
}

} // extern "C"
