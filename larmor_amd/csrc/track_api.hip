// Host side of streamline tracking: the C-ABI entry points lsd_track_calibrate and
// lsd_track_batch (include/larmor_sd.h), which replace track_batch and the algorithm parameter
// classes of larmor's compiled core (cpp/src/python/tracking.cpp), and the calibration of the
// probabilistic algorithm's rejection sampling (cpp/src/tracking/calibrate.cpp), made once per
// (lmax, layout, angle) on the host and kept on each device.

#include "launch.hpp"
#include "track_wave.hpp"

#include <array>
#include <charconv>
#include <cmath>
#include <limits>
#include <map>
#include <mutex>
#include <tuple>

namespace lsd {

int sh_constants(ShConstantsDev* out);                // sd_api.hip
bool track_supported(int lmax, bool full);            // inst_track.hip
int launch_track(int algo, int lmax, bool full, const TrackArgs& a, int lanes,
                 hipStream_t stream);

namespace {

constexpr int MAX_LMAX = 20;

// The reference binding's number text (std::to_chars, fixed; NaN and inf spelled out).
std::string number_text(double value) {
    if (std::isnan(value)) return "NaN";
    if (std::isinf(value)) return value > 0.0 ? "inf" : "-inf";
    std::array<char, 512> buffer{};
    const auto [end, ec] = std::to_chars(buffer.data(), buffer.data() + buffer.size(), value,
                                         std::chars_format::fixed);
    if (ec != std::errc()) return std::to_string(value);
    return std::string(buffer.data(), end);
}

// --- The series' amplitude on the host, for the calibration (harmonics.cpp:73-89, 187-219). ---

using Table = std::array<std::array<double, MAX_LMAX + 1>, MAX_LMAX + 1>;

struct Recurrence {
    double diagonal[MAX_LMAX + 1];
    double factor[MAX_LMAX + 1][MAX_LMAX + 1];
    Recurrence() {
        double ratio = 1.0;
        for (int m = 0; m <= MAX_LMAX; ++m) {
            const double mf = static_cast<double>(m);
            if (m > 0) ratio = ratio * ((2.0 * mf - 1.0) / (2.0 * mf));
            diagonal[m] = (m % 2 == 1 ? -1.0 : 1.0) * std::sqrt((2.0 * mf + 1.0) / (4.0 * M_PI) * ratio);
            for (int l = 0; l <= MAX_LMAX; ++l) {
                const double lf = static_cast<double>(l);
                factor[l][m] = l > m ? std::sqrt((4.0 * lf * lf - 1.0) / (lf * lf - mf * mf)) : 0.0;
            }
        }
    }
};

double power_int(double base, int exponent) {
    double result = 1.0;
    while (true) {
        if (exponent % 2 == 1) result = result * base;
        exponent /= 2;
        if (exponent == 0) break;
        base = base * base;
    }
    return result;
}

double host_amplitude(const std::vector<double>& c, int lmax, bool full, const double d[3]) {
    static const Recurrence r;
    const double xy = std::sqrt(d[0] * d[0] + d[1] * d[1]);
    Table p{};
    for (int m = 0; m <= lmax; ++m) {
        p[m][m] = r.diagonal[m] * power_int(xy, m);
        if (m < lmax) p[m + 1][m] = std::sqrt(2.0 * m + 3.0) * d[2] * p[m][m];
        for (int l = m + 2; l <= lmax; ++l) {
            p[l][m] = r.factor[l][m] * (d[2] * p[l - 1][m] - p[l - 2][m] / r.factor[l - 1][m]);
        }
    }
    auto index = [full](int l, int m) { return full ? l * l + l + m : l * (l + 1) / 2 + m; };
    double cos_phi = 1.0;
    double sin_phi = 0.0;
    if (xy > 0.0) {
        cos_phi = d[0] / xy;
        sin_phi = d[1] / xy;
    }
    const int step = full ? 1 : 2;
    double value = 0.0;
    for (int l = 0; l <= lmax; l += step) value = value + c[index(l, 0)] * p[l][0];
    double cm = 1.0;
    double sm = 0.0;
    for (int m = 1; m <= lmax; ++m) {
        const double nc = cm * cos_phi - sm * sin_phi;
        const double ns = sm * cos_phi + cm * sin_phi;
        cm = nc;
        sm = ns;
        for (int l = (full || m % 2 == 0) ? m : m + 1; l <= lmax; l += step) {
            value = value + M_SQRT2 * p[l][m] * (c[index(l, m)] * cm + c[index(l, -m)] * sm);
        }
    }
    return value;
}

// --- The calibration (calibrate.cpp:24-108). ---

long saturating_index(double value) {
    if (std::isnan(value)) return 0;
    constexpr double limit = static_cast<double>(std::numeric_limits<long>::max());
    if (value >= limit) return std::numeric_limits<long>::max();
    if (value <= -limit) return std::numeric_limits<long>::min();
    return static_cast<long>(value);
}

// A hexagonal lattice of spacing `spacing` mapped to the sphere around +z out to max_angle, in
// lattice order (rows i, then columns j); its size only, when out is null.
long direction_grid(double max_angle, double spacing, std::vector<double>* out) {
    const double radius = max_angle / spacing;
    const double radius2 = radius * radius;
    const long extent = saturating_index(std::ceil(radius));
    long n = 0;
    for (long i = -extent; i <= extent; ++i) {
        for (long j = -extent; j <= extent; ++j) {
            const double x = static_cast<double>(i) + 0.5 * static_cast<double>(j);
            const double y = 0.866025403784439 * static_cast<double>(j);
            const double r2 = x * x + y * y;
            if (r2 > radius2) continue;
            n += 1;
            if (out) {
                const double angle = spacing * std::sqrt(r2);
                const double scale = angle > 0.0 ? std::sin(angle) / std::sqrt(r2) : 0.0;
                out->push_back(scale * x);
                out->push_back(scale * y);
                out->push_back(std::cos(angle));
            }
        }
    }
    return n;
}

struct Calibration {
    std::vector<double> grid; // (n, 3)
    double ratio = 0.0;
};

// The lobe of a unit fibre along +z sampled every 1e-3 rad out to the cone's edge or its first
// non-positive amplitude; of the spacings sqrt(3) * elevation, the one with the fewest expected
// amplitude evaluations per step (draws in the cone + grid directions) is kept.
Calibration calibrate(int lmax, bool full, double max_angle) {
    std::vector<double> fibre(static_cast<size_t>(full ? (lmax + 1) * (lmax + 1)
                                                        : (lmax + 1) * (lmax + 2) / 2),
                              0.0);
    for (int l = 0; l <= lmax; l += full ? 1 : 2) {
        fibre[static_cast<size_t>(full ? l * l + l : l * (l + 1) / 2)] =
            std::sqrt(static_cast<double>(2 * l + 1) / (4.0 * M_PI));
    }
    std::vector<std::pair<double, double>> lobe;
    double elevation = 0.0;
    while (elevation < max_angle) {
        const double d[3] = {std::sin(elevation), 0.0, std::cos(elevation)};
        const double a = host_amplitude(fibre, lmax, full, d);
        lobe.emplace_back(elevation, a);
        if (!std::isfinite(a) || a <= 0.0) break;
        elevation += 0.001;
    }
    const double nan = std::numeric_limits<double>::quiet_NaN();
    Calibration out;
    if (lobe.empty()) {
        out.ratio = nan;
        return out;
    }
    const double zero = lobe.back().first;
    const double on_axis = lobe.front().second;
    double best = std::numeric_limits<double>::infinity();
    double theta = nan;
    double ratio = nan;
    for (size_t s = 1; s < lobe.size(); ++s) {
        const double e = lobe[s].first;
        const double a = lobe[s].second;
        const double n_cone = max_angle * max_angle * (1.0 + on_axis / a) / (2.0 * zero * zero);
        const double n_eval =
            n_cone + static_cast<double>(direction_grid(max_angle + e, std::sqrt(3.0) * e, nullptr));
        if (n_eval > 0.0 && n_eval < best) {
            best = n_eval;
            theta = e;
            ratio = on_axis / a;
        }
    }
    direction_grid(max_angle + theta, std::sqrt(3.0) * theta, &out.grid);
    out.ratio = ratio;
    return out;
}

// Calibrations already made, with their device copies per device.
struct CachedCalibration {
    Calibration c;
    std::map<int, double*> device;
};

std::mutex g_calibration_mu;
std::map<std::tuple<int, bool, double>, CachedCalibration> g_calibrations;

CachedCalibration& cached_calibration(int lmax, bool full, double max_angle) {
    const auto key = std::make_tuple(lmax, full, max_angle);
    auto it = g_calibrations.find(key);
    if (it == g_calibrations.end()) {
        it = g_calibrations.emplace(key, CachedCalibration{calibrate(lmax, full, max_angle), {}}).first;
    }
    return it->second;
}

int check_prob_angle(double max_angle_deg) {
    if (!(max_angle_deg * (M_PI / 180.0) > 0.001 && max_angle_deg <= 180.0)) {
        return fail("the maximum angle must be above 0.001 rad (0.057 degrees) and at most 180, got " +
                    number_text(max_angle_deg));
    }
    return 0;
}

// lmax_of (cpp/src/python/tracking.cpp:152-172).
int lmax_of(int64_t n_coeffs, bool full, int* lmax) {
    auto root = [](int64_t v) { return static_cast<int64_t>(std::sqrt(static_cast<double>(v))); };
    const int64_t l = full ? std::max<int64_t>(root(n_coeffs) - 1, 0)
                           : std::max<int64_t>(root(8 * n_coeffs + 1) - 3, 0) / 2;
    if (l > MAX_LMAX || (!full && l % 2 != 0)) {
        return fail("lmax must be at most " + std::to_string(MAX_LMAX) +
                    ", and even unless full, got " + std::to_string(l));
    }
    const int64_t expect = full ? (l + 1) * (l + 1) : (l + 1) * (l + 2) / 2;
    if (expect != n_coeffs) {
        return fail("lmax " + std::to_string(l) + " has " + std::to_string(expect) +
                    " coefficients, got " + std::to_string(n_coeffs));
    }
    *lmax = static_cast<int>(l);
    return 0;
}

VoxelMap voxel_map(const double* inv) {
    VoxelMap m{};
    for (int i = 0; i < 12; ++i) m.m[i] = inv[i];
    return m;
}

} // namespace

} // namespace lsd

using namespace lsd;

extern "C" {

int lsd_track_calibrate(int64_t lmax, int full, double max_angle_rad, double* grid, int64_t cap,
                        int64_t* n, double* ratio) {
    if (lmax < 0 || lmax > MAX_LMAX || (!full && lmax % 2 != 0)) {
        return fail("lmax must be at most " + std::to_string(MAX_LMAX) +
                    ", and even unless full, got " + std::to_string(lmax));
    }
    int rc = check_prob_angle(max_angle_rad * (180.0 / M_PI));
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g_calibration_mu);
    const Calibration& c = cached_calibration(static_cast<int>(lmax), full != 0, max_angle_rad).c;
    const int64_t rows = static_cast<int64_t>(c.grid.size() / 3);
    for (int64_t i = 0; i < std::min(rows, cap) * 3; ++i) grid[i] = c.grid[static_cast<size_t>(i)];
    *n = rows;
    *ratio = c.ratio;
    return 0;
}

int lsd_track_batch(const float* fod, int64_t nx, int64_t ny, int64_t nz, int64_t nc,
                    const double* fod_inv_affine, const float* mask, int64_t mx, int64_t my,
                    int64_t mz, const double* mask_inv_affine, const double* seeds,
                    const double* directions, const int64_t* indices, int64_t n_seeds,
                    uint64_t run_seed, double step_size, int64_t max_steps, int bidirectional,
                    const lsd_track_params* params, float* positions, int64_t positions_rows,
                    int64_t* counts, uint8_t* status, double* followed, lsd_track_stats* stats,
                    void* stream_) {
    auto stream = static_cast<hipStream_t>(stream_);
    const lsd_track_params& p = *params;
    if (p.algorithm < LSD_TRACK_DET || p.algorithm > LSD_TRACK_TEST) {
        return fail("unknown tracking algorithm " + std::to_string(p.algorithm));
    }
    const bool full = p.full != 0;
    int lmax = 0;
    if (p.algorithm != LSD_TRACK_TEST) {
        if (!std::isfinite(p.cutoff) || p.cutoff < 0.0) {
            return fail("the cutoff must be non-negative and finite, got " + number_text(p.cutoff));
        }
        if (p.algorithm == LSD_TRACK_DET) {
            if (!(p.max_angle_deg > 0.0 && p.max_angle_deg <= 180.0)) {
                return fail("the maximum angle must be in (0, 180], got " +
                            number_text(p.max_angle_deg));
            }
        } else {
            int rc = check_prob_angle(p.max_angle_deg);
            if (rc) return rc;
            if (p.trials < 1 || p.attempts < 1) {
                return fail("trials and attempts must be at least 1, got " +
                            std::to_string(p.trials) + " and " + std::to_string(p.attempts));
            }
        }
    } else if (p.test_rule < 0 || p.test_rule > 3 ||
               (p.test_rule == 1 && (p.test_stop < 1 || p.test_stop > 6))) {
        return fail("a stop is an EndStatus from 1 to 6, got " + std::to_string(p.test_stop));
    }
    if (!std::isfinite(step_size) || step_size <= 0.0 || max_steps <= 0) {
        return fail("need a positive step size and max_steps, got " + number_text(step_size) +
                    " and " + std::to_string(max_steps));
    }
    if (positions_rows < n_seeds) {
        return fail("positions must be (at least " + std::to_string(n_seeds) + ", " +
                    std::to_string(max_steps + 1) + ", 3), got (" +
                    std::to_string(positions_rows) + ", " + std::to_string(max_steps + 1) +
                    ", 3)");
    }
    if (p.algorithm != LSD_TRACK_TEST) {
        int rc = lmax_of(nc, full, &lmax);
        if (rc) return rc;
        if (!track_supported(lmax, full)) {
            return fail("the GPU tracking kernels take lmax <= 20, got " + std::to_string(lmax));
        }
    }
    if (nx > INT32_MAX || ny > INT32_MAX || nz > INT32_MAX || mx > INT32_MAX || my > INT32_MAX ||
        mz > INT32_MAX || n_seeds > INT32_MAX - (1 << 24)) {
        return fail("the tracking kernels index images and seed queues with 32-bit integers");
    }

    TrackArgs a{};
    a.fod = fod;
    a.nx = static_cast<int>(nx);
    a.ny = static_cast<int>(ny);
    a.nz = static_cast<int>(nz);
    a.nc = static_cast<int>(nc);
    a.fod_map = voxel_map(fod_inv_affine);
    a.mask = mask;
    if (mask != nullptr) {
        a.mx = static_cast<int>(mx);
        a.my = static_cast<int>(my);
        a.mz = static_cast<int>(mz);
        a.mask_map = voxel_map(mask_inv_affine);
    }
    a.seeds = seeds;
    a.directions = directions;
    a.indices = reinterpret_cast<const long long*>(indices);
    a.n_seeds = n_seeds;
    a.run_seed = run_seed;
    a.step = step_size;
    a.max_steps = max_steps;
    a.bidirectional = bidirectional != 0;
    a.cutoff = p.cutoff;
    const double rad = p.max_angle_deg * (M_PI / 180.0);
    a.cos_max_angle = p.max_angle_deg >= 180.0 ? -std::numeric_limits<double>::infinity()
                                               : std::cos(rad);
    a.max_angle_rad = rad;
    a.sin_bound = rad >= M_PI / 2.0 ? 1.0 : std::sin(rad);
    a.trials = p.trials;
    a.attempts = p.attempts;
    a.test_rule = p.test_rule;
    a.test_stop = p.test_stop;
    a.test_past_x = p.test_refuses_past_x;
    a.positions = positions;
    a.counts = reinterpret_cast<long long*>(counts);
    a.status = status;
    a.followed = followed;

    uint64_t n_grid = 0;
    if (p.algorithm == LSD_TRACK_PROB) {
        int dev = 0;
        LSD_HIP(hipGetDevice(&dev));
        std::lock_guard<std::mutex> lock(g_calibration_mu);
        CachedCalibration& cc = cached_calibration(lmax, full, rad);
        auto it = cc.device.find(dev);
        if (it == cc.device.end()) {
            double* d = nullptr;
            LSD_HIP(hipMalloc(reinterpret_cast<void**>(&d),
                              std::max<size_t>(8, cc.c.grid.size() * sizeof(double))));
            if (!cc.c.grid.empty()) {
                LSD_HIP(hipMemcpy(d, cc.c.grid.data(), cc.c.grid.size() * sizeof(double),
                                  hipMemcpyHostToDevice));
            }
            it = cc.device.emplace(dev, d).first;
        }
        a.grid = it->second;
        a.n_grid = static_cast<int>(cc.c.grid.size() / 3);
        a.ratio = cc.c.ratio;
        n_grid = static_cast<uint64_t>(a.n_grid);
    }
    if (p.algorithm != LSD_TRACK_TEST) {
        int rc = sh_constants(&a.k);
        if (rc) return rc;
    }

    Workspace mem(stream);
    unsigned long long* counters = nullptr; // 4 stats, then the queue
    int rc = mem.get(&counters, 5);
    if (rc) return rc;
    LSD_HIP(hipMemsetAsync(counters, 0, 5 * sizeof(unsigned long long), stream));
    a.stats = counters;
    a.next_seed = reinterpret_cast<int*>(counters + 4);
    if (n_seeds > 0) {
        rc = launch_track(p.algorithm, lmax, full, a, 0, stream);
        if (rc) return rc;
    }
    if (stats != nullptr) {
        unsigned long long host[4] = {0, 0, 0, 0};
        LSD_HIP(hipMemcpyAsync(host, counters, sizeof(host), hipMemcpyDeviceToHost, stream));
        LSD_HIP(hipStreamSynchronize(stream));
        stats->sampled_steps = host[0];
        stats->accepted_index_sum = host[1];
        stats->truncations = host[2];
        double worst = 0.0;
        std::memcpy(&worst, &host[3], sizeof(double));
        stats->max_truncation = worst;
        stats->calibration_directions = n_grid;
    }
    return 0;
}

} // extern "C"
