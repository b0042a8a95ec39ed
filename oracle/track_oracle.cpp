// Test infrastructure only: the CPU FP64 oracle of larmor's streamline tracking through a field of
// FODs, the consumer of the deconvolution's output.
//
// Only tests/ and bench.py's CPU-baseline leg load this code (it is linked into
// liblarmor_oracle.so), as the checker or the timed CPU baseline; the product path never does.
//
// Each function restates the reference's algorithm with its step order and tie rules (paths
// relative to the reference repository):
//   cpp/src/tracking/rng.cpp        the counter-based SplitMix64 draws of a streamline
//   cpp/src/tracking/grid.cpp       world-to-voxel maps, the image's extent, trilinear
//                                   interpolation of the FOD, the nearest-voxel mask
//   cpp/src/common/maths.cpp        rotate_to, norm3
//   cpp/src/sh/harmonics.cpp        amplitude (the series along a unit vector)
//   cpp/src/tracking/calibrate.cpp  direction_grid, calibrate (rejection-sampling bound)
//   cpp/src/tracking/det.cpp        deterministic tracking (SD_STREAM: climb to the nearest peak)
//   cpp/src/tracking/prob.cpp       probabilistic tracking (iFOD1: rejection sampling in a cone)
//   cpp/src/tracking/testing.cpp    the fixed-rule test algorithm of the step loop
//   cpp/src/tracking/tracker.cpp    track_one_way, track_seed, track_batch
// The series' derivatives and the Newton climb are the shared ones of sh_series.hpp.
//
// Compiled with -O3 -ffp-contract=off (no fused multiply-add), as the reference is.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <limits>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "sh_series.hpp"

namespace {

using namespace oracle_sh;
using u64 = std::uint64_t;

thread_local std::string g_track_error;

// ---------------------------------------------------------------------------------------------
// Random numbers (rng.cpp): the k-th draw of key K is SplitMix64's finaliser of K + (k + 1) G.

constexpr u64 SPLITMIX_GAMMA = 0x9E3779B97F4A7C15ull;
constexpr u64 SPLITMIX_M1 = 0xBF58476D1CE4E5B9ull;
constexpr u64 SPLITMIX_M2 = 0x94D049BB133111EBull;
constexpr u64 SEARCH_COUNTER = u64{1} << 63;

// SplitMix64's finaliser (rng.cpp:21-27); unsigned products wrap.
u64 splitmix_finalise(u64 z) {
    z = (z ^ (z >> 30)) * SPLITMIX_M1;
    z = (z ^ (z >> 27)) * SPLITMIX_M2;
    return z ^ (z >> 31);
}

struct Draws {
    u64 key = 0;
    u64 counter = 0;

    // Rng::streamline (rng.cpp:33-35): key = f(f(run_seed) + index), counter 0.
    static Draws streamline(u64 run_seed, u64 index) {
        return {splitmix_finalise(splitmix_finalise(run_seed) + index), 0};
    }
    // Rng::search (rng.cpp:37-39): the same key from counter 2^63.
    Draws search() const { return {key, SEARCH_COUNTER}; }

    // Rng::uniform (rng.cpp:42-46): the top 53 bits times 2^-53.
    double uniform() {
        const u64 bits = splitmix_finalise(key + (counter + 1) * SPLITMIX_GAMMA);
        counter += 1;
        return static_cast<double>(bits >> 11) * (1.0 / 9007199254740992.0);
    }

    // Rng::unit_vector (rng.cpp:48-55): z then the azimuth.
    void unit_vector(double out[3]) {
        const double z = 2.0 * uniform() - 1.0;
        const double phi = 2.0 * M_PI * uniform();
        const double r = std::sqrt(1.0 - z * z);
        out[0] = r * std::cos(phi);
        out[1] = r * std::sin(phi);
        out[2] = z;
    }

    // Rng::unit_vector_in_cone (rng.cpp:57-73): the azimuth, then the polar angle by rejection
    // against its sine.
    void unit_vector_in_cone(double max_angle, double sin_bound, double out[3]) {
        const double phi = 2.0 * M_PI * uniform();
        double theta = 0.0;
        while (true) {
            theta = max_angle * uniform();
            if (sin_bound * uniform() <= std::sin(theta)) {
                break;
            }
        }
        const double st = std::sin(theta);
        const double ct = std::cos(theta);
        out[0] = st * std::cos(phi);
        out[1] = st * std::sin(phi);
        out[2] = ct;
    }
};

// ---------------------------------------------------------------------------------------------
// Grids (grid.cpp).

struct Grid {
    Index shape[3] = {0, 0, 0};
    double inv[16] = {};

    // Grid::to_voxel (grid.cpp:29-36): each coordinate summed from the x term to the translation.
    void to_voxel(const double p[3], double v[3]) const {
        for (int a = 0; a < 3; ++a) {
            v[a] = inv[4 * a + 0] * p[0] + inv[4 * a + 1] * p[1] + inv[4 * a + 2] * p[2] +
                   inv[4 * a + 3];
        }
    }

    // Grid::contains (grid.cpp:38-50): within half a voxel of the outer centres; NaN is outside.
    bool contains(const double p[3]) const {
        double v[3];
        to_voxel(p, v);
        for (int a = 0; a < 3; ++a) {
            const bool on_axis = -0.5 <= v[a] && v[a] <= static_cast<double>(shape[a]) - 0.5;
            if (shape[a] < 1 || !on_axis) {
                return false;
            }
        }
        return true;
    }
};

// clamped_index (grid.cpp:13-21): NaN and anything not above 0 give 0.
Index clamped_index(double index, Index last) {
    if (!(index > 0.0)) {
        return 0;
    }
    if (index >= static_cast<double>(last)) {
        return last;
    }
    return static_cast<Index>(index);
}

struct AxisWeights {
    Index lower, upper;
    double upper_weight;
};

// interpolation_axis (grid.cpp:52-59).
AxisWeights interpolation_axis(double coordinate, Index size) {
    const double lower = std::floor(coordinate);
    return {clamped_index(lower, size - 1), clamped_index(lower + 1.0, size - 1),
            coordinate - lower};
}

struct Field {
    const float* data = nullptr;
    Index channels = 0;
    Grid grid;

    // Field::trilinear (grid.cpp:66-95): out[0..channels) = sum over the 8 corners, x then y then
    // z, lower before upper, of weight * value, each corner's channels in order.
    void trilinear(const double p[3], double* out) const {
        double v[3];
        grid.to_voxel(p, v);
        const AxisWeights ax = interpolation_axis(v[0], grid.shape[0]);
        const AxisWeights ay = interpolation_axis(v[1], grid.shape[1]);
        const AxisWeights az = interpolation_axis(v[2], grid.shape[2]);
        for (Index c = 0; c < channels; ++c) out[c] = 0.0;
        const Index xs[2] = {ax.lower, ax.upper};
        const Index ys[2] = {ay.lower, ay.upper};
        const Index zs[2] = {az.lower, az.upper};
        const double wx[2] = {1.0 - ax.upper_weight, ax.upper_weight};
        const double wy[2] = {1.0 - ay.upper_weight, ay.upper_weight};
        const double wz[2] = {1.0 - az.upper_weight, az.upper_weight};
        for (int i = 0; i < 2; ++i) {
            for (int j = 0; j < 2; ++j) {
                for (int k = 0; k < 2; ++k) {
                    const double w = wx[i] * wy[j] * wz[k];
                    const float* corner =
                        data + ((xs[i] * grid.shape[1] + ys[j]) * grid.shape[2] + zs[k]) * channels;
                    for (Index c = 0; c < channels; ++c) {
                        out[c] = out[c] + w * static_cast<double>(corner[c]);
                    }
                }
            }
        }
    }
};

struct MaskImage {
    const float* data = nullptr;
    Grid grid;

    // Mask::inside (grid.cpp:100-115): the nearest voxel, halves rounded away from zero, must be
    // at least 0.5; a coordinate at or past half a voxel beyond the outer centres is outside.
    bool inside(const double p[3]) const {
        double v[3];
        grid.to_voxel(p, v);
        Index flat = 0;
        for (int a = 0; a < 3; ++a) {
            const bool on_axis = v[a] > -0.5 && v[a] < static_cast<double>(grid.shape[a]) - 0.5;
            if (!on_axis) {
                return false;
            }
            flat = flat * grid.shape[a] + static_cast<Index>(std::round(v[a]));
        }
        return static_cast<double>(data[flat]) >= 0.5;
    }
};

// ---------------------------------------------------------------------------------------------
// Directions and amplitudes.

double norm3(const double v[3]) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

// rotate_to (maths.cpp:23-42): the rotation taking +z to the unit vector r, applied to d.
void rotate_to(const double r[3], const double d[3], double out[3]) {
    const double xy = std::sqrt(r[0] * r[0] + r[1] * r[1]);
    if (xy == 0.0) {
        const double s = r[2] < 0.0 ? -1.0 : 1.0;
        for (int a = 0; a < 3; ++a) out[a] = s < 0.0 ? -d[a] : d[a];
        return;
    }
    const double h[3] = {r[0] / xy, r[1] / xy, 0.0};
    const double t[3] = {r[2] * h[0], r[2] * h[1], -xy};
    const double along_z = d[2];
    const double along_h = d[0] * h[0] + d[1] * h[1];
    out[0] = d[0] + along_z * r[0] + along_h * (t[0] - h[0]);
    out[1] = d[1] + along_z * r[1] + along_h * (t[1] - h[1]);
    out[2] = d[2] + along_z * (r[2] - 1.0) + along_h * (t[2] - h[2]);
}

// amplitude (harmonics.cpp:187-219): the zonal terms in degree order, then per order m the
// angle-addition cos(m phi), sin(m phi) and that order's degrees.
double sh_amplitude(const double* c, Index lmax, bool full, const double d[3]) {
    const double xy = std::sqrt(d[0] * d[0] + d[1] * d[1]);
    LegendreTable p;
    fill_legendre(lmax, d[2], xy, p);
    double cos_phi = 1.0;
    double sin_phi = 0.0;
    if (xy > 0.0) {
        cos_phi = d[0] / xy;
        sin_phi = d[1] / xy;
    }
    const Index step = full ? 1 : 2;
    double value = 0.0;
    for (Index l = 0; l <= lmax; l += step) {
        value = value + c[sh_index(l, 0, full)] * p[l][0];
    }
    double cm = 1.0;
    double sm = 0.0;
    for (Index m = 1; m <= lmax; ++m) {
        const double next_c = cm * cos_phi - sm * sin_phi;
        const double next_s = sm * cos_phi + cm * sin_phi;
        cm = next_c;
        sm = next_s;
        const Index first = full || m % 2 == 0 ? m : m + 1;
        for (Index l = first; l <= lmax; l += step) {
            value = value + M_SQRT2 * p[l][m] *
                                (c[sh_index(l, m, full)] * cm + c[sh_index(l, -m, full)] * sm);
        }
    }
    return value;
}

// ---------------------------------------------------------------------------------------------
// Calibration of the rejection sampling (calibrate.cpp).

// saturating_index (calibrate.cpp:24-37).
Index saturating_index(double value) {
    if (std::isnan(value)) return 0;
    constexpr double limit = static_cast<double>(std::numeric_limits<Index>::max());
    if (value >= limit) return std::numeric_limits<Index>::max();
    if (value <= -limit) return std::numeric_limits<Index>::min();
    return static_cast<Index>(value);
}

// direction_grid (calibrate.cpp:41-60): a hexagonal lattice of spacing `spacing` mapped to the
// sphere around +z, out to max_angle, lattice rows i then columns j.
std::vector<std::array<double, 3>> direction_grid(double max_angle, double spacing) {
    const double radius = max_angle / spacing;
    const double radius2 = radius * radius;
    const Index extent = saturating_index(std::ceil(radius));
    std::vector<std::array<double, 3>> grid;
    for (Index i = -extent; i <= extent; ++i) {
        for (Index j = -extent; j <= extent; ++j) {
            const double x = static_cast<double>(i) + 0.5 * static_cast<double>(j);
            const double y = 0.866025403784439 * static_cast<double>(j);
            const double r2 = x * x + y * y;
            if (r2 > radius2) continue;
            const double angle = spacing * std::sqrt(r2);
            const double scale = angle > 0.0 ? std::sin(angle) / std::sqrt(r2) : 0.0;
            grid.push_back({scale * x, scale * y, std::cos(angle)});
        }
    }
    return grid;
}

struct Calibration {
    std::vector<std::array<double, 3>> grid;
    double ratio = 0.0;
};

// calibrate (calibrate.cpp:63-108): a unit-peak fibre along +z sampled every 1e-3 rad to the
// cone's edge or its first non-positive amplitude; the spacing with the fewest expected
// evaluations per step (cone draws + grid size) wins.
Calibration calibrate(Index lmax, bool full, double max_angle) {
    std::vector<double> fibre(static_cast<size_t>(full ? (lmax + 1) * (lmax + 1)
                                                        : (lmax + 1) * (lmax + 2) / 2),
                              0.0);
    for (Index l = 0; l <= lmax; l += full ? 1 : 2) {
        fibre[static_cast<size_t>(sh_index(l, 0, full))] =
            std::sqrt(static_cast<double>(2 * l + 1) / (4.0 * M_PI));
    }
    std::vector<std::pair<double, double>> samples;
    double elevation = 0.0;
    while (elevation < max_angle) {
        const double d[3] = {std::sin(elevation), 0.0, std::cos(elevation)};
        const double a = sh_amplitude(fibre.data(), lmax, full, d);
        samples.emplace_back(elevation, a);
        if (!std::isfinite(a) || a <= 0.0) break;
        elevation += 0.001;
    }
    const double nan = std::numeric_limits<double>::quiet_NaN();
    if (samples.empty()) return {{}, nan};
    const double zero = samples.back().first;
    const double on_axis = samples.front().second;
    double best = std::numeric_limits<double>::infinity();
    double theta = nan;
    double ratio = nan;
    for (size_t s = 1; s < samples.size(); ++s) {
        const double e = samples[s].first;
        const double a = samples[s].second;
        const double n_cone = max_angle * max_angle * (1.0 + on_axis / a) / (2.0 * zero * zero);
        const double n_eval =
            n_cone + static_cast<double>(direction_grid(max_angle + e, std::sqrt(3.0) * e).size());
        if (n_eval > 0.0 && n_eval < best) {
            best = n_eval;
            theta = e;
            ratio = on_axis / a;
        }
    }
    return {direction_grid(max_angle + theta, std::sqrt(3.0) * theta), ratio};
}

// ---------------------------------------------------------------------------------------------
// Algorithms (det.cpp, prob.cpp, testing.cpp) and the step loop (tracker.cpp).

enum Stop : std::uint8_t {
    NONE = 0,
    MAX_LENGTH = 1,
    LEFT_VOLUME = 2,
    LEFT_MASK = 3,
    LOW_FOD = 4,
    SHARP_TURN = 5,
    NO_DIRECTION = 6,
};

struct Stats {
    u64 sampled_steps = 0, accepted_index_sum = 0, truncations = 0;
    double max_truncation = 0.0;
    void add(const Stats& o) {
        sampled_steps += o.sampled_steps;
        accepted_index_sum += o.accepted_index_sum;
        truncations += o.truncations;
        max_truncation = std::fmax(max_truncation, o.max_truncation);
    }
};

} // namespace

// The algorithm's settings, laid out as lsd_track_params of include/larmor_sd.h.
struct oracle_track_params {
    std::int32_t algorithm; // 0 det, 1 prob, 2 test
    std::int32_t full;
    double cutoff;
    double max_angle_deg;
    std::int64_t trials;
    std::int64_t attempts;
    std::int32_t test_rule; // 0 straight, 1 refuse, 2 quarter turn, 3 wander
    std::int32_t test_stop; // the refusal's stop, 1..6
    double test_refuses_past_x;
};

namespace {

struct Algorithm {
    oracle_track_params p{};
    Index lmax = 0;
    bool full = false;
    double cos_max_angle = 0.0; // det
    double max_angle_rad = 0.0; // prob
    double sin_bound = 0.0;     // prob
    Calibration calibration;    // prob

    // first_direction of det.cpp:36-50, prob.cpp:48-58 and testing.cpp:10-17.
    bool first_direction(const Field& fod, const double pos[3], Draws& rng, double out[3]) const {
        if (p.algorithm == 2) {
            rng.unit_vector(out);
            return pos[0] <= p.test_refuses_past_x;
        }
        std::vector<double> c(static_cast<size_t>(fod.channels));
        fod.trilinear(pos, c.data());
        if (p.algorithm == 0) {
            for (int it = 0; it < 20; ++it) {
                double start[3];
                rng.unit_vector(start);
                ShPeak peak;
                if (sh_climb(c.data(), lmax, full, start, 50, 1e-4, 0.2, peak) &&
                    peak.amplitude >= p.cutoff) {
                    for (int a = 0; a < 3; ++a) out[a] = peak.dir[a];
                    return true;
                }
            }
            return false;
        }
        for (Index attempt = 0; attempt < p.attempts; ++attempt) {
            double d[3];
            rng.unit_vector(d);
            if (sh_amplitude(c.data(), lmax, full, d) > p.cutoff) {
                for (int a = 0; a < 3; ++a) out[a] = d[a];
                return true;
            }
        }
        return false;
    }

    // next_direction of det.cpp:52-74, prob.cpp:61-96 and testing.cpp:19-39: NONE and the new
    // direction in `dir`, or the stop.
    Stop next_direction(const Field& fod, const double pos[3], double dir[3], Draws& rng,
                        Stats& stats) const {
        if (p.algorithm == 2) {
            switch (p.test_rule) {
            case 0:
                return NONE;
            case 1:
                return static_cast<Stop>(p.test_stop);
            case 2: {
                const double x = -dir[1];
                const double y = dir[0];
                dir[0] = x;
                dir[1] = y;
                dir[2] = 0.0;
                return NONE;
            }
            default: {
                const double x = dir[0] + 0.5 * (rng.uniform() - 0.5);
                const double y = dir[1] + 0.5 * (rng.uniform() - 0.5);
                const double z = dir[2] + 0.5 * (rng.uniform() - 0.5);
                const double v[3] = {x, y, z};
                const double n = norm3(v);
                dir[0] = x / n;
                dir[1] = y / n;
                dir[2] = z / n;
                return NONE;
            }
            }
        }
        std::vector<double> c(static_cast<size_t>(fod.channels));
        fod.trilinear(pos, c.data());
        if (p.algorithm == 0) {
            ShPeak peak;
            if (!sh_climb(c.data(), lmax, full, dir, 50, 1e-4, 0.2, peak)) {
                return NO_DIRECTION;
            }
            const double cos_turn =
                peak.dir[0] * dir[0] + peak.dir[1] * dir[1] + peak.dir[2] * dir[2];
            if (peak.amplitude < p.cutoff) return LOW_FOD;
            if (cos_turn < cos_max_angle) return SHARP_TURN;
            for (int a = 0; a < 3; ++a) dir[a] = peak.dir[a];
            return NONE;
        }
        double largest = 0.0;
        for (const auto& g : calibration.grid) {
            double r[3];
            rotate_to(dir, g.data(), r);
            largest = std::fmax(largest, sh_amplitude(c.data(), lmax, full, r));
        }
        if (largest <= 0.0) return LOW_FOD;
        const double bound = largest * calibration.ratio;
        stats.sampled_steps += 1;
        for (Index trial = 0; trial < p.trials; ++trial) {
            double cone[3];
            rng.unit_vector_in_cone(max_angle_rad, sin_bound, cone);
            double cand[3];
            rotate_to(dir, cone, cand);
            const double a = sh_amplitude(c.data(), lmax, full, cand);
            if (std::isnan(a) || a <= p.cutoff) continue;
            if (a > bound) {
                stats.truncations += 1;
                stats.max_truncation = std::fmax(stats.max_truncation, a / bound);
            }
            if (rng.uniform() < a / bound) {
                stats.accepted_index_sum += static_cast<u64>(trial);
                const double n = norm3(cand);
                for (int a2 = 0; a2 < 3; ++a2) dir[a2] = cand[a2] / n;
                return NONE;
            }
        }
        return NO_DIRECTION;
    }
};

struct Setup {
    Field fod;
    bool has_mask = false;
    MaskImage mask;
    double step = 0.0;
    Index max_steps = 0;
    bool bidirectional = true;
    u64 run_seed = 0;
};

void write_point(float* row, Index i, const double p[3]) {
    for (int a = 0; a < 3; ++a) row[3 * i + a] = static_cast<float>(p[a]);
}

// track_one_way (tracker.cpp:42-71): step, drop a point off the image or the mask, stop at
// max_steps + 1 points, else ask the algorithm.
Stop track_one_way(const Setup& s, const Algorithm& alg, const double start[3],
                   const double direction[3], Index& n_points, float* row, Draws& rng,
                   Stats& stats) {
    double pos[3] = {start[0], start[1], start[2]};
    double dir[3] = {direction[0], direction[1], direction[2]};
    while (true) {
        for (int a = 0; a < 3; ++a) pos[a] = pos[a] + s.step * dir[a];
        if (!s.fod.grid.contains(pos)) return LEFT_VOLUME;
        if (s.has_mask && !s.mask.inside(pos)) return LEFT_MASK;
        write_point(row, n_points, pos);
        n_points += 1;
        if (n_points == s.max_steps + 1) return MAX_LENGTH;
        const Stop stop = alg.next_direction(s.fod, pos, dir, rng, stats);
        if (stop != NONE) return stop;
    }
}

struct Track {
    Index n_points = 0;
    std::uint8_t ends[2] = {0, 0};
    double direction[3] = {0.0, 0.0, 0.0};
    Stats stats;
};

// track_seed (tracker.cpp:73-121).
Track track_seed(const Setup& s, const Algorithm& alg, const double seed[3],
                 const double* given, u64 index, float* row) {
    Track t;
    if (!s.fod.grid.contains(seed) || (s.has_mask && !s.mask.inside(seed))) {
        return t;
    }
    Draws rng = Draws::streamline(s.run_seed, index);
    double dir[3];
    if (given != nullptr) {
        for (int a = 0; a < 3; ++a) dir[a] = given[a];
    } else {
        Draws search = rng.search();
        if (!alg.first_direction(s.fod, seed, search, dir)) {
            return t;
        }
    }
    write_point(row, 0, seed);
    Index n = 1;
    Stats first_stats;
    const Stop first = track_one_way(s, alg, seed, dir, n, row, rng, first_stats);
    for (int a = 0; a < 3; ++a) t.direction[a] = dir[a];
    if (!s.bidirectional) {
        t.n_points = n;
        t.ends[1] = first;
        t.stats = first_stats;
        return t;
    }
    for (Index i = 0; i < n / 2; ++i) {
        for (int a = 0; a < 3; ++a) std::swap(row[3 * i + a], row[3 * (n - 1 - i) + a]);
    }
    t.ends[0] = first;
    if (n == s.max_steps + 1) {
        t.n_points = n;
        t.ends[1] = MAX_LENGTH;
        t.stats = first_stats;
        return t;
    }
    const double back[3] = {-dir[0], -dir[1], -dir[2]};
    Stats second_stats;
    t.ends[1] = track_one_way(s, alg, seed, back, n, row, rng, second_stats);
    t.n_points = n;
    t.stats = first_stats;
    t.stats.add(second_stats);
    return t;
}

void copy_affine(const double* src, double dst[16]) {
    for (int i = 0; i < 16; ++i) dst[i] = src[i];
}

} // namespace

extern "C" {

const char* oracle_track_last_error() { return g_track_error.c_str(); }

// Rng draws for tests: the key of (run_seed, index) and n uniforms from counter 0 (or from the
// search's 2^63 when search != 0).
std::uint64_t oracle_rng_uniforms(std::uint64_t run_seed, std::uint64_t index, int search,
                                  std::int64_t n, double* out) {
    Draws d = Draws::streamline(run_seed, index);
    if (search) d = d.search();
    for (std::int64_t i = 0; i < n; ++i) out[i] = d.uniform();
    return d.key;
}

// Amplitudes of one series along n unit vectors (n, 3).
void oracle_sh_amplitude(const double* coeffs, std::int64_t lmax, int full, const double* dirs,
                         std::int64_t n, double* out) {
    for (std::int64_t i = 0; i < n; ++i) out[i] = sh_amplitude(coeffs, lmax, full != 0, dirs + 3 * i);
}

// rotate_to of n directions (n, 3) by the rotation taking +z to reference.
void oracle_rotate_to(const double* reference, const double* dirs, std::int64_t n, double* out) {
    for (std::int64_t i = 0; i < n; ++i) rotate_to(reference, dirs + 3 * i, out + 3 * i);
}

// direction_grid into grid (cap rows of 3); returns its size (rows past cap are not written).
std::int64_t oracle_direction_grid(double max_angle, double spacing, double* grid,
                                   std::int64_t cap) {
    const auto g = direction_grid(max_angle, spacing);
    for (size_t i = 0; i < g.size() && static_cast<std::int64_t>(i) < cap; ++i) {
        for (int a = 0; a < 3; ++a) grid[3 * i + a] = g[i][a];
    }
    return static_cast<std::int64_t>(g.size());
}

// calibrate: the grid's size (rows past cap are not written) and the ratio.
std::int64_t oracle_calibrate(std::int64_t lmax, int full, double max_angle, double* grid,
                              std::int64_t cap, double* ratio) {
    const Calibration c = calibrate(lmax, full != 0, max_angle);
    for (size_t i = 0; i < c.grid.size() && static_cast<std::int64_t>(i) < cap; ++i) {
        for (int a = 0; a < 3; ++a) grid[3 * i + a] = c.grid[i][a];
    }
    *ratio = c.ratio;
    return static_cast<std::int64_t>(c.grid.size());
}

// track_batch (tracker.cpp:124-139 with the binding's outputs, cpp/src/python/tracking.cpp):
// fod (nx, ny, nz, nc) float32; mask (mx, my, mz) float32 or null; affines (4, 4) world-to-voxel;
// seeds (n, 3), directions (n, 3) or null, indices (n); positions (n, max_steps + 1, 3) float32;
// counts (n), status (n, 2), followed (n, 3); stats: sampled steps, accepted index sum,
// truncations (uint64) and the largest truncation (double), then the calibration's size.
int oracle_track_batch(const float* fod, std::int64_t nx, std::int64_t ny, std::int64_t nz,
                       std::int64_t nc, const double* fod_inv, const float* mask, std::int64_t mx,
                       std::int64_t my, std::int64_t mz, const double* mask_inv,
                       const double* seeds, const double* directions, const std::int64_t* indices,
                       std::int64_t n_seeds, std::uint64_t run_seed, double step_size,
                       std::int64_t max_steps, int bidirectional,
                       const oracle_track_params* params, int n_threads, float* positions,
                       std::int64_t* counts, std::uint8_t* status, double* followed,
                       std::uint64_t* stats_counts, double* max_truncation,
                       std::int64_t* calibration_size) {
    try {
        Setup s;
        s.fod.data = fod;
        s.fod.channels = nc;
        s.fod.grid.shape[0] = nx;
        s.fod.grid.shape[1] = ny;
        s.fod.grid.shape[2] = nz;
        copy_affine(fod_inv, s.fod.grid.inv);
        if (mask != nullptr) {
            s.has_mask = true;
            s.mask.data = mask;
            s.mask.grid.shape[0] = mx;
            s.mask.grid.shape[1] = my;
            s.mask.grid.shape[2] = mz;
            copy_affine(mask_inv, s.mask.grid.inv);
        }
        s.step = step_size;
        s.max_steps = max_steps;
        s.bidirectional = bidirectional != 0;
        s.run_seed = run_seed;

        Algorithm alg;
        alg.p = *params;
        alg.full = params->full != 0;
        if (params->algorithm != 2) {
            const Index lmax = alg.full ? static_cast<Index>(std::sqrt(static_cast<double>(nc))) - 1
                                        : (static_cast<Index>(std::sqrt(8.0 * nc + 1.0)) - 3) / 2;
            alg.lmax = std::max<Index>(lmax, 0);
        }
        const double deg = M_PI / 180.0;
        alg.cos_max_angle = params->max_angle_deg >= 180.0
                                ? -std::numeric_limits<double>::infinity()
                                : std::cos(params->max_angle_deg * deg);
        alg.max_angle_rad = params->max_angle_deg * deg;
        alg.sin_bound = alg.max_angle_rad >= M_PI / 2.0 ? 1.0 : std::sin(alg.max_angle_rad);
        if (params->algorithm == 1) {
            alg.calibration = calibrate(alg.lmax, alg.full, alg.max_angle_rad);
        }
        *calibration_size = static_cast<std::int64_t>(alg.calibration.grid.size());

        const Index row_len = (max_steps + 1) * 3;
        std::vector<Track> tracks(static_cast<size_t>(n_seeds));
        std::atomic<Index> next{0};
        std::exception_ptr error;
        std::mutex error_mu;
        auto worker = [&] {
            try {
                while (true) {
                    const Index begin = next.fetch_add(4);
                    if (begin >= n_seeds) break;
                    const Index end = std::min<Index>(begin + 4, n_seeds);
                    for (Index i = begin; i < end; ++i) {
                        tracks[static_cast<size_t>(i)] =
                            track_seed(s, alg, seeds + 3 * i,
                                       directions != nullptr ? directions + 3 * i : nullptr,
                                       static_cast<u64>(indices[i]), positions + i * row_len);
                    }
                }
            } catch (...) {
                std::lock_guard<std::mutex> lock(error_mu);
                if (!error) error = std::current_exception();
            }
        };
        const int n = std::max(1, n_threads);
        std::vector<std::thread> pool;
        for (int t = 1; t < n; ++t) pool.emplace_back(worker);
        worker();
        for (auto& th : pool) th.join();
        if (error) std::rethrow_exception(error);

        Stats total;
        for (Index i = 0; i < n_seeds; ++i) {
            const Track& t = tracks[static_cast<size_t>(i)];
            counts[i] = t.n_points;
            status[2 * i] = t.ends[0];
            status[2 * i + 1] = t.ends[1];
            for (int a = 0; a < 3; ++a) followed[3 * i + a] = t.direction[a];
            total.add(t.stats);
        }
        stats_counts[0] = total.sampled_steps;
        stats_counts[1] = total.accepted_index_sum;
        stats_counts[2] = total.truncations;
        *max_truncation = total.max_truncation;
        return 0;
    } catch (const std::exception& e) {
        g_track_error = e.what();
        return 1;
    }
}

} // extern "C"
