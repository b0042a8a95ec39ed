// Streamline tracking through a field of FODs: one streamline per lane, persistent waves fed by a
// device seed queue.
//
// Restates larmor's tracking (cpp/src/tracking/):
//  - tracker.cpp:42-121 the step loop: step along the current direction, drop a point off the
//    FOD's image (LeftVolume) or outside the mask (LeftMask), stop at max_steps + 1 points
//    (MaxLength), else ask the algorithm; bidirectionally the first half is reversed in place and
//    the second half tracks from the seed along the opposite direction;
//  - grid.cpp the world-to-voxel maps, the image's extent, trilinear interpolation of the float32
//    SH coefficients into FP64 (corners x, y, z lower before upper), the nearest-voxel mask;
//  - rng.cpp the counter-based SplitMix64 draws keyed by (run seed, seed index): a streamline's
//    numbers do not depend on which lane or wave tracks it;
//  - det.cpp (SD_STREAM) a Newton climb from the previous direction to the nearest FOD peak;
//  - prob.cpp (iFOD1) the rejection-sampling bound from the calibration grid turned to the
//    previous direction, then candidates drawn in the cone until one is accepted;
//  - testing.cpp the fixed-rule test algorithm (straight, refuse, quarter turn, wander).
//
// Work mapping.  Streamline lengths differ by orders of magnitude, so lanes are not given fixed
// seeds: each pass of the kernel's loop takes one step of every lane's streamline, and a lane
// whose streamline has ended takes the next seed from the queue in the same pass (one atomic per
// wave per pass, a ballot numbering the lanes that need one).  A wave exits when the queue is
// empty and all its lanes have finished.  The coefficients interpolated at a position live in
// registers (LMAX and FULL are template parameters; the lmax-20 instantiations, which serve every
// lmax above 12, or 8 with all degrees, keep them in the lane's private memory), and the Legendre
// values of a climb or an amplitude are made a degree or an order at a time.  The transcendental functions (acos, atan2, sin, cos) are the device library's, so
// det and prob match the CPU oracle to rounding; the step loop and trilinear interpolation are
// the same arithmetic as the oracle's.
#pragma once

#include "peaks_wave.hpp"

namespace lsd {

enum TrackAlgo : int { TRACK_DET = 0, TRACK_PROB = 1, TRACK_TEST = 2 };

// The stop codes of larmor's EndStatus; 0 is an end that was not tracked.
enum TrackStop : int {
    STOP_NONE = 0,
    STOP_MAX_LENGTH = 1,
    STOP_LEFT_VOLUME = 2,
    STOP_LEFT_MASK = 3,
    STOP_LOW_FOD = 4,
    STOP_SHARP_TURN = 5,
    STOP_NO_DIRECTION = 6,
};

// The rows of a world-to-voxel matrix (the fourth row of the affine is not used).
struct VoxelMap {
    double m[12];
};

struct TrackArgs {
    const float* fod; // (nx, ny, nz, nc), channel fastest
    int nx, ny, nz, nc;
    VoxelMap fod_map;
    const float* mask; // (mx, my, mz) or null
    int mx, my, mz;
    VoxelMap mask_map;
    const double* seeds;      // (n, 3)
    const double* directions; // (n, 3) or null
    const long long* indices; // (n)
    long n_seeds;
    unsigned long long run_seed;
    double step;
    long max_steps;
    int bidirectional;
    // algorithm
    double cutoff;
    double cos_max_angle; // det: -inf at 180 degrees
    double max_angle_rad; // prob
    double sin_bound;     // prob
    long trials, attempts;
    const double* grid; // prob calibration directions (n_grid, 3)
    int n_grid;
    double ratio;
    int test_rule, test_stop;
    double test_past_x;
    ShConstantsDev k;
    // outputs
    float* positions; // (n, max_steps + 1, 3)
    long long* counts;
    unsigned char* status; // (n, 2)
    double* followed;      // (n, 3)
    unsigned long long* stats; // sampled steps, accepted index sum, truncations, max truncation bits
    int* next_seed;
    int lanes; // lanes of each wave that take seeds from the queue (1-64)
};

// ---------------------------------------------------------------------------------------------
// Counter-based draws (rng.cpp).

constexpr unsigned long long SPLITMIX_GAMMA = 0x9E3779B97F4A7C15ull;

__device__ __forceinline__ unsigned long long splitmix_finalise(unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct LaneDraws {
    unsigned long long key;
    unsigned long long counter;

    __device__ __forceinline__ double uniform() {
        const unsigned long long bits = splitmix_finalise(key + (counter + 1ull) * SPLITMIX_GAMMA);
        counter += 1ull;
        return static_cast<double>(bits >> 11) * (1.0 / 9007199254740992.0);
    }

    __device__ __forceinline__ void unit_vector(double (&out)[3]) {
        const double z = 2.0 * uniform() - 1.0;
        const double phi = 2.0 * M_PI * uniform();
        const double r = sqrt(1.0 - z * z);
        out[0] = r * cos(phi);
        out[1] = r * sin(phi);
        out[2] = z;
    }

    __device__ __forceinline__ void unit_vector_in_cone(double max_angle, double sin_bound,
                                                        double (&out)[3]) {
        const double phi = 2.0 * M_PI * uniform();
        double theta = 0.0;
        while (true) {
            theta = max_angle * uniform();
            if (sin_bound * uniform() <= sin(theta)) {
                break;
            }
        }
        const double st = sin(theta);
        const double ct = cos(theta);
        out[0] = st * cos(phi);
        out[1] = st * sin(phi);
        out[2] = ct;
    }
};

// ---------------------------------------------------------------------------------------------
// Grids (grid.cpp).

__device__ __forceinline__ void to_voxel(const VoxelMap& g, const double (&p)[3], double (&v)[3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        v[a] = g.m[4 * a + 0] * p[0] + g.m[4 * a + 1] * p[1] + g.m[4 * a + 2] * p[2] +
               g.m[4 * a + 3];
    }
}

__device__ __forceinline__ bool grid_contains(const VoxelMap& g, int nx, int ny, int nz,
                                              const double (&p)[3]) {
    double v[3];
    to_voxel(g, p, v);
    const int n[3] = {nx, ny, nz};
    bool in = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        in = in && n[a] >= 1 && -0.5 <= v[a] && v[a] <= static_cast<double>(n[a]) - 0.5;
    }
    return in;
}

__device__ __forceinline__ bool mask_inside(const TrackArgs& A, const double (&p)[3]) {
    double v[3];
    to_voxel(A.mask_map, p, v);
    const int n[3] = {A.mx, A.my, A.mz};
    long flat = 0;
    bool on = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        on = on && v[a] > -0.5 && v[a] < static_cast<double>(n[a]) - 0.5;
        flat = flat * n[a] + (on ? static_cast<long>(round(v[a])) : 0);
    }
    return on && static_cast<double>(A.mask[flat]) >= 0.5;
}

__device__ __forceinline__ int clamped_index(double index, int last) {
    if (!(index > 0.0)) return 0;
    if (index >= static_cast<double>(last)) return last;
    return static_cast<int>(index);
}

// Field::trilinear (grid.cpp:66-95) into c[0..nc): nc = NC, or for the wide instantiation
// (WIDE) the image's channel count, at most NC.
template <int NC, bool WIDE>
__device__ __forceinline__ void trilinear(const TrackArgs& A, const double (&p)[3],
                                          double (&c)[NC]) {
    const int nc = WIDE ? A.nc : NC;
    double v[3];
    to_voxel(A.fod_map, p, v);
    const int n[3] = {A.nx, A.ny, A.nz};
    int lo[3], hi[3];
    double w[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double f = floor(v[a]);
        w[a] = v[a] - f;
        lo[a] = clamped_index(f, n[a] - 1);
        hi[a] = clamped_index(f + 1.0, n[a] - 1);
    }
    if constexpr (WIDE) {
        for (int q = 0; q < nc; ++q) c[q] = 0.0;
    } else {
#pragma unroll
        for (int q = 0; q < NC; ++q) c[q] = 0.0;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
            for (int kk = 0; kk < 2; ++kk) {
                const double wx = i ? w[0] : 1.0 - w[0];
                const double wy = j ? w[1] : 1.0 - w[1];
                const double wz = kk ? w[2] : 1.0 - w[2];
                const double weight = wx * wy * wz;
                const long x = i ? hi[0] : lo[0];
                const long y = j ? hi[1] : lo[1];
                const long z = kk ? hi[2] : lo[2];
                const float* corner = A.fod + ((x * A.ny + y) * A.nz + z) * nc;
                if constexpr (WIDE) {
                    for (int q = 0; q < nc; ++q) {
                        c[q] = c[q] + weight * static_cast<double>(corner[q]);
                    }
                } else {
#pragma unroll
                    for (int q = 0; q < NC; ++q) {
                        c[q] = c[q] + weight * static_cast<double>(corner[q]);
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Directions and amplitudes.

__device__ __forceinline__ double norm3(const double (&v)[3]) {
    return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

// rotate_to (maths.cpp:23-42).
__device__ __forceinline__ void rotate_to(const double (&r)[3], const double (&d)[3],
                                          double (&out)[3]) {
    const double xy = sqrt(r[0] * r[0] + r[1] * r[1]);
    if (xy == 0.0) {
        const bool flip = r[2] < 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) out[a] = flip ? -d[a] : d[a];
        return;
    }
    const double h0 = r[0] / xy;
    const double h1 = r[1] / xy;
    const double t0 = r[2] * h0;
    const double t1 = r[2] * h1;
    const double t2 = -xy;
    const double along_z = d[2];
    const double along_h = d[0] * h0 + d[1] * h1;
    out[0] = d[0] + along_z * r[0] + along_h * (t0 - h0);
    out[1] = d[1] + along_z * r[1] + along_h * (t1 - h1);
    out[2] = d[2] + along_z * (r[2] - 1.0) + along_h * (t2 - 0.0);
}

// amplitude (harmonics.cpp:187-219) of the series c along the unit vector d: the zonal terms in
// increasing l, then order by order (m = 1..lmax) the terms in increasing l, each order's
// Legendre values made by its recurrence in l as the sum needs them (fill_legendre's expressions,
// so the same values as a table).
template <int LMAX, bool FULL, int NC>
__device__ __forceinline__ double sh_amplitude_lane(const double (&c)[NC], const double (&d)[3],
                                                    const ShConstantsDev& k) {
    const int lmax = series_lmax<LMAX>(k);
    const double xy = sqrt(d[0] * d[0] + d[1] * d[1]);
    const double cos_t = d[2];
    double cos_phi = 1.0;
    double sin_phi = 0.0;
    if (xy > 0.0) {
        cos_phi = d[0] / xy;
        sin_phi = d[1] / xy;
    }
    double value = 0.0;
    double cm = 1.0;
    double sm = 0.0;
    auto add = [&](int l, int m, double p) {
        if (FULL || l % 2 == 0) {
            if (m == 0) {
                value = value + c[sh_index(l, 0, FULL)] * p;
            } else {
                value = value + M_SQRT2 * p *
                                    (c[sh_index(l, m, FULL)] * cm + c[sh_index(l, -m, FULL)] * sm);
            }
        }
    };
#pragma unroll
    for (int m = 0; m <= LMAX; ++m) {
        if (m > lmax) {
            break;
        }
        if (m > 0) {
            const double next_c = cm * cos_phi - sm * sin_phi;
            const double next_s = sm * cos_phi + cm * sin_phi;
            cm = next_c;
            sm = next_s;
        }
        double lm2 = 0.0;
        double lm1 = k.diagonal[m] * power_int_dev(xy, m);
        add(m, m, lm1);
        if (m < lmax) {
            const double next = k.offdiag[m] * cos_t * lm1;
            lm2 = lm1;
            lm1 = next;
            add(m + 1, m, lm1);
        }
#pragma unroll
        for (int l = m + 2; l <= LMAX; ++l) {
            if (l > lmax) {
                break;
            }
            const double next = k.factor[l * SH_TABLE + m] *
                                (cos_t * lm1 - lm2 / k.factor[(l - 1) * SH_TABLE + m]);
            lm2 = lm1;
            lm1 = next;
            add(l, m, lm1);
        }
    }
    return value;
}

// ---------------------------------------------------------------------------------------------
// Algorithms.

struct LaneStats {
    unsigned long long sampled_steps = 0, accepted_index_sum = 0, truncations = 0;
    double max_truncation = 0.0;
};

// first_direction of det.cpp:36-50, prob.cpp:48-58, testing.cpp:10-17.
template <int LMAX, bool FULL, int ALGO>
__device__ __forceinline__ bool first_direction(const TrackArgs& A, const double (&pos)[3],
                                                LaneDraws& rng, double (&out)[3]) {
    if constexpr (ALGO == TRACK_TEST) {
        rng.unit_vector(out);
        return pos[0] <= A.test_past_x;
    } else {
        constexpr int NC = FULL ? (LMAX + 1) * (LMAX + 1) : (LMAX + 1) * (LMAX + 2) / 2;
        double c[NC];
        trilinear<NC, LMAX == SH_WIDE_LMAX>(A, pos, c);
        if constexpr (ALGO == TRACK_DET) {
            for (int it = 0; it < 20; ++it) {
                double d[3];
                rng.unit_vector(d);
                double amp = 0.0;
                if (sh_climb_lane<LMAX, FULL>(c, d, amp, 50, 1e-4, 0.2, A.k) && amp >= A.cutoff) {
                    out[0] = d[0];
                    out[1] = d[1];
                    out[2] = d[2];
                    return true;
                }
            }
            return false;
        } else {
            for (long attempt = 0; attempt < A.attempts; ++attempt) {
                double d[3];
                rng.unit_vector(d);
                if (sh_amplitude_lane<LMAX, FULL, NC>(c, d, A.k) > A.cutoff) {
                    out[0] = d[0];
                    out[1] = d[1];
                    out[2] = d[2];
                    return true;
                }
            }
            return false;
        }
    }
}

// next_direction of det.cpp:52-74, prob.cpp:61-96, testing.cpp:19-39: STOP_NONE with the new
// direction in dir, or the stop.
template <int LMAX, bool FULL, int ALGO>
__device__ __forceinline__ int next_direction(const TrackArgs& A, const double (&pos)[3],
                                              double (&dir)[3], LaneDraws& rng, LaneStats& st) {
    if constexpr (ALGO == TRACK_TEST) {
        if (A.test_rule == 1) return A.test_stop;
        if (A.test_rule == 2) {
            const double x = -dir[1];
            const double y = dir[0];
            dir[0] = x;
            dir[1] = y;
            dir[2] = 0.0;
        } else if (A.test_rule == 3) {
            const double x = dir[0] + 0.5 * (rng.uniform() - 0.5);
            const double y = dir[1] + 0.5 * (rng.uniform() - 0.5);
            const double z = dir[2] + 0.5 * (rng.uniform() - 0.5);
            const double v[3] = {x, y, z};
            const double n = norm3(v);
            dir[0] = x / n;
            dir[1] = y / n;
            dir[2] = z / n;
        }
        return STOP_NONE;
    } else {
        constexpr int NC = FULL ? (LMAX + 1) * (LMAX + 1) : (LMAX + 1) * (LMAX + 2) / 2;
        double c[NC];
        trilinear<NC, LMAX == SH_WIDE_LMAX>(A, pos, c);
        if constexpr (ALGO == TRACK_DET) {
            double d[3] = {dir[0], dir[1], dir[2]};
            double amp = 0.0;
            if (!sh_climb_lane<LMAX, FULL>(c, d, amp, 50, 1e-4, 0.2, A.k)) {
                return STOP_NO_DIRECTION;
            }
            const double cos_turn = d[0] * dir[0] + d[1] * dir[1] + d[2] * dir[2];
            if (amp < A.cutoff) return STOP_LOW_FOD;
            if (cos_turn < A.cos_max_angle) return STOP_SHARP_TURN;
            dir[0] = d[0];
            dir[1] = d[1];
            dir[2] = d[2];
            return STOP_NONE;
        } else {
            double largest = 0.0;
            for (int g = 0; g < A.n_grid; ++g) {
                const double gd[3] = {A.grid[3 * g], A.grid[3 * g + 1], A.grid[3 * g + 2]};
                double r[3];
                rotate_to(dir, gd, r);
                largest = fmax(largest, sh_amplitude_lane<LMAX, FULL, NC>(c, r, A.k));
            }
            if (largest <= 0.0) return STOP_LOW_FOD;
            const double bound = largest * A.ratio;
            st.sampled_steps += 1ull;
            for (long trial = 0; trial < A.trials; ++trial) {
                double cone[3];
                rng.unit_vector_in_cone(A.max_angle_rad, A.sin_bound, cone);
                double cand[3];
                rotate_to(dir, cone, cand);
                const double a = sh_amplitude_lane<LMAX, FULL, NC>(c, cand, A.k);
                if (isnan(a) || a <= A.cutoff) continue;
                if (a > bound) {
                    st.truncations += 1ull;
                    st.max_truncation = fmax(st.max_truncation, a / bound);
                }
                if (rng.uniform() < a / bound) {
                    st.accepted_index_sum += static_cast<unsigned long long>(trial);
                    const double n = norm3(cand);
                    dir[0] = cand[0] / n;
                    dir[1] = cand[1] / n;
                    dir[2] = cand[2] / n;
                    return STOP_NONE;
                }
            }
            return STOP_NO_DIRECTION;
        }
    }
}

// ---------------------------------------------------------------------------------------------
// The step loop.

__device__ __forceinline__ void write_point(const TrackArgs& A, long row, long i,
                                            const double (&p)[3]) {
    float* out = A.positions + row * (A.max_steps + 1) * 3 + i * 3;
    out[0] = static_cast<float>(p[0]);
    out[1] = static_cast<float>(p[1]);
    out[2] = static_cast<float>(p[2]);
}

__device__ __forceinline__ void finish_seed(const TrackArgs& A, long s, long n_points, int first,
                                            int second, const double (&dir)[3]) {
    A.counts[s] = n_points;
    A.status[2 * s] = static_cast<unsigned char>(first);
    A.status[2 * s + 1] = static_cast<unsigned char>(second);
    A.followed[3 * s + 0] = dir[0];
    A.followed[3 * s + 1] = dir[1];
    A.followed[3 * s + 2] = dir[2];
}

template <int LMAX, bool FULL, int ALGO>
__global__ void __launch_bounds__(WAVE * 4) track_kernel(TrackArgs A) {
    const int l = lane_id();
    bool active = false;  // the lane holds a streamline in progress
    bool drained = l >= A.lanes; // the queue had no seed left for this lane (or it takes none)
    long s = 0;           // the lane's seed
    double seed_pos[3] = {0.0, 0.0, 0.0};
    double start_dir[3] = {0.0, 0.0, 0.0};
    double pos[3] = {0.0, 0.0, 0.0};
    double dir[3] = {0.0, 0.0, 0.0};
    long n_points = 0;
    int half = 0;
    int first_stop = STOP_NONE;
    LaneDraws rng{0ull, 0ull};
    LaneStats st;

    while (true) {
        // Lanes without a streamline take seeds until one starts or the queue is empty.
        while (true) {
            const bool need = !active && !drained;
            const unsigned long long want = ballot(need);
            if (want == 0ull) break;
            const int leader = __ffsll(static_cast<long long>(want)) - 1;
            int base = 0;
            if (l == leader) {
                base = atomicAdd(A.next_seed, __popcll(want));
            }
            base = readlane(base, leader);
            if (need) {
                s = static_cast<long>(base) + __popcll(want & ((1ull << l) - 1ull));
                if (s >= A.n_seeds) {
                    drained = true;
                } else {
                    seed_pos[0] = A.seeds[3 * s];
                    seed_pos[1] = A.seeds[3 * s + 1];
                    seed_pos[2] = A.seeds[3 * s + 2];
                    const double zero[3] = {0.0, 0.0, 0.0};
                    bool started = grid_contains(A.fod_map, A.nx, A.ny, A.nz, seed_pos) &&
                                   (A.mask == nullptr || mask_inside(A, seed_pos));
                    if (started) {
                        rng.key = splitmix_finalise(splitmix_finalise(A.run_seed) +
                                                    static_cast<unsigned long long>(A.indices[s]));
                        rng.counter = 0ull;
                        if (A.directions != nullptr) {
                            start_dir[0] = A.directions[3 * s];
                            start_dir[1] = A.directions[3 * s + 1];
                            start_dir[2] = A.directions[3 * s + 2];
                        } else {
                            LaneDraws search{rng.key, 1ull << 63};
                            started = first_direction<LMAX, FULL, ALGO>(A, seed_pos, search,
                                                                         start_dir);
                        }
                    }
                    if (started) {
                        write_point(A, s, 0, seed_pos);
                        n_points = 1;
                        half = 0;
                        first_stop = STOP_NONE;
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            pos[a] = seed_pos[a];
                            dir[a] = start_dir[a];
                        }
                        active = true;
                    } else {
                        finish_seed(A, s, 0, STOP_NONE, STOP_NONE, zero);
                    }
                }
            }
        }
        if (ballot(active) == 0ull) break;

        if (active) {
            // One step of track_one_way (tracker.cpp:42-71).
#pragma unroll
            for (int a = 0; a < 3; ++a) pos[a] = pos[a] + A.step * dir[a];
            int stop = STOP_NONE;
            if (!grid_contains(A.fod_map, A.nx, A.ny, A.nz, pos)) {
                stop = STOP_LEFT_VOLUME;
            } else if (A.mask != nullptr && !mask_inside(A, pos)) {
                stop = STOP_LEFT_MASK;
            } else {
                write_point(A, s, n_points, pos);
                n_points += 1;
                if (n_points == A.max_steps + 1) {
                    stop = STOP_MAX_LENGTH;
                } else {
                    stop = next_direction<LMAX, FULL, ALGO>(A, pos, dir, rng, st);
                }
            }
            if (stop != STOP_NONE) {
                // The end of a half (track_seed, tracker.cpp:73-121).
                if (half == 1) {
                    finish_seed(A, s, n_points, first_stop, stop, start_dir);
                    active = false;
                } else if (!A.bidirectional) {
                    finish_seed(A, s, n_points, STOP_NONE, stop, start_dir);
                    active = false;
                } else {
                    float* row = A.positions + s * (A.max_steps + 1) * 3;
                    for (long i = 0; i < n_points / 2; ++i) {
                        const long j = n_points - 1 - i;
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            const float t = row[3 * i + a];
                            row[3 * i + a] = row[3 * j + a];
                            row[3 * j + a] = t;
                        }
                    }
                    first_stop = stop;
                    if (n_points == A.max_steps + 1) {
                        finish_seed(A, s, n_points, stop, STOP_MAX_LENGTH, start_dir);
                        active = false;
                    } else {
                        half = 1;
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            pos[a] = seed_pos[a];
                            dir[a] = -start_dir[a];
                        }
                    }
                }
            }
        }
    }

    // The wave's counts, then one atomic each.
    unsigned long long sampled = st.sampled_steps;
    unsigned long long index_sum = st.accepted_index_sum;
    unsigned long long truncations = st.truncations;
    double worst = st.max_truncation;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        sampled += __shfl_xor(sampled, o);
        index_sum += __shfl_xor(index_sum, o);
        truncations += __shfl_xor(truncations, o);
        worst = fmax(worst, __shfl_xor(worst, o));
    }
    if (l == 0) {
        if (sampled) atomicAdd(A.stats + 0, sampled);
        if (index_sum) atomicAdd(A.stats + 1, index_sum);
        if (truncations) atomicAdd(A.stats + 2, truncations);
        // Non-negative doubles order as their bit patterns.
        if (worst > 0.0) atomicMax(A.stats + 3, static_cast<unsigned long long>(__double_as_longlong(worst)));
    }
}

} // namespace lsd
