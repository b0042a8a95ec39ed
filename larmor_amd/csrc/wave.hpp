// Wavefront-level building blocks for the CDNA4 (gfx950) spherical-deconvolution kernels.
//
// One 64-lane wavefront works on one voxel.  Two layouts spread a voxel's vectors over the
// lanes:
//  - RowVec<RPL>: a vector indexed by measurement (row) r lives in lane r % 64, register slot
//    r / 64, so a voxel's m measurements take RPL = ceil(m / 64) registers per lane;
//  - LaneVec<NS>: a vector indexed by solver column or factor position p lives in lane p % 64,
//    slot p / 64.
// A uniform (wave-wide) index into either is resolved with readlane from an ext_vector_type
// register vector, so no register array is ever indexed dynamically.
//
// Sums over rows use one fixed order, the "wave sum": each lane adds its rows in increasing order
// from 0, then the 64 partials are combined by an XOR butterfly p = p + p[lane ^ o] for
// o = 32, 16, 8, 4, 2, 1.  The CPU oracle reproduces this order, which makes the GPU results
// bit-identical to it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

// The arithmetic type of the solve kernels: double (namespace lsd), or float (namespace lsd32) in
// the translation units of the plain FP32 form, which define LSD_FP32 before the first include.
#ifdef LSD_FP32
#define LSD_NS lsd32
#else
#define LSD_NS lsd
#endif

namespace LSD_NS {

#ifdef LSD_FP32
using real = float;
using real2 = float2;
#else
using real = double;
using real2 = double2;
#endif

constexpr int WAVE = 64;

__device__ __forceinline__ int lane_id() { return static_cast<int>(__lane_id()); }

// The value of v in lane src (uniform src).
__device__ __forceinline__ double readlane(double v, int src) {
    const auto bits = static_cast<unsigned long long>(__double_as_longlong(v));
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(bits & 0xffffffffull), src);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(bits >> 32), src);
    const unsigned long long out =
        (static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) |
        static_cast<unsigned>(lo);
    return __longlong_as_double(static_cast<long long>(out));
}

__device__ __forceinline__ int readlane(int v, int src) { return __builtin_amdgcn_readlane(v, src); }
__device__ __forceinline__ float readlane(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

// The value of v in lane src, src per lane.
__device__ __forceinline__ double gather(double v, int src) { return __shfl(v, src); }
__device__ __forceinline__ float gather(float v, int src) { return __shfl(v, src); }
__device__ __forceinline__ int gather(int v, int src) { return __shfl(v, src); }

__device__ __forceinline__ unsigned long long ballot(bool p) { return __ballot(p); }

// The value of x in lane lane_id() ^ o, for o = 1, 2, 4, 8, 16 or 32 (a constant once the caller is
// unrolled), with all 64 lanes active.  o = 1, 2 and 8 are DPP moves within a 16-lane row (the
// quad permutations and the rotation by 8); o = 4 and 16 are ds_swizzle in its bit mode (lane ^ o
// within each 32-lane half), which runs on the LDS pipe without an address register and so leaves
// the VALU, which bounds the solve kernels, free; o = 32 is gfx950's swap of the upper half of the
// first operand with the lower half of the second: with x as both operands, a lane of the lower
// half finds its partner in the second result, one of the upper half in the first.  The partner
// is exactly lane ^ o, as __shfl_xor's is.
__device__ __forceinline__ unsigned xor_lanes(unsigned x, int o) {
    const int l = lane_id();
    const int xi = static_cast<int>(x);
    if (o == 1) {
        return static_cast<unsigned>(__builtin_amdgcn_mov_dpp(xi, 0xB1, 0xf, 0xf, false));  // quad_perm [1,0,3,2]
    }
    if (o == 2) {
        return static_cast<unsigned>(__builtin_amdgcn_mov_dpp(xi, 0x4E, 0xf, 0xf, false));  // quad_perm [2,3,0,1]
    }
    if (o == 4) {
        // and_mask 0x1f, or_mask 0, xor_mask 4
        return static_cast<unsigned>(__builtin_amdgcn_ds_swizzle(xi, (4 << 10) | 0x1f));
    }
    if (o == 8) {
        return static_cast<unsigned>(__builtin_amdgcn_mov_dpp(xi, 0x128, 0xf, 0xf, false));  // row_ror:8
    }
    if (o == 16) {
        // and_mask 0x1f, or_mask 0, xor_mask 16
        return static_cast<unsigned>(__builtin_amdgcn_ds_swizzle(xi, (16 << 10) | 0x1f));
    }
    const auto s = __builtin_amdgcn_permlane32_swap(x, x, false, false);
    return (l & 32) ? s[0] : s[1];
}

__device__ __forceinline__ int xor_lanes(int x, int o) {
    return static_cast<int>(xor_lanes(static_cast<unsigned>(x), o));
}

__device__ __forceinline__ float xor_lanes(float x, int o) {
    return __uint_as_float(xor_lanes(__float_as_uint(x), o));
}

__device__ __forceinline__ double xor_lanes(double x, int o) {
    const auto bits = static_cast<unsigned long long>(__double_as_longlong(x));
    const unsigned lo = xor_lanes(static_cast<unsigned>(bits & 0xffffffffull), o);
    const unsigned hi = xor_lanes(static_cast<unsigned>(bits >> 32), o);
    return __longlong_as_double(
        static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// Orders this wave's LDS accesses across lanes: writes before it are seen by reads after it.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// The canonical butterfly of 64 lane partials.  Every lane computes the same sum (each step adds
// the same two values, in either order); lane 0's, read back, is provably uniform to the
// compiler, which keeps the control flow that depends on it scalar.
__device__ __forceinline__ real butterfly_sum(real p) {
#pragma unroll
    for (int o = WAVE / 2; o >= 1; o >>= 1) {
        p = p + xor_lanes(p, o);
    }
    return readlane(p, 0);
}

// max over lanes (exact, order-free; with NaN values, lane 0's result).
__device__ __forceinline__ real wave_max(real p) {
#pragma unroll
    for (int o = WAVE / 2; o >= 1; o >>= 1) {
        const real q = xor_lanes(p, o);
        p = q > p ? q : p;
    }
    return readlane(p, 0);
}

// The largest value and the smallest index holding it, over lanes; idx < 0 means "none".
// (With NaN values the lanes' results could differ; lane 0's is the one returned.)
__device__ __forceinline__ void wave_argmax_first(real& value, int& idx) {
#pragma unroll
    for (int o = WAVE / 2; o >= 1; o >>= 1) {
        const real ov = xor_lanes(value, o);
        const int oi = xor_lanes(idx, o);
        const bool take = oi >= 0 && (idx < 0 || ov > value || (ov == value && oi < idx));
        value = take ? ov : value;
        idx = take ? oi : idx;
    }
    // Every lane holds the same result; lane 0's, read back, is provably uniform.
    value = readlane(value, 0);
    idx = readlane(idx, 0);
}

// The smallest value and the smallest index holding it, over lanes; idx < 0 means "none".
__device__ __forceinline__ void wave_argmin_first(real& value, int& idx) {
#pragma unroll
    for (int o = WAVE / 2; o >= 1; o >>= 1) {
        const real ov = xor_lanes(value, o);
        const int oi = xor_lanes(idx, o);
        const bool take = oi >= 0 && (idx < 0 || ov < value || (ov == value && oi < idx));
        value = take ? ov : value;
        idx = take ? oi : idx;
    }
    // Every lane holds the same result; lane 0's, read back, is provably uniform.
    value = readlane(value, 0);
    idx = readlane(idx, 0);
}

// The largest value among the flagged entries of a column-spread vector (column l + 64 t in
// lane l, slot t) and the smallest column holding it: idx = -1 when none is flagged.  One max
// butterfly, then a ballot per slot for the first column that holds the maximum.  For values
// without NaN this is the in-order scan's "first strictly larger" choice; a NaN value is never
// chosen.
template <int NS>
__device__ __forceinline__ void wave_argmax_cols(const real (&v)[NS], const bool (&flag)[NS],
                                                 real& value, int& idx) {
    real lane_max = -__builtin_inff();
    bool any = false;
#pragma unroll
    for (int t = 0; t < NS; ++t) {
        any = any || flag[t];
        lane_max = (flag[t] && v[t] > lane_max) ? v[t] : lane_max;
    }
    value = wave_max(lane_max);
    idx = -1;
    if (ballot(any) != 0ull) {
#pragma unroll
        for (int t = NS - 1; t >= 0; --t) {
            const unsigned long long hit = ballot(flag[t] && v[t] == value);
            idx = hit != 0ull ? WAVE * t + __ffsll(static_cast<long long>(hit)) - 1 : idx;
        }
    }
}

// The same for the smallest value.
template <int NS>
__device__ __forceinline__ void wave_argmin_cols(const real (&v)[NS], const bool (&flag)[NS],
                                                 real& value, int& idx) {
    real lane_min = __builtin_inff();
    bool any = false;
#pragma unroll
    for (int t = 0; t < NS; ++t) {
        any = any || flag[t];
        lane_min = (flag[t] && v[t] < lane_min) ? v[t] : lane_min;
    }
#pragma unroll
    for (int o = WAVE / 2; o >= 1; o >>= 1) {
        const real q = xor_lanes(lane_min, o);
        lane_min = q < lane_min ? q : lane_min;
    }
    value = readlane(lane_min, 0);
    idx = -1;
    if (ballot(any) != 0ull) {
#pragma unroll
        for (int t = NS - 1; t >= 0; --t) {
            const unsigned long long hit = ballot(flag[t] && v[t] == value);
            idx = hit != 0ull ? WAVE * t + __ffsll(static_cast<long long>(hit)) - 1 : idx;
        }
    }
}

// Register vectors.  Element selection by a run-time (uniform) index stays in registers for
// ext_vector_type values, where a plain array would be moved to scratch memory.
template <int N>
using dvec = double __attribute__((ext_vector_type(N)));
template <class T, int N>
using vec_t = T __attribute__((ext_vector_type(N)));
template <int N>
using rvec = vec_t<real, N>;

// A vector over rows: row r in lane r % 64, slot r / 64.
template <int RPL>
struct RowVec {
    rvec<RPL> v;

    __device__ __forceinline__ void fill(real x) {
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            v[t] = x;
        }
    }
    // Row r (uniform r) as a wave-wide value.
    __device__ __forceinline__ real get(int r) const {
        const real x = v[r / WAVE];
        return readlane(x, r % WAVE);
    }
    // Sets row r (uniform r) to x (uniform x).
    __device__ __forceinline__ void set(int r, real x) {
        if (lane_id() == r % WAVE) {
            v[r / WAVE] = x;
        }
    }
};

// The wave sum over rows [begin, end) of the per-row terms held in slots.
template <int RPL>
__device__ __forceinline__ real wave_sum_rows(const real (&terms)[RPL], int begin, int end) {
    const int l = lane_id();
    real s = 0;
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
        const int r = l + WAVE * t;
        if (r >= begin && r < end) {
            s = s + terms[t];
        }
    }
    return butterfly_sum(s);
}

// A vector over columns or positions: index p in lane p % 64, slot p / 64.
template <int NS, class T = real>
struct LaneVec {
    vec_t<T, NS> v;

    __device__ __forceinline__ void fill(T x) {
#pragma unroll
        for (int t = 0; t < NS; ++t) {
            v[t] = x;
        }
    }
    __device__ __forceinline__ T get(int p) const {
        const T x = v[p / WAVE];
        return readlane(x, p % WAVE);
    }
    __device__ __forceinline__ void set(int p, T x) {
        if (lane_id() == p % WAVE) {
            v[p / WAVE] = x;
        }
    }
};

// hypot as m * sqrt(1 + (s / m)^2), m the larger magnitude; the oracle uses the same formula.
__device__ __forceinline__ real hypot_fixed(real a, real b) {
    const real x = fabs(a);
    const real y = fabs(b);
    const real big = x > y ? x : y;
    const real small = x > y ? y : x;
    if (big == real(0)) {
        return real(0);
    }
    const real r = small / big;
    return big * sqrt(real(1) + r * r);
}

#ifdef LSD_FP32
// The plain FP32 form divides with the compiler's single-precision division everywhere; these
// keep the FP64 form's call sites.
__device__ __forceinline__ bool divisor_in_range(float d) { return d > 0.0f; }
__device__ __forceinline__ bool dividend_in_range(float) { return true; }
__device__ __forceinline__ float division_reciprocal(float d) { return 1.0f / d; }
__device__ __forceinline__ float divide_in_range(float a, float d, float) { return a / d; }
__device__ __forceinline__ float divide_by(float a, float d, float) { return a / d; }
__device__ __forceinline__ float divide_nonzero_by(float a, float d, float, bool) {
    return a != 0.0f ? a / d : a;
}
#else
// Division by a divisor used many times, split in the two halves of gfx950's double-division
// sequence (V_DIV_SCALE, V_RCP and two Newton steps on the reciprocal; then the quotient a y, the
// exact residual a - d (a y), the correction by V_DIV_FMAS and V_DIV_FIXUP): the reciprocal is
// refined once per divisor and each quotient then costs three operations.  Where the sequence's
// scaling and fix-up steps are the identity -- a divisor in [2^-100, 2^100) and a dividend of
// magnitude in [2^-800, 2^600), so that no operand, the reciprocal or the quotient is subnormal or
// near overflow and the exponents differ by less than 768 -- the three operations are the
// sequence's own and the quotient is a / d to the bit.  Outside that range (zero, subnormal, huge
// or non-finite operands) the callers divide.
__device__ __forceinline__ int exponent_field(double x) {
    return static_cast<int>((__double_as_longlong(x) >> 52) & 0x7ff);
}

__device__ __forceinline__ bool divisor_in_range(double d) {
    return static_cast<unsigned>(exponent_field(d) - (1023 - 100)) < 200u && d > 0.0;
}

__device__ __forceinline__ bool dividend_in_range(double a) {
    return static_cast<unsigned>(exponent_field(a) - (1023 - 800)) < 1400u;
}

// The refined reciprocal of d.
__device__ __forceinline__ double division_reciprocal(double d) {
    double y = __builtin_amdgcn_rcp(d);
    double e = fma(-d, y, 1.0);
    y = fma(y, e, y);
    e = fma(-d, y, 1.0);
    return fma(y, e, y);
}

// a / d from y = division_reciprocal(d), both operands in range.
__device__ __forceinline__ double divide_in_range(double a, double d, double y) {
    const double q = a * y;
    const double r = fma(-d, q, a);
    return fma(r, y, q);
}

// a / d given y = division_reciprocal(d).
__device__ __forceinline__ double divide_by(double a, double d, double y) {
    return (divisor_in_range(d) && dividend_in_range(a)) ? divide_in_range(a, d, y) : a / d;
}

// The triangular solves' a != 0 ? a / d : a, with d_ok = divisor_in_range(d) known (a zero
// dividend is never in range).
__device__ __forceinline__ double divide_nonzero_by(double a, double d, double y, bool d_ok) {
    if (d_ok && dividend_in_range(a)) {
        return divide_in_range(a, d, y);
    }
    return a != 0.0 ? a / d : a;
}
#endif

} // namespace LSD_NS
