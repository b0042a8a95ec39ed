// The projection of FOD amplitudes onto SH (model.py:412-440): coeffs = amps P, with P the
// least-squares projection of the basis on the amplitude directions, scaled by the density;
// with exact_integral, c00 = sum(amps) / sqrt(4 pi).
//
// One wave computes a tile of 16 voxels x 16 TC coefficients with V_MFMA_F64_16X16X4_F64 (TC
// 16 x 16 accumulators, TC up to 6: lmax 12).  The sum over the amplitude index k is taken in an order that lets
// each lane load a contiguous 32-byte chunk of its voxel's row: in a block of 16 k, lane group
// q = lane / 16 holds k0 + 4q .. k0 + 4q + 3, and MFMA step s contracts k = k0 + 4q + s over
// q, so the B operand is P's row k0 + 4q + s in lane (q, j).  Layouts (wave64, f64 16x16x4):
// A[i][q] in lane 16 q + i, B[q][j] in lane 16 q + j, D[l / 16 + 4 r][l % 16] in lane l,
// element r (checked on the device with known operands).  c00's row sum is the lane's partial sums combined across the four lane groups.
#include "kernels.hpp"

namespace lsd {

namespace {

using d4 = __attribute__((ext_vector_type(4))) double;

struct ProjectArgs {
    const double* amps; // (n, k)
    long n;
    int k;
    const double* p;    // (k, c)
    int c;
    int exact_integral;
    double* out;        // (n, c)
};

// TC tiles of 16 coefficients.
template <int TC>
__global__ void __launch_bounds__(WAVE * WAVES_PER_BLOCK) sh_project_kernel(ProjectArgs A) {
    constexpr int SH_TILE_COLS = TC;
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / WAVE);
    const long tile = static_cast<long>(blockIdx.x) * WAVES_PER_BLOCK + w;
    const long v0 = tile * 16;
    if (v0 >= A.n) {
        return;
    }
    const int l = lane_id();
    const int i = l % 16;
    const int q = l / 16;
    const long row = v0 + i;
    const bool row_ok = row < A.n;
    const double* a_row = A.amps + (row_ok ? row : 0) * A.k;
    d4 acc[SH_TILE_COLS];
#pragma unroll
    for (int t = 0; t < SH_TILE_COLS; ++t) {
        acc[t] = d4{0.0, 0.0, 0.0, 0.0};
    }
    double row_sum = 0.0;
    for (int k0 = 0; k0 < A.k; k0 += 16) {
        double a[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int kk = k0 + 4 * q + s;
            a[s] = (row_ok && kk < A.k) ? a_row[kk] : 0.0;
            row_sum = row_sum + a[s];
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int kk = k0 + 4 * q + s;
#pragma unroll
            for (int t = 0; t < SH_TILE_COLS; ++t) {
                const int j = 16 * t + i;
                const double b = (kk < A.k && j < A.c) ? A.p[static_cast<long>(kk) * A.c + j] : 0.0;
                acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b, acc[t], 0, 0, 0);
            }
        }
    }
    // The row sums: lanes i, i + 16, i + 32, i + 48 hold row i's partials.
    row_sum = row_sum + __shfl_xor(row_sum, 16);
    row_sum = row_sum + __shfl_xor(row_sum, 32);
    const double inv_norm = 1.0 / sqrt(4.0 * M_PI);
#pragma unroll
    for (int t = 0; t < SH_TILE_COLS; ++t) {
        const int j = 16 * t + i;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int out_row = q + 4 * r;
            const long v = v0 + out_row;
            // Every lane takes part in the shuffle (lane out_row holds that row's sum).
            const double total = __shfl(row_sum, out_row);
            double value = acc[t][r];
            if (A.exact_integral && j == 0) {
                value = total * inv_norm;
            }
            if (v < A.n && j < A.c) {
                A.out[v * A.c + j] = value;
            }
        }
    }
}

} // namespace

int launch_sh_project(const double* amps, long n, int k, const double* p, int c,
                      int exact_integral, double* out, hipStream_t stream) {
    ProjectArgs a{amps, n, k, p, c, exact_integral, out};
    const int waves = static_cast<int>((n + 15) / 16);
    const int tiles = (c + 15) / 16;
    switch (tiles) {
    case 0:
    case 1: return launch("sh_project", sh_project_kernel<1>, a, waves, 0, stream);
    case 2: return launch("sh_project", sh_project_kernel<2>, a, waves, 0, stream);
    case 3: return launch("sh_project", sh_project_kernel<3>, a, waves, 0, stream);
    case 4:
    case 5: return launch("sh_project", sh_project_kernel<5>, a, waves, 0, stream);
    case 6: return launch("sh_project", sh_project_kernel<6>, a, waves, 0, stream);
    default:
        return fail("the SH projection takes at most 96 coefficients (lmax 12), got " +
                    std::to_string(c));
    }
}

} // namespace lsd
