// The mixed-precision form's low-resolution correlations on the matrix cores.
//
// corr = K^T a for a list of voxels, a = max(w s - W_iso p, 0) (multi_tissue.cpp:20-24), as one
// GEMM (voxels x m) (m x n) with V_MFMA_F32_16X16X32_BF16 and FP32 accumulation.  Both operands
// are split into BF16 halves, x = hi + lo with hi = bf16(x) and lo = bf16(x - hi), and the
// refinement step adds the cross products: corr = hi_a . hi_K + (lo_a . hi_K + hi_a . lo_K),
// the two sums in separate FP32 accumulators added in FP64 at the end.  Each product then carries
// about 2^-17 of relative error instead of BF16's 2^-8 (the dropped lo . lo term and the lo halves'
// own rounding); the FP64 refinement rounds of the fit remove what remains.
//
// Layout.  A wave computes 16 voxels (list positions) by up to 64 columns: 16 x 16 output tiles,
// NT of them.  In the instruction's operand layout lane l holds, for its A operand, voxel l % 16
// and the 8 consecutive rows 8 (l / 16) + i of the 32-row chunk; for its B operand, column
// l % 16 of the tile and the same 8 rows; and of the 16 x 16 result, column l % 16 and voxels
// 4 (l / 16) + i.  The kernel's halves are packed on the host in exactly that B order, chunk by
// chunk and tile by tile (64 lanes x 8 BF16 per tile), so each lane loads its 16 bytes directly.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace lsd {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));

constexpr int MFMA_ROWS = 32;   // rows (measurements) per MFMA step
constexpr int MFMA_TILE = 16;   // voxels and columns per output tile
constexpr int MFMA_MAX_TILES = 4;

// The BF16 bit pattern nearest to f (round to nearest, ties to even; finite inputs).
__host__ __device__ __forceinline__ unsigned short bf16_bits(float f) {
    unsigned u = 0;
    __builtin_memcpy(&u, &f, sizeof(u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<unsigned short>(u >> 16);
}

__host__ __device__ __forceinline__ float bf16_value(unsigned short b) {
    const unsigned u = static_cast<unsigned>(b) << 16;
    float f = 0.0f;
    __builtin_memcpy(&f, &u, sizeof(f));
    return f;
}

// The halves of x: hi = bf16(x), lo = bf16(x - hi).
__host__ __device__ __forceinline__ void bf16_split(double x, unsigned short& hi,
                                                    unsigned short& lo) {
    hi = bf16_bits(static_cast<float>(x));
    lo = bf16_bits(static_cast<float>(x - static_cast<double>(bf16_value(hi))));
}

struct CorrMfmaArgs {
    const double* signals;    // (n_vox, m)
    int m;
    const double* weights;    // (m)
    const double* w_iso;      // (m, ci)
    const double* iso;        // (n_vox, ci)
    int ci;
    const unsigned short* k_hi; // (chunks, NT, 64, 8) the kernel's BF16 halves in B order
    const unsigned short* k_lo;
    int n;
    int chunks;               // ceil(m / 32)
    const int* list;          // voxels by list position, or null for all
    const int* list_count;
    long n_vox;
    double* corr;             // (list position, n)
};

template <int NT>
__global__ void __launch_bounds__(WAVE * WAVES_PER_BLOCK) corr_bf16_kernel(CorrMfmaArgs A) {
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / WAVE);
    const int l = lane_id();
    const int vi = l % MFMA_TILE;        // the lane's voxel in the A operand
    const int kg = l / MFMA_TILE;        // its 8-row group within a chunk
    const long count = A.list ? static_cast<long>(*A.list_count) : A.n_vox;
    const long gw = static_cast<long>(blockIdx.x) * WAVES_PER_BLOCK + w;
    const long stride = static_cast<long>(gridDim.x) * WAVES_PER_BLOCK * MFMA_TILE;
    const bf16x8_t* khi = reinterpret_cast<const bf16x8_t*>(A.k_hi);
    const bf16x8_t* klo = reinterpret_cast<const bf16x8_t*>(A.k_lo);
    for (long base = gw * MFMA_TILE; base < count; base += stride) {
        const long it = base + vi;
        const bool on = it < count;
        const long v = on ? (A.list ? A.list[it] : it) : 0;
        double p[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            p[q] = (on && q < A.ci) ? A.iso[v * A.ci + q] : 0.0;
        }
        floatx4_t hh[NT];
        floatx4_t cross[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            hh[t] = floatx4_t{0.0f, 0.0f, 0.0f, 0.0f};
            cross[t] = floatx4_t{0.0f, 0.0f, 0.0f, 0.0f};
        }
        for (int c = 0; c < A.chunks; ++c) {
            bf16x8_t ah;
            bf16x8_t al;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int r = c * MFMA_ROWS + kg * 8 + i;
                double a = 0.0;
                if (on && r < A.m) {
                    const double ws = A.signals[v * A.m + r] * A.weights[r];
                    double wp = 0.0;
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        if (q < A.ci) {
                            wp = wp + A.w_iso[static_cast<long>(r) * A.ci + q] * p[q];
                        }
                    }
                    a = fmax(ws - wp, 0.0);
                }
                unsigned short hi = 0;
                unsigned short lo = 0;
                bf16_split(a, hi, lo);
                ah[i] = __builtin_bit_cast(__bf16, hi);
                al[i] = __builtin_bit_cast(__bf16, lo);
            }
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const long at = (static_cast<long>(c) * NT + t) * WAVE + l;
                const bf16x8_t bh = khi[at];
                const bf16x8_t bl = klo[at];
                hh[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, hh[t], 0, 0, 0);
                cross[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, cross[t], 0, 0, 0);
                cross[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, cross[t], 0, 0, 0);
            }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int col = t * MFMA_TILE + l % MFMA_TILE;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const long pos = base + (l / MFMA_TILE) * 4 + i;
                if (pos < count && col < A.n) {
                    A.corr[pos * A.n + col] =
                        static_cast<double>(hh[t][i]) + static_cast<double>(cross[t][i]);
                }
            }
        }
    }
}

// The packed BF16 halves of a kernel K (m x n row-major, n <= 64) in the B operand order above:
// element i of lane l of tile t of chunk c is K[32 c + 8 (l / 16) + i][16 t + l % 16], zero
// outside K.  Returns the number of tiles.
inline int pack_bf16_halves(const double* k, int m, int n, std::vector<unsigned short>& hi,
                            std::vector<unsigned short>& lo) {
    const int nt = (n + MFMA_TILE - 1) / MFMA_TILE;
    const int chunks = (m + MFMA_ROWS - 1) / MFMA_ROWS;
    hi.assign(static_cast<size_t>(chunks) * nt * WAVE * 8, 0);
    lo.assign(hi.size(), 0);
    for (int c = 0; c < chunks; ++c) {
        for (int t = 0; t < nt; ++t) {
            for (int ln = 0; ln < WAVE; ++ln) {
                for (int i = 0; i < 8; ++i) {
                    const int r = c * MFMA_ROWS + (ln / MFMA_TILE) * 8 + i;
                    const int j = t * MFMA_TILE + ln % MFMA_TILE;
                    const size_t at = ((static_cast<size_t>(c) * nt + t) * WAVE + ln) * 8 + i;
                    if (r < m && j < n) {
                        bf16_split(k[static_cast<long>(r) * n + j], hi[at], lo[at]);
                    }
                }
            }
        }
    }
    return nt;
}

} // namespace lsd
