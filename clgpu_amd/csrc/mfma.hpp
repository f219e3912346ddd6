// Device-side building blocks for the gfx950 (CDNA4) level-3 kernels:
// vector types, MFMA wrappers, the FP32 -> BF16 hi/lo split and the
// XCD-aware workgroup -> tile mapping.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "level3_tiles.hpp"

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

namespace icl {

// v_mfma_f32_32x32x2_f32. Lane l supplies a = A[l%32][l/32] (32x2 operand) and
// b = B[l/32][l%32] (2x32 operand). D[i][j] for i = 8*(r/4) + 4*(l/32) + r%4,
// j = l%32 is register r of lane l.
__device__ __forceinline__ f32x16 mfma_32x32x2_f32(float a, float b, f32x16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}

// v_mfma_f32_32x32x16_bf16. Lane l supplies A[l%32][8*(l/32) .. +7] and
// B[8*(l/32) .. +7][l%32]; D has the same lane layout as the f32 32x32 form.
__device__ __forceinline__ f32x16 mfma_32x32x16_bf16(bf16x8 a, bf16x8 b, f32x16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// v_mfma_f64_16x16x4_f64. Lane l supplies A[l%16][l/16] and B[l/16][l%16];
// D[i][j] for i = 4*r + l/16, j = l%16 is register r of lane l (one row per
// register, unlike the FP32 16x16 form whose registers hold 4 consecutive rows).
__device__ __forceinline__ f64x4 mfma_16x16x4_f64(double a, double b, f64x4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Row index (within a 32x32 MFMA block) held by register r of lane `lane`.
__device__ __forceinline__ int mfma32_row(int lane, int r)
{
    return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
}

// x = hi + lo + e with hi = bf16(x), lo = bf16(x - hi) (both round-to-nearest-even)
// and |e| <= 2^-17 |x| for normal x. x - hi is exact in FP32. Non-finite x
// gives lo = 0 so that inf stays inf instead of turning into inf - inf = NaN.
__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo)
{
    hi = (__bf16)x;
    float r = x - (float)hi;
    lo = (__bf16)(__builtin_isfinite(x) ? r : 0.0f);
}

// Physical 16-byte chunk slot of logical chunk c of a 128-byte LDS row r is
// c ^ lds_swizzle(r). A 16-byte LDS read serves 16 lanes per cycle from 64
// four-byte banks (256 bytes), i.e. from two 128-byte rows' worth of banks; the
// 16 lanes of a fragment read take 16 consecutive rows at one logical chunk,
// and (r & 1, c ^ ((r >> 1) & 7)) is distinct over any 16 consecutive rows
// starting at a multiple of 16, so those reads are free of bank conflicts.
__device__ __forceinline__ int lds_swizzle(int r)
{
    return (r >> 1) & 7;
}

__device__ __forceinline__ bf16x8 bf16x8_neg(bf16x8 v)
{
    u32x4 u = __builtin_bit_cast(u32x4, v);
    u ^= 0x80008000u;
    return __builtin_bit_cast(bf16x8, u);
}

// Maps blockIdx.x to a (tile_m, tile_n) pair.
// 1) The dispatcher hands consecutive workgroup ids to the 8 XCDs round-robin,
//    and each XCD has a private L2. Renumbering id -> (id % 8) * (n / 8) + id / 8
//    gives every XCD a contiguous range of logical ids.
// 2) Logical ids are rastered in groups of GROUP tile-rows so that workgroups
//    running together on one XCD share A row-panels and B column-panels in L2.
template <int GROUP>
__device__ __forceinline__ void tile_coords(int tiles_m, int tiles_n, int& tm, int& tn)
{
    const int nb = gridDim.x;
    int b = blockIdx.x;
    if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
    const int per_group = GROUP * tiles_n;
    const int g = b / per_group;
    const int first = g * GROUP;
    const int gsz = min(tiles_m - first, GROUP);
    const int r = b - g * per_group;
    tm = first + r % gsz;
    tn = r / gsz;
}

// Tile numbering of the level-3 launches (the grid holds only the tiles that
// do work, so it must come from tri_tile_count on the host):
//  - write_tri: the tiles meeting the written triangle, tile-row by tile-row;
//    every XCD takes a contiguous share of that list (equal work per XCD,
//    whereas contiguous bands of tile rows would load one XCD with most of the
//    triangle);
//  - kband: all tiles, longest k window first, so that the round-robin
//    dispatch over the XCDs spreads the long and the short tiles evenly and
//    the short ones fill the tail.
// Returns false (the workgroup has no tile) only for ids beyond that count.
template <int BM, int BN>
__device__ __forceinline__ bool tile_coords_level3(int write_tri, int kband, int tiles_m, int tiles_n, int& tm,
                                                   int& tn)
{
    const int nb = gridDim.x;
    int b = blockIdx.x;
    if (write_tri != 0) {
        if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
        for (int i = 0; i < tiles_m; ++i) {
            int lo, cnt;
            tri_tile_row(write_tri, i, tiles_n, BM, BN, lo, cnt);
            if (b < cnt) {
                tm = i;
                tn = lo + b;
                return true;
            }
            b -= cnt;
        }
        return false;
    }
    if (b >= tiles_m * tiles_n) return false;
    // Values, not stores through tm / tn per case: stores to different
    // references in different cases get merged into one store through a
    // selected pointer, which keeps tm and tn in scratch memory.
    const bool by_m = kband == 1 || kband == 2;
    const int q = b / (by_m ? tiles_n : tiles_m), r = b % (by_m ? tiles_n : tiles_m);
    const int m_idx = kband == 1 ? tiles_m - 1 - q : kband == 2 ? q : r; // work grows / shrinks with tm
    const int n_idx = kband == 1 || kband == 2 ? r : kband == 3 ? tiles_n - 1 - q : q; // ... with tn
    tm = m_idx;
    tn = n_idx;
    return true;
}

} // namespace icl
