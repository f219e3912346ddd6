// Level-3 GEMM kernels for gfx950: FP32 MFMA SGEMM, split-BF16 (x3) SGEMM,
// FP32 MFMA CGEMM and split-BF16 CGEMM.
//
// Semantics follow iclblasSgemm / iclblasCgemm (reference include/iclBLAS/iclBLAS.h:1648-1685,
// 1762-1799; CPU definition tests/blas_funcs/test_gemm.hpp:83-201): column-major
// C = alpha * op(A) * op(B) + beta * C with op(A) m x k, op(B) k x n; beta == 0 never reads C.
//
// Structure shared by all four kernels:
//  * One workgroup computes a BM x BN tile of C; WAVES_M x WAVES_N wavefronts of 64
//    lanes each own a (BM/WAVES_M) x (BN/WAVES_N) sub-tile made of 32x32 MFMA blocks.
//  * K is walked in slabs of BK. A slab of op(A) (BM x BK) and op(B) (BN x BK) is
//    fetched with 16-byte global loads along the operand's contiguous dimension into
//    registers, one slab ahead of the MFMA work (double-buffered LDS, one barrier
//    per slab).
//  * The MFMA's first operand is the op(B) fragment and the second the op(A) fragment,
//    so the accumulator's lane index runs along m: C stores (column-major) are
//    128-byte coalesced per half-wave.
//  * EDGE = false kernels assume M % BM == N % BN == K % BK == 0 and 16-byte aligned
//    operand columns; EDGE = true kernels bounds-check every element.
#pragma once
#include "mfma.hpp"
#include "kernel_abi.h"

namespace icl {

using GemmArgs = icl_gemm_args;

enum : int {
    kMaskNone = 0,
    kMaskLower = 1,       // keep m >= n
    kMaskLowerStrict = 2, // keep m > n
    kMaskUpper = 3,       // keep m <= n
    kMaskUpperStrict = 4  // keep m < n
};

// Triangle mask decoded once per tile: keep(m, n) = !on || (lower ? m - n : n - m) >= strict.
struct TriMask {
    bool on, lower;
    int strict;
    __device__ __forceinline__ explicit TriMask(int mask)
        : on(mask != kMaskNone), lower(mask == kMaskLower || mask == kMaskLowerStrict),
          strict((mask == kMaskLowerStrict || mask == kMaskUpperStrict) ? 1 : 0)
    {
    }
    __device__ __forceinline__ bool keep(int m, int n) const { return !on || (lower ? m - n : n - m) >= strict; }
};

// Level-3 tile window. Returns false when the tile [m0, m0 + BM) x [n0, n0 + BN)
// lies wholly outside the triangle of C that g.write_tri selects (the whole
// workgroup then exits before touching memory); otherwise sets [kb, ke), the k
// range in which the triangular operand named by g.kband can be non-zero. kb is
// m0 or n0, a multiple of BK because BM and BN are.
template <int BM, int BN>
__device__ __forceinline__ bool tile_window(const GemmArgs& g, int m0, int n0, int& kb, int& ke)
{
    if ((g.write_tri & 3) == ICL_TRI_LOWER && m0 + BM - 1 < n0) return false;
    if ((g.write_tri & 3) == ICL_TRI_UPPER && m0 > n0 + BN - 1) return false;
    // Both bounds as values (see tile_coords_level3 on stores per case).
    const int kband = g.kband;
    kb = kband == ICL_KBAND_K_FROM_M0 ? min(g.K, m0) : kband == ICL_KBAND_K_FROM_N0 ? min(g.K, n0) : 0;
    ke = kband == ICL_KBAND_K_BELOW_M_END   ? min(g.K, m0 + BM)
         : kband == ICL_KBAND_K_BELOW_N_END ? min(g.K, n0 + BN)
                                            : g.K;
    return true;
}

// Store predicate of g.write_tri (its band flags select tiles, not elements).
__device__ __forceinline__ bool tri_writes(int write_tri, int m, int n)
{
    const int tri = write_tri & 3;
    return tri == ICL_TRI_NONE || (tri == ICL_TRI_LOWER ? m >= n : m <= n);
}

// ---------------------------------------------------------------------------
// Operand slab loaders.
// An operand slab is ROWS x BK where "row" is m for op(A) and n for op(B).
// KCONTIG: element (r, k) at X[(row0 + r) * ld + k0 + k] (k contiguous in memory);
// otherwise at X[(k0 + k) * ld + row0 + r] (rows contiguous).
// ---------------------------------------------------------------------------

// FP32 slab stored to LDS as S[k * LD + r] (rows contiguous).
template <int ROWS, int BK, int NT, bool KCONTIG, bool EDGE>
struct F32Slab {
    static constexpr int UNITS = ROWS * BK / 4;
    static constexpr int PT = (UNITS + NT - 1) / NT;
    static constexpr bool FULL = UNITS % NT == 0;
    // LDS leading dimension: row-contiguous slabs are written with 16-byte stores
    // (need LD % 4 == 0); k-contiguous slabs are transposed with 4-byte stores whose
    // 32-lane pattern is bank-conflict free when LD % 8 == 32 / BK.
    static constexpr int LD = KCONTIG ? ROWS + 32 / BK : ROWS + 4;
    f32x4 v[PT];

    __device__ __forceinline__ void load(const float* __restrict__ X, long ld, int row0, int k0, int rows,
                                         int K, int tid)
    {
#pragma unroll
        for (int p = 0; p < PT; ++p) {
            const int u = tid + p * NT;
            if (!FULL && u >= UNITS) break;
            if constexpr (!KCONTIG) {
                const int r = (u % (ROWS / 4)) * 4;
                const int k = u / (ROWS / 4);
                const float* src = X + (long)(k0 + k) * ld + (row0 + r);
                if constexpr (!EDGE) {
                    v[p] = *reinterpret_cast<const f32x4*>(src);
                } else {
                    const bool kin = (k0 + k) < K;
#pragma unroll
                    for (int i = 0; i < 4; ++i) v[p][i] = (kin && row0 + r + i < rows) ? src[i] : 0.0f;
                }
            } else {
                const int k = (u % (BK / 4)) * 4;
                const int r = u / (BK / 4);
                const float* src = X + (long)(row0 + r) * ld + (k0 + k);
                if constexpr (!EDGE) {
                    v[p] = *reinterpret_cast<const f32x4*>(src);
                } else {
                    const bool rin = (row0 + r) < rows;
#pragma unroll
                    for (int i = 0; i < 4; ++i) v[p][i] = (rin && k0 + k + i < K) ? src[i] : 0.0f;
                }
            }
        }
    }

    __device__ __forceinline__ void store(float* S, int tid) const
    {
#pragma unroll
        for (int p = 0; p < PT; ++p) {
            const int u = tid + p * NT;
            if (!FULL && u >= UNITS) break;
            if constexpr (!KCONTIG) {
                const int r = (u % (ROWS / 4)) * 4;
                const int k = u / (ROWS / 4);
                *reinterpret_cast<f32x4*>(S + k * LD + r) = v[p];
            } else {
                const int k = (u % (BK / 4)) * 4;
                const int r = u / (BK / 4);
#pragma unroll
                for (int i = 0; i < 4; ++i) S[(k + i) * LD + r] = v[p][i];
            }
        }
    }
};

// FP32 slab split into BF16 hi/lo, stored as Sh[r * LD + k], Sl[r * LD + k]
// (k contiguous, so an MFMA fragment of 8 consecutive k is one 16-byte LDS read).
// LD = BK + 8: a 16-byte row pad makes the 8-lane groups of a 16-byte read hit
// distinct banks.
template <int ROWS, int BK, int NT, bool KCONTIG, bool EDGE>
struct SplitSlab {
    static constexpr int LD = BK + 8;
    // k-contiguous: one 16-byte load (4 k) per unit.
    // row-contiguous: a 4 (rows) x 4 (k) block per unit, 4 loads, transposed in registers.
    static constexpr int UNITS = KCONTIG ? ROWS * BK / 4 : ROWS * BK / 16;
    static constexpr int PT = (UNITS + NT - 1) / NT;
    static constexpr bool FULL = UNITS % NT == 0;
    static constexpr int NV = KCONTIG ? 1 : 4;
    f32x4 v[PT][NV];

    __device__ __forceinline__ void load(const float* __restrict__ X, long ld, int row0, int k0, int rows,
                                         int K, int tid)
    {
#pragma unroll
        for (int p = 0; p < PT; ++p) {
            const int u = tid + p * NT;
            if (!FULL && u >= UNITS) break;
            if constexpr (KCONTIG) {
                const int k = (u % (BK / 4)) * 4;
                const int r = u / (BK / 4);
                const float* src = X + (long)(row0 + r) * ld + (k0 + k);
                if constexpr (!EDGE) {
                    v[p][0] = *reinterpret_cast<const f32x4*>(src);
                } else {
                    const bool rin = (row0 + r) < rows;
#pragma unroll
                    for (int i = 0; i < 4; ++i) v[p][0][i] = (rin && k0 + k + i < K) ? src[i] : 0.0f;
                }
            } else {
                const int r = (u % (ROWS / 4)) * 4;
                const int k = (u / (ROWS / 4)) * 4;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float* src = X + (long)(k0 + k + j) * ld + (row0 + r);
                    if constexpr (!EDGE) {
                        v[p][j] = *reinterpret_cast<const f32x4*>(src);
                    } else {
                        const bool kin = (k0 + k + j) < K;
#pragma unroll
                        for (int i = 0; i < 4; ++i) v[p][j][i] = (kin && row0 + r + i < rows) ? src[i] : 0.0f;
                    }
                }
            }
        }
    }

    __device__ __forceinline__ void store(__bf16* Sh, __bf16* Sl, int tid) const
    {
#pragma unroll
        for (int p = 0; p < PT; ++p) {
            const int u = tid + p * NT;
            if (!FULL && u >= UNITS) break;
            if constexpr (KCONTIG) {
                const int k = (u % (BK / 4)) * 4;
                const int r = u / (BK / 4);
                bf16x4 h, l;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    __bf16 hh, ll;
                    split_bf16(v[p][0][i], hh, ll);
                    h[i] = hh;
                    l[i] = ll;
                }
                *reinterpret_cast<bf16x4*>(Sh + r * LD + k) = h;
                *reinterpret_cast<bf16x4*>(Sl + r * LD + k) = l;
            } else {
                const int r = (u % (ROWS / 4)) * 4;
                const int k = (u / (ROWS / 4)) * 4;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    bf16x4 h, l;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        __bf16 hh, ll;
                        split_bf16(v[p][j][i], hh, ll);
                        h[j] = hh;
                        l[j] = ll;
                    }
                    *reinterpret_cast<bf16x4*>(Sh + (r + i) * LD + k) = h;
                    *reinterpret_cast<bf16x4*>(Sl + (r + i) * LD + k) = l;
                }
            }
        }
    }
};

// Complex FP32 slab (interleaved {re, im} in memory) stored to LDS as separate
// real and imaginary planes Sr[k * LD + r], Si[k * LD + r]. CONJ negates the
// imaginary part (op == ICLBLAS_OP_C).
template <int ROWS, int BK, int NT, bool KCONTIG, bool CONJ, bool EDGE>
struct C32Slab {
    static constexpr int UNITS = ROWS * BK / 2; // one 16-byte load = 2 complex per unit
    static constexpr int PT = (UNITS + NT - 1) / NT;
    static constexpr bool FULL = UNITS % NT == 0;
    static constexpr int LD = KCONTIG ? ROWS + 2 : ROWS + 4;
    f32x4 v[PT];

    __device__ __forceinline__ void load(const float* __restrict__ X, long ld, int row0, int k0, int rows,
                                         int K, int tid)
    {
#pragma unroll
        for (int p = 0; p < PT; ++p) {
            const int u = tid + p * NT;
            if (!FULL && u >= UNITS) break;
            if constexpr (!KCONTIG) {
                const int r = (u % (ROWS / 2)) * 2;
                const int k = u / (ROWS / 2);
                const float* src = X + 2 * ((long)(k0 + k) * ld + (row0 + r));
                if constexpr (!EDGE) {
                    v[p] = *reinterpret_cast<const f32x4*>(src);
                } else {
                    const bool kin = (k0 + k) < K;
#pragma unroll
                    for (int i = 0; i < 2; ++i) {
                        const bool in = kin && (row0 + r + i) < rows;
                        v[p][2 * i] = in ? src[2 * i] : 0.0f;
                        v[p][2 * i + 1] = in ? src[2 * i + 1] : 0.0f;
                    }
                }
            } else {
                const int k = (u % (BK / 2)) * 2;
                const int r = u / (BK / 2);
                const float* src = X + 2 * ((long)(row0 + r) * ld + (k0 + k));
                if constexpr (!EDGE) {
                    v[p] = *reinterpret_cast<const f32x4*>(src);
                } else {
                    const bool rin = (row0 + r) < rows;
#pragma unroll
                    for (int i = 0; i < 2; ++i) {
                        const bool in = rin && (k0 + k + i) < K;
                        v[p][2 * i] = in ? src[2 * i] : 0.0f;
                        v[p][2 * i + 1] = in ? src[2 * i + 1] : 0.0f;
                    }
                }
            }
        }
    }

    __device__ __forceinline__ void store(float* Sr, float* Si, int tid) const
    {
        constexpr float s = CONJ ? -1.0f : 1.0f;
#pragma unroll
        for (int p = 0; p < PT; ++p) {
            const int u = tid + p * NT;
            if (!FULL && u >= UNITS) break;
            if constexpr (!KCONTIG) {
                const int r = (u % (ROWS / 2)) * 2;
                const int k = u / (ROWS / 2);
                f32x2 re = {v[p][0], v[p][2]};
                f32x2 im = {s * v[p][1], s * v[p][3]};
                *reinterpret_cast<f32x2*>(Sr + k * LD + r) = re;
                *reinterpret_cast<f32x2*>(Si + k * LD + r) = im;
            } else {
                const int k = (u % (BK / 2)) * 2;
                const int r = u / (BK / 2);
                Sr[k * LD + r] = v[p][0];
                Si[k * LD + r] = s * v[p][1];
                Sr[(k + 1) * LD + r] = v[p][2];
                Si[(k + 1) * LD + r] = s * v[p][3];
            }
        }
    }
};

// Complex slab split into BF16 hi/lo planes for re and im, stored k-contiguous:
// S[plane][r * LD + k], planes {re_hi, re_lo, im_hi, im_lo}.
template <int ROWS, int BK, int NT, bool KCONTIG, bool CONJ, bool EDGE>
struct CSplitSlab {
    static constexpr int LD = BK + 8;
    // k-contiguous: unit = 2 complex along k at one row (1 load).
    // row-contiguous: unit = 2 rows x 4 k (4 loads, each 2 complex along rows).
    static constexpr int UNITS = KCONTIG ? ROWS * BK / 2 : ROWS * BK / 8;
    static constexpr int PT = (UNITS + NT - 1) / NT;
    static constexpr bool FULL = UNITS % NT == 0;
    static constexpr int NV = KCONTIG ? 1 : 4;
    f32x4 v[PT][NV];

    __device__ __forceinline__ void load(const float* __restrict__ X, long ld, int row0, int k0, int rows,
                                         int K, int tid)
    {
#pragma unroll
        for (int p = 0; p < PT; ++p) {
            const int u = tid + p * NT;
            if (!FULL && u >= UNITS) break;
            if constexpr (KCONTIG) {
                const int k = (u % (BK / 2)) * 2;
                const int r = u / (BK / 2);
                const float* src = X + 2 * ((long)(row0 + r) * ld + (k0 + k));
                if constexpr (!EDGE) {
                    v[p][0] = *reinterpret_cast<const f32x4*>(src);
                } else {
                    const bool rin = (row0 + r) < rows;
#pragma unroll
                    for (int i = 0; i < 2; ++i) {
                        const bool in = rin && (k0 + k + i) < K;
                        v[p][0][2 * i] = in ? src[2 * i] : 0.0f;
                        v[p][0][2 * i + 1] = in ? src[2 * i + 1] : 0.0f;
                    }
                }
            } else {
                const int r = (u % (ROWS / 2)) * 2;
                const int k = (u / (ROWS / 2)) * 4;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float* src = X + 2 * ((long)(k0 + k + j) * ld + (row0 + r));
                    if constexpr (!EDGE) {
                        v[p][j] = *reinterpret_cast<const f32x4*>(src);
                    } else {
                        const bool kin = (k0 + k + j) < K;
#pragma unroll
                        for (int i = 0; i < 2; ++i) {
                            const bool in = kin && (row0 + r + i) < rows;
                            v[p][j][2 * i] = in ? src[2 * i] : 0.0f;
                            v[p][j][2 * i + 1] = in ? src[2 * i + 1] : 0.0f;
                        }
                    }
                }
            }
        }
    }

    __device__ __forceinline__ void store(__bf16* S, int plane_stride, int tid) const
    {
        constexpr float s = CONJ ? -1.0f : 1.0f;
        __bf16* Srh = S;
        __bf16* Srl = S + plane_stride;
        __bf16* Sih = S + 2 * plane_stride;
        __bf16* Sil = S + 3 * plane_stride;
#pragma unroll
        for (int p = 0; p < PT; ++p) {
            const int u = tid + p * NT;
            if (!FULL && u >= UNITS) break;
            if constexpr (KCONTIG) {
                const int k = (u % (BK / 2)) * 2;
                const int r = u / (BK / 2);
                __bf16 rh0, rl0, rh1, rl1, ih0, il0, ih1, il1;
                split_bf16(v[p][0][0], rh0, rl0);
                split_bf16(s * v[p][0][1], ih0, il0);
                split_bf16(v[p][0][2], rh1, rl1);
                split_bf16(s * v[p][0][3], ih1, il1);
                typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
                *reinterpret_cast<bf16x2*>(Srh + r * LD + k) = bf16x2{rh0, rh1};
                *reinterpret_cast<bf16x2*>(Srl + r * LD + k) = bf16x2{rl0, rl1};
                *reinterpret_cast<bf16x2*>(Sih + r * LD + k) = bf16x2{ih0, ih1};
                *reinterpret_cast<bf16x2*>(Sil + r * LD + k) = bf16x2{il0, il1};
            } else {
                const int r = (u % (ROWS / 2)) * 2;
                const int k = (u / (ROWS / 2)) * 4;
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    bf16x4 rh, rl, ih, il;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        __bf16 a, b;
                        split_bf16(v[p][j][2 * i], a, b);
                        rh[j] = a;
                        rl[j] = b;
                        split_bf16(s * v[p][j][2 * i + 1], a, b);
                        ih[j] = a;
                        il[j] = b;
                    }
                    *reinterpret_cast<bf16x4*>(Srh + (r + i) * LD + k) = rh;
                    *reinterpret_cast<bf16x4*>(Srl + (r + i) * LD + k) = rl;
                    *reinterpret_cast<bf16x4*>(Sih + (r + i) * LD + k) = ih;
                    *reinterpret_cast<bf16x4*>(Sil + (r + i) * LD + k) = il;
                }
            }
        }
    }
};

// ---------------------------------------------------------------------------
// Epilogue helpers
// ---------------------------------------------------------------------------
// TRI_CHECK: honour g.write_tri (the FP32 kernels always test it; the uniform
// branch also keeps the compiler from hoisting every store address of the tile
// ahead of the stores).
// Register r of a lane holds column n = ncol0 + 4 (lane / 32) + dr(r) with the
// compile-time dr(r) = 8 (r / 4) + r % 4, so every store address is the lane's
// column-0 pointer plus the wave-uniform dr(r) * ldc: one 64-bit pointer per lane
// is live across the tile instead of one address computation per register.
__device__ __forceinline__ constexpr int mfma32_reg_row(int r) { return 8 * (r >> 2) + (r & 3); }

template <bool EDGE, bool TRI_CHECK = false>
__device__ __forceinline__ void store_real_tile(const GemmArgs& g, const f32x16& acc, int mrow, int ncol0, int lane)
{
    // mrow: m index of this lane; ncol0: n index of register 0's 32x32 block origin.
    const TriMask tri(g.mask);
    const int nl = ncol0 + 4 * (lane >> 5);
    float* const c0 = g.C + (long)nl * g.ldc + mrow;
    // beta == 0 (C not read) and beta != 0 as two straight-line loops under one
    // uniform branch, rather than a branch around each load of C.
    if (g.beta_re == 0.0f) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int dr = mfma32_reg_row(r), n = nl + dr;
            if (EDGE && (mrow >= g.M || n >= g.N)) continue;
            if constexpr (TRI_CHECK)
                if (!tri_writes(g.write_tri, mrow, n)) continue;
            c0[(long)dr * g.ldc] = tri.keep(mrow, n) ? g.alpha_re * acc[r] : 0.0f;
        }
    } else {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int dr = mfma32_reg_row(r), n = nl + dr;
            if (EDGE && (mrow >= g.M || n >= g.N)) continue;
            if constexpr (TRI_CHECK)
                if (!tri_writes(g.write_tri, mrow, n)) continue;
            float* c = c0 + (long)dr * g.ldc;
            const float v = fmaf(g.beta_re, *c, g.alpha_re * acc[r]);
            *c = tri.keep(mrow, n) ? v : 0.0f;
        }
    }
}

// c = alpha * acc + beta * c with the complex product of the reference's
// complex.h.cl:32-34 (cmul(a, b) = (fma(a.re, b.re, -a.im*b.im), fma(a.re, b.im, a.im*b.re))).
template <bool EDGE, bool L3 = false>
__device__ __forceinline__ void store_complex_tile(const GemmArgs& g, const f32x16& accr, const f32x16& acci,
                                                   int mrow, int ncol0, int lane)
{
    const bool beta_zero = (g.beta_re == 0.0f && g.beta_im == 0.0f);
    const float conj_sign = (L3 && (g.cflags & ICL_CFLAG_CONJ_ACC)) ? -1.0f : 1.0f;
    const bool real_diag = L3 && (g.cflags & ICL_CFLAG_REAL_DIAG) != 0;
    const int nl = ncol0 + 4 * (lane >> 5);
    float* const c0 = g.C + 2 * ((long)nl * g.ldc + mrow);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int dr = mfma32_reg_row(r), n = nl + dr;
        if (EDGE && (mrow >= g.M || n >= g.N)) continue;
        if constexpr (L3)
            if (!tri_writes(g.write_tri, mrow, n)) continue;
        float* c = c0 + 2 * (long)dr * g.ldc;
        const float xr = accr[r], xi = conj_sign * acci[r];
        float vr = fmaf(g.alpha_re, xr, -g.alpha_im * xi);
        float vi = fmaf(g.alpha_re, xi, g.alpha_im * xr);
        if (!beta_zero) {
            const f32x2 old = *reinterpret_cast<const f32x2*>(c);
            vr += fmaf(g.beta_re, old[0], -g.beta_im * old[1]);
            vi += fmaf(g.beta_re, old[1], g.beta_im * old[0]);
        }
        if (real_diag && mrow == n) vi = 0.0f;
        *reinterpret_cast<f32x2*>(c) = f32x2{vr, vi};
    }
}

// ---------------------------------------------------------------------------
// FP32 MFMA SGEMM
// ---------------------------------------------------------------------------
template <int TA, int TB, int BM, int BN, int BK, int WAVES_M, int WAVES_N, bool EDGE>
struct SgemmF32 {
    static constexpr int NT = WAVES_M * WAVES_N * 64;
    static constexpr int WM = BM / WAVES_M, WN = BN / WAVES_N;
    static constexpr int FM = WM / 32, FN = WN / 32;
    using SlabA = F32Slab<BM, BK, NT, TA != 0, EDGE>;
    using SlabB = F32Slab<BN, BK, NT, TB == 0, EDGE>;
    static constexpr int A_WORDS = BK * SlabA::LD;
    static constexpr int B_WORDS = BK * SlabB::LD;
    static constexpr size_t LDS_BYTES = 2 * (A_WORDS + B_WORDS) * sizeof(float);
    static_assert(WM % 32 == 0 && WN % 32 == 0 && BK % 4 == 0, "bad tile");
};

template <int TA, int TB, int BM, int BN, int BK, int WAVES_M, int WAVES_N, bool EDGE, bool L3 = false>
__global__ void __launch_bounds__(WAVES_M* WAVES_N * 64)
sgemm_f32_kernel(GemmArgs g)
{
    using Cfg = SgemmF32<TA, TB, BM, BN, BK, WAVES_M, WAVES_N, EDGE>;
    extern __shared__ __attribute__((aligned(16))) float lds[];
    auto As = [&](int s) { return lds + s * Cfg::A_WORDS; };
    auto Bs = [&](int s) { return lds + 2 * Cfg::A_WORDS + s * Cfg::B_WORDS; };

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave % WAVES_M) * Cfg::WM;
    const int wn0 = (wave / WAVES_M) * Cfg::WN;
    int tm, tn;
    if constexpr (L3) {
        if (!tile_coords_level3<BM, BN>(g.write_tri, g.kband, (g.M + BM - 1) / BM, (g.N + BN - 1) / BN, tm, tn))
            return;
    } else
        tile_coords<8>((g.M + BM - 1) / BM, (g.N + BN - 1) / BN, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    int kb = 0, ke = g.K;
    if constexpr (L3)
        if (!tile_window<BM, BN>(g, m0, n0, kb, ke)) return;
    // The k window starts the operands at column / row kb; the loop below then
    // runs over a K of ke - kb exactly as for a plain GEMM.
    const float* const Aw = g.A + ((TA != 0) ? (long)kb : (long)kb * g.lda);
    const float* const Bw = g.B + ((TB == 0) ? (long)kb : (long)kb * g.ldb);
    const int Kw = ke - kb;

    typename Cfg::SlabA sa;
    typename Cfg::SlabB sb;
    f32x16 acc[Cfg::FM][Cfg::FN];
#pragma unroll
    for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
        for (int j = 0; j < Cfg::FN; ++j) acc[i][j] = f32x16{};

    const int KT = (Kw + BK - 1) / BK;
    if (KT > 0) {
        sa.load(Aw, g.lda, m0, 0, g.M, Kw, tid);
        sb.load(Bw, g.ldb, n0, 0, g.N, Kw, tid);
        sa.store(As(0), tid);
        sb.store(Bs(0), tid);
        __syncthreads();
        for (int kt = 0; kt < KT; ++kt) {
            const int cur = kt & 1;
            if (kt + 1 < KT) {
                sa.load(Aw, g.lda, m0, (kt + 1) * BK, g.M, Kw, tid);
                sb.load(Bw, g.ldb, n0, (kt + 1) * BK, g.N, Kw, tid);
            }
            const float* as = As(cur);
            const float* bs = Bs(cur);
            // Fragments of step kk + 2 are read into the second register set
            // before step kk's products are issued, so each LDS read latency
            // hides behind FM * FN MFMAs instead of stalling the wave.
            float a[2][Cfg::FM], b[2][Cfg::FN];
            auto fetch = [&](int kk, float(&fa)[Cfg::FM], float(&fb)[Cfg::FN]) {
                const int kr = kk + (lane >> 5);
#pragma unroll
                for (int i = 0; i < Cfg::FM; ++i) fa[i] = as[kr * Cfg::SlabA::LD + wm0 + 32 * i + (lane & 31)];
#pragma unroll
                for (int j = 0; j < Cfg::FN; ++j) fb[j] = bs[kr * Cfg::SlabB::LD + wn0 + 32 * j + (lane & 31)];
            };
            fetch(0, a[0], b[0]);
#pragma unroll
            for (int kk = 0; kk < BK; kk += 2) {
                const int s = (kk >> 1) & 1;
                if (kk + 2 < BK) fetch(kk + 2, a[s ^ 1], b[s ^ 1]);
                // Keeps the scheduler from sinking the reads back below the
                // products (it would, to save the second register set).
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
                    for (int j = 0; j < Cfg::FN; ++j) acc[i][j] = mfma_32x32x2_f32(b[s][j], a[s][i], acc[i][j]);
            }
            if (kt + 1 < KT) {
                sa.store(As(cur ^ 1), tid);
                sb.store(Bs(cur ^ 1), tid);
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
        for (int j = 0; j < Cfg::FN; ++j)
            store_real_tile<EDGE, true>(g, acc[i][j], m0 + wm0 + 32 * i + (lane & 31), n0 + wn0 + 32 * j, lane);
}

// ---------------------------------------------------------------------------
// Split-BF16 x3 SGEMM: acc += Bh*Ah + Bl*Ah + Bh*Al on v_mfma_f32_32x32x16_bf16.
// ---------------------------------------------------------------------------
template <int TA, int TB, int BM, int BN, int BK, int WAVES_M, int WAVES_N, bool EDGE>
struct SgemmBf16x3 {
    static constexpr int NT = WAVES_M * WAVES_N * 64;
    static constexpr int WM = BM / WAVES_M, WN = BN / WAVES_N;
    static constexpr int FM = WM / 32, FN = WN / 32;
    using SlabA = SplitSlab<BM, BK, NT, TA != 0, EDGE>;
    using SlabB = SplitSlab<BN, BK, NT, TB == 0, EDGE>;
    static constexpr int A_ELEMS = BM * SlabA::LD; // per hi or lo plane
    static constexpr int B_ELEMS = BN * SlabB::LD;
    static constexpr int STAGE_ELEMS = 2 * (A_ELEMS + B_ELEMS);
    static constexpr size_t LDS_BYTES = 2 * STAGE_ELEMS * sizeof(__bf16);
    static_assert(WM % 32 == 0 && WN % 32 == 0 && BK % 16 == 0, "bad tile");
};

template <int TA, int TB, int BM, int BN, int BK, int WAVES_M, int WAVES_N, bool EDGE>
__global__ void __launch_bounds__(WAVES_M* WAVES_N * 64)
sgemm_bf16x3_kernel(GemmArgs g)
{
    using Cfg = SgemmBf16x3<TA, TB, BM, BN, BK, WAVES_M, WAVES_N, EDGE>;
    extern __shared__ __attribute__((aligned(16))) __bf16 ldsb[];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave % WAVES_M) * Cfg::WM;
    const int wn0 = (wave / WAVES_M) * Cfg::WN;
    int tm, tn;
    tile_coords<8>((g.M + BM - 1) / BM, (g.N + BN - 1) / BN, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;

    typename Cfg::SlabA sa;
    typename Cfg::SlabB sb;
    f32x16 acc[Cfg::FM][Cfg::FN];
#pragma unroll
    for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
        for (int j = 0; j < Cfg::FN; ++j) acc[i][j] = f32x16{};

    auto Ah = [&](int s) { return ldsb + s * Cfg::STAGE_ELEMS; };
    auto Al = [&](int s) { return ldsb + s * Cfg::STAGE_ELEMS + Cfg::A_ELEMS; };
    auto Bh = [&](int s) { return ldsb + s * Cfg::STAGE_ELEMS + 2 * Cfg::A_ELEMS; };
    auto Bl = [&](int s) { return ldsb + s * Cfg::STAGE_ELEMS + 2 * Cfg::A_ELEMS + Cfg::B_ELEMS; };

    const int KT = (g.K + BK - 1) / BK;
    if (KT > 0) {
        sa.load(g.A, g.lda, m0, 0, g.M, g.K, tid);
        sb.load(g.B, g.ldb, n0, 0, g.N, g.K, tid);
        sa.store(Ah(0), Al(0), tid);
        sb.store(Bh(0), Bl(0), tid);
        __syncthreads();
        for (int kt = 0; kt < KT; ++kt) {
            const int cur = kt & 1;
            if (kt + 1 < KT) {
                sa.load(g.A, g.lda, m0, (kt + 1) * BK, g.M, g.K, tid);
                sb.load(g.B, g.ldb, n0, (kt + 1) * BK, g.N, g.K, tid);
            }
            const __bf16* ah_s = Ah(cur);
            const __bf16* al_s = Al(cur);
            const __bf16* bh_s = Bh(cur);
            const __bf16* bl_s = Bl(cur);
#pragma unroll
            for (int kk = 0; kk < BK; kk += 16) {
                const int kc = kk + 8 * (lane >> 5);
                bf16x8 ah[Cfg::FM], al[Cfg::FM], bh[Cfg::FN], bl[Cfg::FN];
#pragma unroll
                for (int i = 0; i < Cfg::FM; ++i) {
                    const int o = (wm0 + 32 * i + (lane & 31)) * Cfg::SlabA::LD + kc;
                    ah[i] = *reinterpret_cast<const bf16x8*>(ah_s + o);
                    al[i] = *reinterpret_cast<const bf16x8*>(al_s + o);
                }
#pragma unroll
                for (int j = 0; j < Cfg::FN; ++j) {
                    const int o = (wn0 + 32 * j + (lane & 31)) * Cfg::SlabB::LD + kc;
                    bh[j] = *reinterpret_cast<const bf16x8*>(bh_s + o);
                    bl[j] = *reinterpret_cast<const bf16x8*>(bl_s + o);
                }
#pragma unroll
                for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
                    for (int j = 0; j < Cfg::FN; ++j) acc[i][j] = mfma_32x32x16_bf16(bh[j], ah[i], acc[i][j]);
#pragma unroll
                for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
                    for (int j = 0; j < Cfg::FN; ++j) acc[i][j] = mfma_32x32x16_bf16(bl[j], ah[i], acc[i][j]);
#pragma unroll
                for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
                    for (int j = 0; j < Cfg::FN; ++j) acc[i][j] = mfma_32x32x16_bf16(bh[j], al[i], acc[i][j]);
            }
            if (kt + 1 < KT) {
                sa.store(Ah(cur ^ 1), Al(cur ^ 1), tid);
                sb.store(Bh(cur ^ 1), Bl(cur ^ 1), tid);
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
        for (int j = 0; j < Cfg::FN; ++j)
            store_real_tile<EDGE>(g, acc[i][j], m0 + wm0 + 32 * i + (lane & 31), n0 + wn0 + 32 * j, lane);
}

// ---------------------------------------------------------------------------
// FP32 MFMA CGEMM: four real MFMAs per complex fragment product
//   Cr += Br*Ar - Bi*Ai,  Ci += Bi*Ar + Br*Ai.
// OA/OB: 0 = N, 1 = T, 2 = C (conjugate transpose).
// ---------------------------------------------------------------------------
template <int OA, int OB, int BM, int BN, int BK, int WAVES_M, int WAVES_N, bool EDGE>
struct CgemmF32 {
    static constexpr int NT = WAVES_M * WAVES_N * 64;
    static constexpr int WM = BM / WAVES_M, WN = BN / WAVES_N;
    static constexpr int FM = WM / 32, FN = WN / 32;
    using SlabA = C32Slab<BM, BK, NT, OA != 0, OA == 2, EDGE>;
    using SlabB = C32Slab<BN, BK, NT, OB == 0, OB == 2, EDGE>;
    static constexpr int A_WORDS = BK * SlabA::LD; // per plane
    static constexpr int B_WORDS = BK * SlabB::LD;
    static constexpr int STAGE_WORDS = 2 * (A_WORDS + B_WORDS);
    static constexpr size_t LDS_BYTES = 2 * STAGE_WORDS * sizeof(float);
};

template <int OA, int OB, int BM, int BN, int BK, int WAVES_M, int WAVES_N, bool EDGE, bool L3 = false>
__global__ void __launch_bounds__(WAVES_M* WAVES_N * 64, WAVES_M* WAVES_N <= 4 ? 2 : 1)
cgemm_f32_kernel(GemmArgs g)
{
    using Cfg = CgemmF32<OA, OB, BM, BN, BK, WAVES_M, WAVES_N, EDGE>;
    extern __shared__ __attribute__((aligned(16))) float lds[];
    auto Ar = [&](int s) { return lds + s * Cfg::STAGE_WORDS; };
    auto Ai = [&](int s) { return lds + s * Cfg::STAGE_WORDS + Cfg::A_WORDS; };
    auto Br = [&](int s) { return lds + s * Cfg::STAGE_WORDS + 2 * Cfg::A_WORDS; };
    auto Bi = [&](int s) { return lds + s * Cfg::STAGE_WORDS + 2 * Cfg::A_WORDS + Cfg::B_WORDS; };

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave % WAVES_M) * Cfg::WM;
    const int wn0 = (wave / WAVES_M) * Cfg::WN;
    int tm, tn;
    if constexpr (L3) {
        if (!tile_coords_level3<BM, BN>(g.write_tri, g.kband, (g.M + BM - 1) / BM, (g.N + BN - 1) / BN, tm, tn))
            return;
    } else
        tile_coords<8>((g.M + BM - 1) / BM, (g.N + BN - 1) / BN, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    int kb = 0, ke = g.K;
    if constexpr (L3)
        if (!tile_window<BM, BN>(g, m0, n0, kb, ke)) return;
    // The k window starts the operands at column / row kb; the loop below then
    // runs over a K of ke - kb exactly as for a plain GEMM.
    const float* const Aw = g.A + 2 * ((OA != 0) ? (long)kb : (long)kb * g.lda);
    const float* const Bw = g.B + 2 * ((OB == 0) ? (long)kb : (long)kb * g.ldb);
    const int Kw = ke - kb;

    typename Cfg::SlabA sa;
    typename Cfg::SlabB sb;
    f32x16 accr[Cfg::FM][Cfg::FN], acci[Cfg::FM][Cfg::FN];
#pragma unroll
    for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
        for (int j = 0; j < Cfg::FN; ++j) {
            accr[i][j] = f32x16{};
            acci[i][j] = f32x16{};
        }

    const int KT = (Kw + BK - 1) / BK;
    if (KT > 0) {
        sa.load(Aw, g.lda, m0, 0, g.M, Kw, tid);
        sb.load(Bw, g.ldb, n0, 0, g.N, Kw, tid);
        sa.store(Ar(0), Ai(0), tid);
        sb.store(Br(0), Bi(0), tid);
        __syncthreads();
        for (int kt = 0; kt < KT; ++kt) {
            const int cur = kt & 1;
            if (kt + 1 < KT) {
                sa.load(Aw, g.lda, m0, (kt + 1) * BK, g.M, Kw, tid);
                sb.load(Bw, g.ldb, n0, (kt + 1) * BK, g.N, Kw, tid);
            }
            const float* ar_s = Ar(cur);
            const float* ai_s = Ai(cur);
            const float* br_s = Br(cur);
            const float* bi_s = Bi(cur);
            // Fragments of step kk + 2 are read into the second register set
            // before step kk's products are issued (as in the real kernel).
            float ar[2][Cfg::FM], ai[2][Cfg::FM], br[2][Cfg::FN], bi[2][Cfg::FN];
            auto fetch = [&](int kk, int s) {
                const int kr = kk + (lane >> 5);
#pragma unroll
                for (int i = 0; i < Cfg::FM; ++i) {
                    const int o = kr * Cfg::SlabA::LD + wm0 + 32 * i + (lane & 31);
                    ar[s][i] = ar_s[o];
                    ai[s][i] = ai_s[o];
                }
#pragma unroll
                for (int j = 0; j < Cfg::FN; ++j) {
                    const int o = kr * Cfg::SlabB::LD + wn0 + 32 * j + (lane & 31);
                    br[s][j] = br_s[o];
                    bi[s][j] = bi_s[o];
                }
            };
            fetch(0, 0);
#pragma unroll
            for (int kk = 0; kk < BK; kk += 2) {
                const int s = (kk >> 1) & 1;
                if (kk + 2 < BK) fetch(kk + 2, s ^ 1);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
                    for (int j = 0; j < Cfg::FN; ++j) {
                        accr[i][j] = mfma_32x32x2_f32(br[s][j], ar[s][i], accr[i][j]);
                        acci[i][j] = mfma_32x32x2_f32(bi[s][j], ar[s][i], acci[i][j]);
                    }
#pragma unroll
                for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
                    for (int j = 0; j < Cfg::FN; ++j) {
                        accr[i][j] = mfma_32x32x2_f32(-bi[s][j], ai[s][i], accr[i][j]);
                        acci[i][j] = mfma_32x32x2_f32(br[s][j], ai[s][i], acci[i][j]);
                    }
            }
            if (kt + 1 < KT) {
                sa.store(Ar(cur ^ 1), Ai(cur ^ 1), tid);
                sb.store(Br(cur ^ 1), Bi(cur ^ 1), tid);
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
        for (int j = 0; j < Cfg::FN; ++j)
            store_complex_tile<EDGE, L3>(g, accr[i][j], acci[i][j], m0 + wm0 + 32 * i + (lane & 31),
                                     n0 + wn0 + 32 * j, lane);
}

// ---------------------------------------------------------------------------
// Split-BF16 x3 CGEMM: each of the four real products of the complex MAC is
// formed as hi*hi + lo*hi + hi*lo (12 BF16 MFMAs per complex fragment product).
// ---------------------------------------------------------------------------
template <int OA, int OB, int BM, int BN, int BK, int WAVES_M, int WAVES_N, bool EDGE>
struct CgemmBf16x3 {
    static constexpr int NT = WAVES_M * WAVES_N * 64;
    static constexpr int WM = BM / WAVES_M, WN = BN / WAVES_N;
    static constexpr int FM = WM / 32, FN = WN / 32;
    using SlabA = CSplitSlab<BM, BK, NT, OA != 0, OA == 2, EDGE>;
    using SlabB = CSplitSlab<BN, BK, NT, OB == 0, OB == 2, EDGE>;
    static constexpr int A_PLANE = BM * SlabA::LD;
    static constexpr int B_PLANE = BN * SlabB::LD;
    static constexpr int STAGE_ELEMS = 4 * (A_PLANE + B_PLANE);
    static constexpr size_t LDS_BYTES = 2 * STAGE_ELEMS * sizeof(__bf16);
    static_assert(WM % 32 == 0 && WN % 32 == 0 && BK % 16 == 0, "bad tile");
};

template <int OA, int OB, int BM, int BN, int BK, int WAVES_M, int WAVES_N, bool EDGE>
__global__ void __launch_bounds__(WAVES_M* WAVES_N * 64)
cgemm_bf16x3_kernel(GemmArgs g)
{
    using Cfg = CgemmBf16x3<OA, OB, BM, BN, BK, WAVES_M, WAVES_N, EDGE>;
    extern __shared__ __attribute__((aligned(16))) __bf16 ldsb[];
    auto Aplane = [&](int s, int p) { return ldsb + s * Cfg::STAGE_ELEMS + p * Cfg::A_PLANE; };
    auto Bplane = [&](int s, int p) { return ldsb + s * Cfg::STAGE_ELEMS + 4 * Cfg::A_PLANE + p * Cfg::B_PLANE; };

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave % WAVES_M) * Cfg::WM;
    const int wn0 = (wave / WAVES_M) * Cfg::WN;
    int tm, tn;
    tile_coords<8>((g.M + BM - 1) / BM, (g.N + BN - 1) / BN, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;

    typename Cfg::SlabA sa;
    typename Cfg::SlabB sb;
    f32x16 accr[Cfg::FM][Cfg::FN], acci[Cfg::FM][Cfg::FN];
#pragma unroll
    for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
        for (int j = 0; j < Cfg::FN; ++j) {
            accr[i][j] = f32x16{};
            acci[i][j] = f32x16{};
        }

    const int KT = (g.K + BK - 1) / BK;
    if (KT > 0) {
        sa.load(g.A, g.lda, m0, 0, g.M, g.K, tid);
        sb.load(g.B, g.ldb, n0, 0, g.N, g.K, tid);
        sa.store(Aplane(0, 0), Cfg::A_PLANE, tid);
        sb.store(Bplane(0, 0), Cfg::B_PLANE, tid);
        __syncthreads();
        for (int kt = 0; kt < KT; ++kt) {
            const int cur = kt & 1;
            if (kt + 1 < KT) {
                sa.load(g.A, g.lda, m0, (kt + 1) * BK, g.M, g.K, tid);
                sb.load(g.B, g.ldb, n0, (kt + 1) * BK, g.N, g.K, tid);
            }
#pragma unroll
            for (int kk = 0; kk < BK; kk += 16) {
                const int kc = kk + 8 * (lane >> 5);
                // planes: 0 re_hi, 1 re_lo, 2 im_hi, 3 im_lo
                bf16x8 a[4][Cfg::FM], b[4][Cfg::FN];
#pragma unroll
                for (int p = 0; p < 4; ++p) {
#pragma unroll
                    for (int i = 0; i < Cfg::FM; ++i)
                        a[p][i] = *reinterpret_cast<const bf16x8*>(
                            Aplane(cur, p) + (wm0 + 32 * i + (lane & 31)) * Cfg::SlabA::LD + kc);
#pragma unroll
                    for (int j = 0; j < Cfg::FN; ++j)
                        b[p][j] = *reinterpret_cast<const bf16x8*>(
                            Bplane(cur, p) + (wn0 + 32 * j + (lane & 31)) * Cfg::SlabB::LD + kc);
                }
#pragma unroll
                for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
                    for (int j = 0; j < Cfg::FN; ++j) {
                        const bf16x8 nbih = bf16x8_neg(b[2][j]);
                        const bf16x8 nbil = bf16x8_neg(b[3][j]);
                        // Cr += Br*Ar
                        accr[i][j] = mfma_32x32x16_bf16(b[0][j], a[0][i], accr[i][j]);
                        accr[i][j] = mfma_32x32x16_bf16(b[1][j], a[0][i], accr[i][j]);
                        accr[i][j] = mfma_32x32x16_bf16(b[0][j], a[1][i], accr[i][j]);
                        // Ci += Bi*Ar
                        acci[i][j] = mfma_32x32x16_bf16(b[2][j], a[0][i], acci[i][j]);
                        acci[i][j] = mfma_32x32x16_bf16(b[3][j], a[0][i], acci[i][j]);
                        acci[i][j] = mfma_32x32x16_bf16(b[2][j], a[1][i], acci[i][j]);
                        // Cr -= Bi*Ai
                        accr[i][j] = mfma_32x32x16_bf16(nbih, a[2][i], accr[i][j]);
                        accr[i][j] = mfma_32x32x16_bf16(nbil, a[2][i], accr[i][j]);
                        accr[i][j] = mfma_32x32x16_bf16(nbih, a[3][i], accr[i][j]);
                        // Ci += Br*Ai
                        acci[i][j] = mfma_32x32x16_bf16(b[0][j], a[2][i], acci[i][j]);
                        acci[i][j] = mfma_32x32x16_bf16(b[1][j], a[2][i], acci[i][j]);
                        acci[i][j] = mfma_32x32x16_bf16(b[0][j], a[3][i], acci[i][j]);
                    }
            }
            if (kt + 1 < KT) {
                sa.store(Aplane(cur ^ 1, 0), Cfg::A_PLANE, tid);
                sb.store(Bplane(cur ^ 1, 0), Cfg::B_PLANE, tid);
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
        for (int j = 0; j < Cfg::FN; ++j)
            store_complex_tile<EDGE>(g, accr[i][j], acci[i][j], m0 + wm0 + 32 * i + (lane & 31),
                                     n0 + wn0 + 32 * j, lane);
}

} // namespace icl
