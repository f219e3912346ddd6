// FP32 MFMA SGEMM fed by LDS DMA: 256 x 256 tiles with 8 waves (4 along m, 2 along n)
// and 128 x 128 tiles with 4 waves (2 x 2).
//
// Same product as sgemm_f32_kernel (reference iclBLAS.h:1648-1685; CPU definition
// tests/blas_funcs/test_gemm.hpp:83-201), for problems without edges (M, N
// multiples of the tile, K of the slab depth BK, 16-byte aligned operand
// columns); problems with edges run sgemm_f32_kernel of the same tile.
//
// Each BK-deep slab (BK = 16, or 32 for the 256 x 256 k32 variant) of op(A) and
// op(B) is copied global -> LDS with global_load_lds_dwordx4, one slab ahead of
// the MFMAs, so no slab passes through VGPRs and no ds_write sits between the
// last MFMA of a slab and the barrier. One DMA instruction writes 64 lanes x 16
// bytes = 1 KB of contiguous LDS, in one of two layouts by the operand's
// contiguous dimension:
//  * rows contiguous (op(A) = A, op(B) = B^T): S[k][rows], one instruction per k
//    row of a 256-row tile, per two k rows of a 128-row tile;
//    a fragment read is one float per lane, 32 consecutive rows per half-wave.
//  * k contiguous (op(A) = A^T, op(B) = B): S[row][BK], rows of BK / 4 16-byte
//    chunks, logical chunk c of row r stored in slot c ^ ((r >> 2) & 3) (BK 16)
//    or c ^ ((r >> 1) & 7) (BK 32); a 16-byte read serves 16 lanes from 256
//    bytes of banks, and over 16 consecutive rows at one logical chunk the slots
//    and rows fall on distinct banks, so the reads are free of bank conflicts.
// A 16-byte read returns four k for one row, so within a slab the MFMAs take k
// in the order k(j, h) = 4 * (2 * (j >> 2) + h) + (j & 3) for MFMA j = 0..BK/2-1
// and lane half h (both operands alike: a common permutation of the summation
// index). The 32-deep slab halves the barriers per k; it gains most where op(B)
// is rows-contiguous (NT, TT: 8192^3 138.6 -> 142.5 and 139.2 -> 141.7 TFLOP/s).
#include "gemm_variants.hpp"

namespace icl {

typedef __attribute__((address_space(1))) void f32dma_global_void;
typedef __attribute__((address_space(3))) void f32dma_lds_void;

template <int BM_, int BN_, int WVM_, int WVN_, int BK_ = 16>
struct F32DmaCfg {
    static constexpr int BM = BM_, BN = BN_, BK = BK_, WVM = WVM_, WVN = WVN_;
    static_assert(BK == 16 || BK == 32, "slab depth");
    static constexpr int NW = WVM * WVN, NT = NW * 64;
    static constexpr int WM = BM / WVM, WN = BN / WVN;
    static constexpr int FM = WM / 32, FN = WN / 32;
    static constexpr int A_BYTES = BM * BK * 4, B_BYTES = BN * BK * 4; // per stage
    static constexpr int STAGE_BYTES = A_BYTES + B_BYTES;
    static constexpr size_t LDS_BYTES = 2 * (size_t)STAGE_BYTES;
    static_assert(A_BYTES % (1024 * NW) == 0 && B_BYTES % (1024 * NW) == 0, "slab copies must divide among the waves");
    static_assert(BM == 128 || BM == 256, "a k row of a rows-contiguous slab fills 1 KB or half of it");
    static_assert(BN == 128 || BN == 256, "a k row of a rows-contiguous slab fills 1 KB or half of it");
};

// Slot of logical 16-byte chunk c of row r of a k-contiguous slab: c ^ swizzle(r)
// (BK / 4 chunks per row; 16 consecutive rows at one logical chunk then use
// distinct banks for the 16-byte reads).
template <int BK>
__device__ __forceinline__ int f32dma_swizzle(int r)
{
    return BK == 16 ? (r >> 2) & 3 : (r >> 1) & 7;
}

// Issue the DMA of one operand's slab (rows row0 .. row0 + ROWS - 1, k0 .. k0 + BK - 1).
template <bool KC, int ROWS, int NW, int BK>
__device__ __forceinline__ void f32dma_issue(const float* X, long ld, int row0, int k0, unsigned char* S, int wave,
                                             int lane)
{
    constexpr int KR = 256 / ROWS; // k rows per 1 KB block (rows-contiguous layout)
    constexpr int RPB = 256 / BK;  // rows per 1 KB block (k-contiguous layout)
#pragma unroll
    for (int i = 0; i < ROWS * BK * 4 / 1024 / NW; ++i) {
        const int blk = wave + NW * i; // 1 KB block of the slab
        const float* g;
        if constexpr (!KC) {
            // block = k rows KR * blk ..; lane covers 4 consecutive rows of one of them
            const int kr = KR * blk + lane / (64 / KR);
            g = X + (long)(k0 + kr) * ld + row0 + 4 * (lane % (64 / KR));
        } else {
            // block = rows RPB * blk ..; lane -> (row, physical slot)
            const int r = RPB * blk + lane / (BK / 4);
            const int c = (lane % (BK / 4)) ^ f32dma_swizzle<BK>(r);
            g = X + (long)(row0 + r) * ld + k0 + 4 * c;
        }
        __builtin_amdgcn_global_load_lds((f32dma_global_void*)g, (f32dma_lds_void*)(S + blk * 1024), 16, 0, 0);
    }
}

// Fragments of k-group q2 (MFMAs j = 4 * q2 .. 4 * q2 + 3) for F 32-row blocks
// starting at row rb: frag[t][f] is row rb + 32 f + (lane & 31) at k = k(4 q2 + t, h).
template <bool KC, int F, int ROWS, int BK>
__device__ __forceinline__ void f32dma_frag(const float* S, int rb, int q2, int lane, float (&frag)[4][F])
{
    const int h = lane >> 5, q = 2 * q2 + h;
#pragma unroll
    for (int f = 0; f < F; ++f) {
        const int r = rb + 32 * f + (lane & 31);
        if constexpr (KC) {
            const f32x4 v = *reinterpret_cast<const f32x4*>(S + r * BK + ((q ^ f32dma_swizzle<BK>(r)) << 2));
#pragma unroll
            for (int t = 0; t < 4; ++t) frag[t][f] = v[t];
        } else {
#pragma unroll
            for (int t = 0; t < 4; ++t) frag[t][f] = S[(4 * q + t) * ROWS + r];
        }
    }
}

// L3: the level-3 instantiation -- tiles of the written triangle only (write_tri),
// the slabs of the tile's k window only (kband); the epilogue always honours
// write_tri. The plain instantiation is compiled without the first two.
template <int BM, int BN, int WVM, int WVN, bool A_KC, bool B_KC, bool L3, int BK = 16>
__global__ void __launch_bounds__(WVM* WVN * 64) sgemm_f32_dma_kernel(GemmArgs g)
{
    using Cfg = F32DmaCfg<BM, BN, WVM, WVN, BK>;
    extern __shared__ __attribute__((aligned(1024))) unsigned char lds_f32dma[];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm0 = (wave % Cfg::WVM) * Cfg::WM;
    const int wn0 = (wave / Cfg::WVM) * Cfg::WN;
    int tm, tn;
    if constexpr (L3) {
        if (!tile_coords_level3<Cfg::BM, Cfg::BN>(g.write_tri, g.kband, g.M / Cfg::BM, g.N / Cfg::BN, tm, tn)) return;
    } else
        tile_coords<8>(g.M / Cfg::BM, g.N / Cfg::BN, tm, tn);
    const int m0 = tm * Cfg::BM, n0 = tn * Cfg::BN;
    int ks0 = 0, ks1 = g.K / Cfg::BK; // slabs [ks0, ks1) of the tile's k window
    if constexpr (L3) {
        int kb = 0, ke = 0;
        if (!tile_window<Cfg::BM, Cfg::BN>(g, m0, n0, kb, ke)) return;
        ks0 = kb / Cfg::BK;
        ks1 = ke > kb ? (ke + Cfg::BK - 1) / Cfg::BK : ks0;
    }

    f32x16 acc[Cfg::FM][Cfg::FN];
#pragma unroll
    for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
        for (int j = 0; j < Cfg::FN; ++j) acc[i][j] = f32x16{};

    auto issue = [&](int ks, int s) {
        unsigned char* stage = lds_f32dma + s * Cfg::STAGE_BYTES;
        f32dma_issue<A_KC, BM, Cfg::NW, BK>(g.A, g.lda, m0, ks * Cfg::BK, stage, wave, lane);
        f32dma_issue<B_KC, BN, Cfg::NW, BK>(g.B, g.ldb, n0, ks * Cfg::BK, stage + Cfg::A_BYTES, wave, lane);
    };

    if (ks0 < ks1) {
        issue(ks0, 0);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        for (int ks = ks0; ks < ks1; ++ks) {
            const int cur = (ks - ks0) & 1;
            if (ks + 1 < ks1) issue(ks + 1, cur ^ 1);
            const float* As = reinterpret_cast<const float*>(lds_f32dma + cur * Cfg::STAGE_BYTES);
            const float* Bs = reinterpret_cast<const float*>(lds_f32dma + cur * Cfg::STAGE_BYTES + Cfg::A_BYTES);
            float af[2][4][Cfg::FM], bf[2][4][Cfg::FN];
            f32dma_frag<A_KC, Cfg::FM, BM, BK>(As, wm0, 0, lane, af[0]);
            f32dma_frag<B_KC, Cfg::FN, BN, BK>(Bs, wn0, 0, lane, bf[0]);
#pragma unroll
            for (int q2 = 0; q2 < BK / 8; ++q2) {
                if (q2 + 1 < BK / 8) {
                    // the next k-group's fragments read while this group's MFMAs issue
                    f32dma_frag<A_KC, Cfg::FM, BM, BK>(As, wm0, q2 + 1, lane, af[(q2 + 1) & 1]);
                    f32dma_frag<B_KC, Cfg::FN, BN, BK>(Bs, wn0, q2 + 1, lane, bf[(q2 + 1) & 1]);
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
                        for (int j = 0; j < Cfg::FN; ++j)
                            acc[i][j] = mfma_32x32x2_f32(bf[q2 & 1][t][j], af[q2 & 1][t][i], acc[i][j]);
            }
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }
    }
#pragma unroll
    for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
        for (int j = 0; j < Cfg::FN; ++j)
            store_real_tile<false, L3>(g, acc[i][j], m0 + wm0 + 32 * i + (lane & 31), n0 + wn0 + 32 * j, lane);
}

template <int BM, int BN, int WVM, int WVN, bool A_KC, bool B_KC, int BK>
hipError_t run_sgemm_f32_dma(const GemmArgs* g, hipStream_t s)
{
    using Cfg = F32DmaCfg<BM, BN, WVM, WVN, BK>;
    if (uses_level3(g))
        return launch_gemm<BM, BN, Cfg::NT, Cfg::LDS_BYTES>(
            sgemm_f32_dma_kernel<BM, BN, WVM, WVN, A_KC, B_KC, true, BK>, g, s);
    return launch_gemm<BM, BN, Cfg::NT, Cfg::LDS_BYTES>(sgemm_f32_dma_kernel<BM, BN, WVM, WVN, A_KC, B_KC, false, BK>,
                                                        g, s);
}

// op(A) = A has rows (m) contiguous; op(B) = B has k contiguous. Problems with
// edges run sgemm_f32_kernel of the same tile and wave layout, and so does op
// N x T at 128 x 128, where both operands are rows-contiguous and that kernel
// measured faster (8192^3: 134.9 against 132.3 TFLOP/s).
template <int BM, int BN, int WVM, int WVN, int BK = 16>
hipError_t sgemm_f32_dma_dispatch(const icl_gemm_args* g, int opa, int opb, int edge, void*, hipStream_t s)
{
    const bool a_kc = opa != 0, b_kc = opb == 0;
    if (edge || (BM == 128 && !a_kc && !b_kc) || g->K % BK != 0)
        return sgemm_dispatch<false, BM, BN, 16, WVM, WVN>(g, opa, opb, edge, nullptr, s);
    if (!a_kc && b_kc) return run_sgemm_f32_dma<BM, BN, WVM, WVN, false, true, BK>(g, s);
    if (!a_kc && !b_kc) return run_sgemm_f32_dma<BM, BN, WVM, WVN, false, false, BK>(g, s);
    if (a_kc && b_kc) return run_sgemm_f32_dma<BM, BN, WVM, WVN, true, true, BK>(g, s);
    return run_sgemm_f32_dma<BM, BN, WVM, WVN, true, false, BK>(g, s);
}

} // namespace icl

static const icl_gemm_variant kTable[] = {
    {"sgemm_f32_dma_256x256w8", ICL_KIND_SGEMM, ICL_MATH_FP32, 256, 256, 16, 512,
     icl::sgemm_f32_dma_dispatch<256, 256, 4, 2>, nullptr, 141.0f, 1, 1},
    {"sgemm_f32_dma_256x256w8k32", ICL_KIND_SGEMM, ICL_MATH_FP32, 256, 256, 32, 512,
     icl::sgemm_f32_dma_dispatch<256, 256, 4, 2, 32>, nullptr, 142.5f, 1, 1},
    {"sgemm_f32_dma_128x128", ICL_KIND_SGEMM, ICL_MATH_FP32, 128, 128, 16, 256,
     icl::sgemm_f32_dma_dispatch<128, 128, 2, 2>, nullptr, 136.5f, 2, 1},
};
ICL_REGISTER_VARIANTS(kTable)
