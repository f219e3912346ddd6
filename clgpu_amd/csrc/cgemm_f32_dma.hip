// FP32 MFMA CGEMM fed by LDS DMA, 128 x 128 complex tiles, 4 waves (2 x 2, 64 x 64
// complex each), two workgroups per CU.
//
// Same product as cgemm_f32_kernel (reference iclBLAS.h:1762-1799; CPU definition
// tests/blas_funcs/test_gemm.hpp:83-201; complex MAC of complex.h.cl:32-34), for
// problems without edges (M, N multiples of 128, K of 8, 16-byte aligned operand
// columns); problems with edges run cgemm_f32_kernel of the same tile.
//
// Data are interleaved (re, im) pairs. Each 8-deep slab of op(A) and op(B) is
// copied global -> LDS with global_load_lds_dwordx4, one slab ahead of the MFMAs.
// One DMA instruction writes 64 lanes x 16 bytes = 1 KB = 128 complex values, in
// one of two layouts by the operand's contiguous dimension:
//  * rows contiguous (op(A) = A; op(B) = B^T or B^H): S[k][128 complex], one
//    instruction per k; a fragment read is one 8-byte (re, im) pair per lane.
//  * k contiguous (op(A) = A^T or A^H; op(B) = B): S[row][8 complex], 64-byte rows
//    of four 16-byte chunks (two k each), logical chunk c of row r in slot
//    c ^ ((r >> 2) & 3), conflict-free for 16-byte reads as in the real kernel.
// A 16-byte read returns two k for one row, so within a slab the MFMAs take k in
// the order k(j, h) = 2 * (2 * (j >> 1) + h) + (j & 1) for MFMA step j = 0..3 and
// lane half h, for both operands alike. Conjugation (op C) negates the imaginary
// part of the fragment as it is read.
#include "gemm_variants.hpp"

namespace icl {

typedef __attribute__((address_space(1))) void c32dma_global_void;
typedef __attribute__((address_space(3))) void c32dma_lds_void;

struct C32DmaCfg {
    static constexpr int BM = 128, BN = 128, BK = 8, WVM = 2, WVN = 2;
    static constexpr int NW = WVM * WVN, NT = NW * 64;
    static constexpr int WM = BM / WVM, WN = BN / WVN;
    static constexpr int FM = WM / 32, FN = WN / 32;
    static constexpr int OPER_BYTES = 128 * BK * 8; // 8 KB per operand per stage
    static constexpr int STAGE_BYTES = 2 * OPER_BYTES;
    static constexpr size_t LDS_BYTES = 2 * (size_t)STAGE_BYTES;
    static constexpr int INSTR = OPER_BYTES / 1024 / NW; // DMA instructions per wave per operand per slab
    static_assert(INSTR * NW * 1024 == OPER_BYTES, "slab copies must divide among the waves");
};

__device__ __forceinline__ int c32dma_swizzle(int r) { return (r >> 2) & 3; }

// Issue the DMA of one operand's slab (rows row0 .. row0 + 127, k0 .. k0 + 7);
// X and ld in complex elements.
template <bool KC>
__device__ __forceinline__ void c32dma_issue(const float* X, long ld, int row0, int k0, unsigned char* S, int wave,
                                             int lane)
{
#pragma unroll
    for (int i = 0; i < C32DmaCfg::INSTR; ++i) {
        const int blk = wave + C32DmaCfg::NW * i; // 1 KB block of the slab, 0..7
        long e;                                   // complex index of the lane's first element
        if constexpr (!KC) {
            e = (long)(k0 + blk) * ld + row0 + 2 * lane; // block = k row
        } else {
            const int r = 16 * blk + (lane >> 2); // block = rows 16 * blk .. + 15
            const int c = (lane & 3) ^ c32dma_swizzle(r);
            e = (long)(row0 + r) * ld + k0 + 2 * c;
        }
        __builtin_amdgcn_global_load_lds((c32dma_global_void*)(X + 2 * e), (c32dma_lds_void*)(S + blk * 1024), 16, 0,
                                         0);
    }
}

// Fragments of k-group q2 (steps j = 2 q2, 2 q2 + 1) for F 32-row blocks starting at
// row rb: (re[t][f], im[t][f]) is row rb + 32 f + (lane & 31) at k = k(2 q2 + t, h).
template <bool KC, bool CONJ, int F>
__device__ __forceinline__ void c32dma_frag(const float* S, int rb, int q2, int lane, float (&re)[2][F],
                                            float (&im)[2][F])
{
    const int h = lane >> 5, q = 2 * q2 + h;
    const float s = CONJ ? -1.0f : 1.0f;
#pragma unroll
    for (int f = 0; f < F; ++f) {
        const int r = rb + 32 * f + (lane & 31);
        if constexpr (KC) {
            const f32x4 v = *reinterpret_cast<const f32x4*>(S + r * 16 + ((q ^ c32dma_swizzle(r)) << 2));
            re[0][f] = v[0];
            im[0][f] = s * v[1];
            re[1][f] = v[2];
            im[1][f] = s * v[3];
        } else {
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const f32x2 v = *reinterpret_cast<const f32x2*>(S + (2 * q + t) * 256 + 2 * r);
                re[t][f] = v[0];
                im[t][f] = s * v[1];
            }
        }
    }
}

// L3: the level-3 instantiation -- tiles of the written triangle only (write_tri),
// the slabs of the tile's k window only (kband), and the level-3 epilogue
// (triangle stores, conjugated accumulator, real diagonal).
template <bool A_KC, bool CA, bool B_KC, bool CB, bool L3>
__global__ void __launch_bounds__(C32DmaCfg::NT, 2) cgemm_f32_dma_kernel(GemmArgs g)
{
    using Cfg = C32DmaCfg;
    extern __shared__ __attribute__((aligned(1024))) unsigned char lds_c32dma[];
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

    f32x16 accr[Cfg::FM][Cfg::FN], acci[Cfg::FM][Cfg::FN];
#pragma unroll
    for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
        for (int j = 0; j < Cfg::FN; ++j) {
            accr[i][j] = f32x16{};
            acci[i][j] = f32x16{};
        }

    auto issue = [&](int ks, int s) {
        unsigned char* stage = lds_c32dma + s * Cfg::STAGE_BYTES;
        c32dma_issue<A_KC>(g.A, g.lda, m0, ks * Cfg::BK, stage, wave, lane);
        c32dma_issue<B_KC>(g.B, g.ldb, n0, ks * Cfg::BK, stage + Cfg::OPER_BYTES, wave, lane);
    };

    if (ks0 < ks1) {
        issue(ks0, 0);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        for (int ks = ks0; ks < ks1; ++ks) {
            const int cur = (ks - ks0) & 1;
            if (ks + 1 < ks1) issue(ks + 1, cur ^ 1);
            const float* As = reinterpret_cast<const float*>(lds_c32dma + cur * Cfg::STAGE_BYTES);
            const float* Bs = reinterpret_cast<const float*>(lds_c32dma + cur * Cfg::STAGE_BYTES + Cfg::OPER_BYTES);
            float ar[2][2][Cfg::FM], ai[2][2][Cfg::FM], br[2][2][Cfg::FN], bi[2][2][Cfg::FN];
            c32dma_frag<A_KC, CA, Cfg::FM>(As, wm0, 0, lane, ar[0], ai[0]);
            c32dma_frag<B_KC, CB, Cfg::FN>(Bs, wn0, 0, lane, br[0], bi[0]);
#pragma unroll
            for (int q2 = 0; q2 < 2; ++q2) {
                if (q2 == 0) {
                    // second k-group's fragments read while the first group's MFMAs issue
                    c32dma_frag<A_KC, CA, Cfg::FM>(As, wm0, 1, lane, ar[1], ai[1]);
                    c32dma_frag<B_KC, CB, Cfg::FN>(Bs, wn0, 1, lane, br[1], bi[1]);
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int t = 0; t < 2; ++t) {
#pragma unroll
                    for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
                        for (int j = 0; j < Cfg::FN; ++j) {
                            accr[i][j] = mfma_32x32x2_f32(br[q2][t][j], ar[q2][t][i], accr[i][j]);
                            acci[i][j] = mfma_32x32x2_f32(bi[q2][t][j], ar[q2][t][i], acci[i][j]);
                        }
#pragma unroll
                    for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
                        for (int j = 0; j < Cfg::FN; ++j) {
                            accr[i][j] = mfma_32x32x2_f32(-bi[q2][t][j], ai[q2][t][i], accr[i][j]);
                            acci[i][j] = mfma_32x32x2_f32(br[q2][t][j], ai[q2][t][i], acci[i][j]);
                        }
                }
            }
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }
    }
#pragma unroll
    for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
        for (int j = 0; j < Cfg::FN; ++j)
            store_complex_tile<false, L3>(g, accr[i][j], acci[i][j], m0 + wm0 + 32 * i + (lane & 31),
                                          n0 + wn0 + 32 * j, lane);
}

template <bool A_KC, bool CA, bool B_KC, bool CB>
hipError_t run_cgemm_f32_dma(const GemmArgs* g, hipStream_t s)
{
    if (uses_level3(g))
        return launch_gemm<C32DmaCfg::BM, C32DmaCfg::BN, C32DmaCfg::NT, C32DmaCfg::LDS_BYTES>(
            cgemm_f32_dma_kernel<A_KC, CA, B_KC, CB, true>, g, s);
    return launch_gemm<C32DmaCfg::BM, C32DmaCfg::BN, C32DmaCfg::NT, C32DmaCfg::LDS_BYTES>(
        cgemm_f32_dma_kernel<A_KC, CA, B_KC, CB, false>, g, s);
}

// op values: 0 = N, 1 = T, 2 = C. op(A) = A has rows (m) contiguous; op(B) = B has k contiguous.
hipError_t cgemm_f32_dma_dispatch(const icl_gemm_args* g, int opa, int opb, int edge, void* ws, hipStream_t s)
{
    if (edge) return cgemm_dispatch<false, 128, 128, 8, 2, 2>(g, opa, opb, edge, ws, s);
    const int oa = opa < 0 || opa > 2 ? 1 : opa;
    const int ob = opb < 0 || opb > 2 ? 1 : opb;
    switch (oa * 3 + ob) {
    case 0: return run_cgemm_f32_dma<false, false, true, false>(g, s);
    case 1: return run_cgemm_f32_dma<false, false, false, false>(g, s);
    case 2: return run_cgemm_f32_dma<false, false, false, true>(g, s);
    case 3: return run_cgemm_f32_dma<true, false, true, false>(g, s);
    case 4: return run_cgemm_f32_dma<true, false, false, false>(g, s);
    case 5: return run_cgemm_f32_dma<true, false, false, true>(g, s);
    case 6: return run_cgemm_f32_dma<true, true, true, false>(g, s);
    case 7: return run_cgemm_f32_dma<true, true, false, false>(g, s);
    default: return run_cgemm_f32_dma<true, true, false, true>(g, s);
    }
}

} // namespace icl

static const icl_gemm_variant kTable[] = {
    {"cgemm_f32_dma_128x128", ICL_KIND_CGEMM, ICL_MATH_FP32, 128, 128, 8, 256, icl::cgemm_f32_dma_dispatch, nullptr,
     142.5f, 2, 1},
};
ICL_REGISTER_VARIANTS(kTable)
