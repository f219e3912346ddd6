// Split-BF16 (x3) CGEMM on pre-split, pre-packed operands.
//
// A prepass reads op(A) (m x k) and op(B)^T (n x k) -- applying the conjugate of
// ICLBLAS_OP_C -- splits the real and imaginary part of every element into
// bf16 hi/lo, and writes the four planes in a slab-major layout that is the LDS
// image of one 16-deep k-slab:
//
//   P[ks][r][p][kk]   ks = k / 16, r = row (m for A, n for B),
//                     p = 0 re_hi, 1 re_lo, 2 im_hi, 3 im_lo,  kk = k % 16
//                     offset ((ks * Rpad + r) * 4 + p) * 16 + kk   (128 bytes per row and slab)
//
// The GEMM streams slabs into LDS with the LDS DMA path (global_load_lds_dwordx4,
// XOR-swizzled 16-byte chunks as in the SGEMM kernel) one slab ahead, and forms
// each complex fragment product with twelve v_mfma_f32_32x32x16_bf16:
//   Cr += Br*Ar - Bi*Ai,  Ci += Bi*Ar + Br*Ai,  each real product as hi*hi + lo*hi + hi*lo.
#include "gemm_variants.hpp"

namespace icl {

constexpr int kCPackK = 16; // complex k per slab

template <bool KCONTIG, bool CONJ>
__global__ void __launch_bounds__(256)
csplit_pack_kernel(const float* __restrict__ X, long ld, int R, int K, int Rpad, __bf16* __restrict__ P)
{
    __shared__ f32x2 T[64][kCPackK + 1];
    const int r0 = blockIdx.x * 64;
    const int ks = blockIdx.y;
    const int k0 = ks * kCPackK;
    const int t = threadIdx.x;
    if constexpr (!KCONTIG) {
        const int r = t & 63, kb = t >> 6;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int k = kb + 4 * i;
            T[r][k] = (r0 + r < R && k0 + k < K)
                          ? *reinterpret_cast<const f32x2*>(X + 2 * ((long)(k0 + k) * ld + r0 + r))
                          : f32x2{0.0f, 0.0f};
        }
    } else {
        const int k = t & 15, rb = t >> 4;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = rb + 16 * i;
            T[r][k] = (r0 + r < R && k0 + k < K)
                          ? *reinterpret_cast<const f32x2*>(X + 2 * ((long)(r0 + r) * ld + k0 + k))
                          : f32x2{0.0f, 0.0f};
        }
    }
    __syncthreads();
    const int r = t >> 2, plane = t & 3;
    const bool imag = plane >= 2, low = plane & 1;
    const float sign = (imag && CONJ) ? -1.0f : 1.0f;
    bf16x8 out[2];
#pragma unroll
    for (int i = 0; i < kCPackK; ++i) {
        const float v = sign * (imag ? T[r][i][1] : T[r][i][0]);
        __bf16 h, l;
        split_bf16(v, h, l);
        out[i >> 3][i & 7] = low ? l : h;
    }
    __bf16* dst = P + (((size_t)ks * Rpad + r0 + r) * 4 + plane) * kCPackK;
    *reinterpret_cast<bf16x8*>(dst) = out[0];
    *reinterpret_cast<bf16x8*>(dst + 8) = out[1];
}

struct CPackedArgs {
    icl_gemm_args g; // M, N, C, ldc, alpha, beta (A/B/K unused here)
    const __bf16* Ap;
    const __bf16* Bp;
    int Mpad, Npad, slabs;
};

typedef __attribute__((address_space(1))) void cglobal_void;
typedef __attribute__((address_space(3))) void clds_void;

template <int BM, int BN, int WVM, int WVN>
struct CDmaCfg {
    static constexpr int NW = WVM * WVN;
    static constexpr int NT = NW * 64;
    static constexpr int WM = BM / WVM, WN = BN / WVN;
    static constexpr int FM = WM / 32, FN = WN / 32;
    static constexpr int ROWB = 4 * kCPackK * 2; // 128 bytes per LDS row
    static constexpr int STAGE_BYTES = (BM + BN) * ROWB;
    static constexpr int IA = BM / 8 / NW, IB = BN / 8 / NW;
    static constexpr size_t LDS_BYTES = 2 * (size_t)STAGE_BYTES;
    static_assert((BM / 8) % NW == 0 && (BN / 8) % NW == 0, "rows must divide among waves");
    static_assert(WM % 32 == 0 && WN % 32 == 0, "bad wave tile");
};

// L3: the level-3 instantiation -- tiles of the written triangle only
// (write_tri), the slabs of the tile's k window only (kband), and the epilogue
// with the triangle, the conjugated accumulator and the real diagonal (cflags);
// the plain instantiation is compiled without it.
template <int BM, int BN, int WVM, int WVN, bool EDGE, bool L3 = false>
__global__ void __launch_bounds__(WVM* WVN * 64) cgemm_bf16x3_dma_kernel(CPackedArgs p)
{
    using Cfg = CDmaCfg<BM, BN, WVM, WVN>;
    extern __shared__ __attribute__((aligned(1024))) unsigned char ldsc[];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm0 = (wave % WVM) * Cfg::WM;
    const int wn0 = (wave / WVM) * Cfg::WN;
    int tm, tn;
    if constexpr (L3) {
        if (!tile_coords_level3<BM, BN>(p.g.write_tri, p.g.kband, p.Mpad / BM, p.Npad / BN, tm, tn)) return;
    } else
        tile_coords<8>(p.Mpad / BM, p.Npad / BN, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    int ks0 = 0, ks1 = p.slabs; // slabs [ks0, ks1) of the tile's k window
    if constexpr (L3) {
        int kb = 0, ke = 0;
        if (!tile_window<BM, BN>(p.g, m0, n0, kb, ke)) return;
        ks0 = kb / kCPackK;
        ks1 = ke > kb ? (ke + kCPackK - 1) / kCPackK : ks0;
    }

    const size_t row_elems = 4 * kCPackK;
    const size_t a_slab = (size_t)p.Mpad * row_elems, b_slab = (size_t)p.Npad * row_elems;
    const __bf16* a_src = p.Ap + (size_t)m0 * row_elems;
    const __bf16* b_src = p.Bp + (size_t)n0 * row_elems;
    const int lrow = lane >> 3;

    auto issue = [&](int ks, int s) {
        unsigned char* stage = ldsc + s * Cfg::STAGE_BYTES;
        const __bf16* a = a_src + ks * a_slab;
        const __bf16* b = b_src + ks * b_slab;
#pragma unroll
        for (int i = 0; i < Cfg::IA; ++i) {
            const int r0 = 8 * (wave + Cfg::NW * i);
            const __bf16* g = a + (size_t)(r0 + lrow) * row_elems + ((lane & 7) ^ lds_swizzle(r0 + lrow)) * 8;
            __builtin_amdgcn_global_load_lds((cglobal_void*)g, (clds_void*)(stage + r0 * Cfg::ROWB), 16, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < Cfg::IB; ++i) {
            const int r0 = 8 * (wave + Cfg::NW * i);
            const __bf16* g = b + (size_t)(r0 + lrow) * row_elems + ((lane & 7) ^ lds_swizzle(r0 + lrow)) * 8;
            __builtin_amdgcn_global_load_lds((cglobal_void*)g, (clds_void*)(stage + (BM + r0) * Cfg::ROWB), 16, 0,
                                             0);
        }
    };

    f32x16 accr[Cfg::FM][Cfg::FN], acci[Cfg::FM][Cfg::FN];
#pragma unroll
    for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
        for (int j = 0; j < Cfg::FN; ++j) {
            accr[i][j] = f32x16{};
            acci[i][j] = f32x16{};
        }

    const int half = lane >> 5; // which 8-k half of the 16-deep slab this lane supplies
    if (ks0 < ks1) {
        issue(ks0, 0);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        for (int ks = ks0; ks < ks1; ++ks) {
            const int cur = (ks - ks0) & 1;
            if (ks + 1 < ks1) issue(ks + 1, cur ^ 1);
            const unsigned char* As = ldsc + cur * Cfg::STAGE_BYTES;
            const unsigned char* Bs = As + BM * Cfg::ROWB;
            // fragments: [plane][block], plane chunk = 2 * plane + half
            bf16x8 a[4][Cfg::FM], b[4][Cfg::FN];
#pragma unroll
            for (int i = 0; i < Cfg::FM; ++i) {
                const int r = wm0 + 32 * i + (lane & 31);
                const unsigned char* row = As + r * Cfg::ROWB;
#pragma unroll
                for (int pl = 0; pl < 4; ++pl)
                    a[pl][i] = *reinterpret_cast<const bf16x8*>(row + (((2 * pl + half) ^ lds_swizzle(r)) << 4));
            }
#pragma unroll
            for (int j = 0; j < Cfg::FN; ++j) {
                const int r = wn0 + 32 * j + (lane & 31);
                const unsigned char* row = Bs + r * Cfg::ROWB;
#pragma unroll
                for (int pl = 0; pl < 4; ++pl)
                    b[pl][j] = *reinterpret_cast<const bf16x8*>(row + (((2 * pl + half) ^ lds_swizzle(r)) << 4));
            }
#pragma unroll
            for (int j = 0; j < Cfg::FN; ++j) {
                const bf16x8 nbih = bf16x8_neg(b[2][j]);
                const bf16x8 nbil = bf16x8_neg(b[3][j]);
#pragma unroll
                for (int i = 0; i < Cfg::FM; ++i) {
                    accr[i][j] = mfma_32x32x16_bf16(b[0][j], a[0][i], accr[i][j]);
                    acci[i][j] = mfma_32x32x16_bf16(b[2][j], a[0][i], acci[i][j]);
                    accr[i][j] = mfma_32x32x16_bf16(b[1][j], a[0][i], accr[i][j]);
                    acci[i][j] = mfma_32x32x16_bf16(b[3][j], a[0][i], acci[i][j]);
                    accr[i][j] = mfma_32x32x16_bf16(b[0][j], a[1][i], accr[i][j]);
                    acci[i][j] = mfma_32x32x16_bf16(b[2][j], a[1][i], acci[i][j]);
                    accr[i][j] = mfma_32x32x16_bf16(nbih, a[2][i], accr[i][j]);
                    acci[i][j] = mfma_32x32x16_bf16(b[0][j], a[2][i], acci[i][j]);
                    accr[i][j] = mfma_32x32x16_bf16(nbil, a[2][i], accr[i][j]);
                    acci[i][j] = mfma_32x32x16_bf16(b[1][j], a[2][i], acci[i][j]);
                    accr[i][j] = mfma_32x32x16_bf16(nbih, a[3][i], accr[i][j]);
                    acci[i][j] = mfma_32x32x16_bf16(b[0][j], a[3][i], acci[i][j]);
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
            store_complex_tile<EDGE, L3>(p.g, accr[i][j], acci[i][j], m0 + wm0 + 32 * i + (lane & 31),
                                         n0 + wn0 + 32 * j, lane);
}

namespace {

inline int cround_up(int x, int m) { return (x + m - 1) / m * m; }

template <int BM, int BN>
size_t cpacked_workspace(const icl_gemm_args* g)
{
    const size_t kpad = (size_t)cround_up(g->K, kCPackK);
    return ((size_t)cround_up(g->M, BM) + (size_t)cround_up(g->N, BN)) * kpad * 4 * sizeof(__bf16);
}

hipError_t cpack(const float* X, long ld, int R, int K, int Rpad, int kcontig, int conj, __bf16* P, hipStream_t s)
{
    const dim3 grid((unsigned)(Rpad / 64), (unsigned)(cround_up(K, kCPackK) / kCPackK));
    if (kcontig) {
        if (conj) hipLaunchKernelGGL((csplit_pack_kernel<true, true>), grid, dim3(256), 0, s, X, ld, R, K, Rpad, P);
        else hipLaunchKernelGGL((csplit_pack_kernel<true, false>), grid, dim3(256), 0, s, X, ld, R, K, Rpad, P);
    } else {
        if (conj) hipLaunchKernelGGL((csplit_pack_kernel<false, true>), grid, dim3(256), 0, s, X, ld, R, K, Rpad, P);
        else hipLaunchKernelGGL((csplit_pack_kernel<false, false>), grid, dim3(256), 0, s, X, ld, R, K, Rpad, P);
    }
    return hipGetLastError();
}

template <int BM, int BN, int WVM, int WVN>
hipError_t cpacked_launch(const icl_gemm_args* g, int opa, int opb, int edge, void* ws, hipStream_t s)
{
    if (!ws) return hipErrorInvalidValue;
    (void)edge;
    CPackedArgs p{};
    p.g = *g;
    p.Mpad = cround_up(g->M, BM);
    p.Npad = cround_up(g->N, BN);
    p.slabs = cround_up(g->K, kCPackK) / kCPackK;
    __bf16* Ap = static_cast<__bf16*>(ws);
    __bf16* Bp = Ap + (size_t)p.Mpad * p.slabs * (4 * kCPackK);
    p.Ap = Ap;
    p.Bp = Bp;
    // op(A)(m, k): N -> A[m + k*lda]; T/C -> A[k + m*lda] (k contiguous); C conjugates
    hipError_t e = cpack(g->A, g->lda, g->M, g->K, p.Mpad, opa != 0, opa == 2, Ap, s);
    if (e != hipSuccess) return e;
    // op(B)(k, n) as rows n: N -> B[k + n*ldb] (k contiguous); T/C -> B[n + k*ldb]
    e = cpack(g->B, g->ldb, g->N, g->K, p.Npad, opb == 0, opb == 2, Bp, s);
    if (e != hipSuccess) return e;
    using C = CDmaCfg<BM, BN, WVM, WVN>;
    const bool even = g->M % BM == 0 && g->N % BN == 0;
    if ((e = icl_timer_mark(s)) != hipSuccess) return e;
    if (uses_level3(g))
        e = even ? launch_gemm<BM, BN, C::NT, C::LDS_BYTES>(cgemm_bf16x3_dma_kernel<BM, BN, WVM, WVN, false, true>, g,
                                                           s, p)
                 : launch_gemm<BM, BN, C::NT, C::LDS_BYTES>(cgemm_bf16x3_dma_kernel<BM, BN, WVM, WVN, true, true>, g,
                                                           s, p);
    else if (even)
        e = launch_gemm<BM, BN, C::NT, C::LDS_BYTES>(cgemm_bf16x3_dma_kernel<BM, BN, WVM, WVN, false>, g, s, p);
    else
        e = launch_gemm<BM, BN, C::NT, C::LDS_BYTES>(cgemm_bf16x3_dma_kernel<BM, BN, WVM, WVN, true>, g, s, p);
    if (e != hipSuccess) return e;
    return icl_timer_mark(s);
}

} // namespace
} // namespace icl

// TF: measured useful TFLOP/s (8mnk / t) at m = n = k = 4096 on one MI355X (0 = not yet measured).
#define ICL_CGEMM_BF16X3_PACKED(name, BM, BN, WVM, WVN, TF, OCC)                                                       \
    {name,                                                                                                             \
     ICL_KIND_CGEMM,                                                                                                   \
     ICL_MATH_BF16X3,                                                                                                  \
     BM,                                                                                                               \
     BN,                                                                                                               \
     icl::kCPackK,                                                                                                     \
     64 * WVM * WVN,                                                                                                   \
     icl::cpacked_launch<BM, BN, WVM, WVN>,                                                                            \
     icl::cpacked_workspace<BM, BN>,                                                                                   \
     TF,                                                                                                               \
     OCC,                                                                                                              \
     1}

static const icl_gemm_variant kTable[] = {
    ICL_CGEMM_BF16X3_PACKED("cgemm_bf16x3d_mfma_256x128w8", 256, 128, 4, 2, 495.0f, 1),
    ICL_CGEMM_BF16X3_PACKED("cgemm_bf16x3d_mfma_128x128", 128, 128, 2, 2, 387.0f, 1),
};
ICL_REGISTER_VARIANTS(kTable)
