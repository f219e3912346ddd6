// X = W * B for one 256 x 256 diagonal panel of the blocked TRSMs (the panel's
// inverse applied to its right-hand sides). As a GEMM this product has only
// 256 x nlines outputs and a depth of 256, so 64 x 64 tiles leave most of the
// GPU idle while the whole solve waits on it. Here a workgroup owns a 32 x 32
// block of X and its four waves take a 64-deep slice of the depth each
// (1024 waves for 1024 lines): a wave stages its W and B slices in its own LDS
// region (all loads issued before the first store), runs its MFMA chain, and
// the four partial blocks are added through LDS.
#include "mfma.hpp"
#include "kernel_abi.h"

namespace {

constexpr int kPR = 32, kPL = 32, kPK = 64; // rows and lines per workgroup, depth per wave
constexpr int kPLd = 33;                    // padded LDS row (32 + 1 floats)
constexpr int kPanelMax = 256;

template <bool CPLX>
struct PanelCfg {
    static constexpr int NP = CPLX ? 2 : 1;                     // planes (re, im)
    static constexpr int WAVE_FLOATS = NP * 2 * kPK * kPLd;     // W slice + B slice, [k][r or l]
    static constexpr int PART_FLOATS = NP * kPR * kPLd;         // one wave's partial block [r][l]
    static constexpr int FLOATS = 4 * (WAVE_FLOATS > PART_FLOATS ? WAVE_FLOATS : PART_FLOATS);
    static constexpr size_t LDS = sizeof(float) * FLOATS;
};

template <bool CPLX>
__global__ void __launch_bounds__(256) panel_apply_kernel(icl_panel_apply_args a)
{
    using Cfg = PanelCfg<CPLX>;
    constexpr int NP = Cfg::NP;
    constexpr int STRIDE = Cfg::FLOATS / 4; // floats per wave region
    extern __shared__ float sm[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r0 = blockIdx.y * kPR, l0 = blockIdx.x * kPL, k0 = wave * kPK;
    float* reg = sm + wave * STRIDE;
    // Ws[plane][k][r], Bs[plane][k][l]
    auto Ws = [&](int pl, int k, int r) -> float& { return reg[(pl * kPK + k) * kPLd + r]; };
    auto Bs = [&](int pl, int k, int l) -> float& { return reg[((NP + pl) * kPK + k) * kPLd + l]; };
    using E = typename std::conditional<CPLX, float2, float>::type;
    const E* W = reinterpret_cast<const E*>(a.W);
    const E* B = reinterpret_cast<const E*>(a.B);

    // W slice: 32 rows x 64 k, B slice: 64 k x 32 lines; 32 elements of each per lane.
    // Element e = lane + 64 q takes the unit-stride dimension fastest.
    const bool w_rfast = a.w_rs == 1, b_kfast = a.b_ks == 1;
    E wv[32], bv[32];
#pragma unroll
    for (int q = 0; q < 32; ++q) {
        const int e = lane + 64 * q;
        const int r = w_rfast ? e % kPR : e / kPK, k = w_rfast ? e / kPR : e % kPK;
        const bool ok = r0 + r < a.p && k0 + k < a.p;
        wv[q] = W[ok ? (long)(r0 + r) * a.w_rs + (long)(k0 + k) * a.w_ks : 0];
    }
#pragma unroll
    for (int q = 0; q < 32; ++q) {
        const int e = lane + 64 * q;
        const int k = b_kfast ? e % kPK : e / kPL, l = b_kfast ? e / kPK : e % kPL;
        const bool ok = k0 + k < a.p && l0 + l < a.nlines;
        bv[q] = B[ok ? (long)(k0 + k) * a.b_ks + (long)(l0 + l) * a.b_ls : 0];
    }
#pragma unroll
    for (int q = 0; q < 32; ++q) {
        const int e = lane + 64 * q;
        {
            const int r = w_rfast ? e % kPR : e / kPK, k = w_rfast ? e / kPR : e % kPK;
            const bool ok = r0 + r < a.p && k0 + k < a.p;
            if constexpr (CPLX) {
                Ws(0, k, r) = ok ? wv[q].x : 0.0f;
                Ws(1, k, r) = ok ? wv[q].y : 0.0f;
            } else {
                Ws(0, k, r) = ok ? wv[q] : 0.0f;
            }
        }
        {
            const int k = b_kfast ? e % kPK : e / kPL, l = b_kfast ? e / kPK : e % kPL;
            const bool ok = k0 + k < a.p && l0 + l < a.nlines;
            if constexpr (CPLX) {
                Bs(0, k, l) = ok ? bv[q].x : 0.0f;
                Bs(1, k, l) = ok ? bv[q].y : 0.0f;
            } else {
                Bs(0, k, l) = ok ? bv[q] : 0.0f;
            }
        }
    }
    __syncthreads();

    // acc(row = lane & 31, line = mfma32_row(lane, i)) over this wave's 64-deep slice
    f32x16 ar = f32x16{}, ai = f32x16{};
#pragma unroll 8
    for (int kk = 0; kk < kPK; kk += 2) {
        const int kr = kk + (lane >> 5), c = lane & 31;
        const float wr = Ws(0, kr, c), br = Bs(0, kr, c);
        ar = icl::mfma_32x32x2_f32(br, wr, ar);
        if constexpr (CPLX) {
            const float wi = Ws(1, kr, c), bi = Bs(1, kr, c);
            ai = icl::mfma_32x32x2_f32(bi, wr, ai);
            ar = icl::mfma_32x32x2_f32(-bi, wi, ar);
            ai = icl::mfma_32x32x2_f32(br, wi, ai);
        }
    }
    __syncthreads(); // every wave is done with its slices before they become partial blocks
    auto Pt = [&](int w, int pl, int r, int l) -> float& {
        return sm[w * STRIDE + (pl * kPR + r) * kPLd + l];
    };
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int r = lane & 31, l = icl::mfma32_row(lane, i);
        Pt(wave, 0, r, l) = ar[i];
        if constexpr (CPLX) Pt(wave, 1, r, l) = ai[i];
    }
    __syncthreads();
    const bool x_rfast = a.x_rs == 1;
    E* X = reinterpret_cast<E*>(a.X);
#pragma unroll
    for (int q = 0; q < kPR * kPL / 256; ++q) {
        const int o = tid + 256 * q;
        const int r = x_rfast ? o % kPR : o / kPL, l = x_rfast ? o / kPR : o % kPL;
        if (r0 + r >= a.p || l0 + l >= a.nlines) continue;
        float sr = Pt(0, 0, r, l) + Pt(1, 0, r, l) + Pt(2, 0, r, l) + Pt(3, 0, r, l);
        E* dst = X + (long)(r0 + r) * a.x_rs + (long)(l0 + l) * a.x_ls;
        if constexpr (CPLX) {
            const float si = Pt(0, 1, r, l) + Pt(1, 1, r, l) + Pt(2, 1, r, l) + Pt(3, 1, r, l);
            *dst = make_float2(sr, si);
        } else {
            *dst = sr;
        }
    }
}

} // namespace

extern "C" hipError_t icl_panel_apply(const icl_panel_apply_args* a, hipStream_t stream)
{
    if (a->p <= 0 || a->nlines <= 0) return hipSuccess;
    if (a->p > kPanelMax) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((a->nlines + kPL - 1) / kPL), (unsigned)((a->p + kPR - 1) / kPR));
    if (a->cplx) {
        static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void*>(panel_apply_kernel<true>),
                                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                                           (int)PanelCfg<true>::LDS);
        if (attr != hipSuccess) return attr;
        hipLaunchKernelGGL(panel_apply_kernel<true>, grid, dim3(256), PanelCfg<true>::LDS, stream, *a);
    } else {
        static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void*>(panel_apply_kernel<false>),
                                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                                           (int)PanelCfg<false>::LDS);
        if (attr != hipSuccess) return attr;
        hipLaunchKernelGGL(panel_apply_kernel<false>, grid, dim3(256), PanelCfg<false>::LDS, stream, *a);
    }
    return hipGetLastError();
}
