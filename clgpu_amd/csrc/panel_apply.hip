// X = W * B for one 256 x 256 diagonal panel of the blocked TRSMs (the panel's
// inverse applied to its right-hand sides). As a GEMM this product has only
// 256 x nlines outputs and a depth of 256, so 64 x 64 tiles leave most of the
// GPU idle while the whole solve waits on it. Here a workgroup owns a 32 x 32
// block of X and its four waves take a 64-deep slice of the depth each
// (1024 waves for 1024 lines). Each lane loads exactly the operand values its
// v_mfma_f32_32x32x2_f32 steps take -- W(r0 + lane % 32, k) and
// B(k, l0 + lane % 32) for k = k0 + 2 j + lane / 32, j < 32 -- all loads issued
// before the first product, so the operands never pass through LDS (lanes that
// share a cache line across the j steps hit it in L1); the four partial blocks
// are added through LDS. When an operand of a complex product has k as its
// contiguous dimension (WV: W, BV: B), its lanes read G = 2 or 4 consecutive k
// (16 or 32 bytes), and both operands take step j at
// k = k0 + 2 G (j / G) + G h + j % G (h = lane / 32) instead: the same sum in
// another order, with 1 / (2 G) of the cache lines per load instruction.
#include <cstdint>

#include "mfma.hpp"
#include "kernel_abi.h"

namespace {

constexpr int kPR = 32, kPL = 32, kPK = 64; // rows and lines per workgroup, depth per wave
constexpr int kPLd = 33;                    // padded LDS row (32 + 1 floats)
constexpr int kPanelMax = 256;

template <bool CPLX>
struct PanelCfg {
    static constexpr int NP = CPLX ? 2 : 1;             // planes (re, im)
    static constexpr int PART_FLOATS = NP * kPR * kPLd; // one wave's partial block [r][l]
    static constexpr size_t LDS = sizeof(float) * 4 * PART_FLOATS;
};

template <bool CPLX, int G = 1, bool WV = false, bool BV = false>
__global__ void __launch_bounds__(256) panel_apply_kernel(icl_panel_apply_args a)
{
    static_assert(CPLX || G == 1, "grouped loads are for complex operands");
    static_assert((G == 1) == !(WV || BV) && (G == 1 || G == 2 || G == 4), "G > 1 exactly with a grouped operand");
    using Cfg = PanelCfg<CPLX>;
    extern __shared__ float sm[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r0 = blockIdx.y * kPR, l0 = blockIdx.x * kPL, k0 = wave * kPK;
    using E = typename std::conditional<CPLX, float2, float>::type;
    const E* W = reinterpret_cast<const E*>(a.W);
    const E* B = reinterpret_cast<const E*>(a.B);

    // Operands of step j (k = kof(j)): W(r, k) for this lane's row r and
    // B(k, l) for its line l; out-of-range values read as 0. A grouped load reads
    // k .. k + G - 1 (k a multiple of G, and so is p: all in range or none).
    const int c = lane & 31, h = lane >> 5;
    const int r = r0 + c, l = l0 + c;
    const bool r_ok = r < a.p, l_ok = l < a.nlines;
    const E* wp = W + (r_ok ? (long)r * a.w_rs : 0);
    const E* bp = B + (l_ok ? (long)l * a.b_ls : 0);
    auto kof = [&](int j) { return k0 + 2 * G * (j / G) + G * h + j % G; };
    E wv[kPK / 2], bv[kPK / 2];
#pragma unroll
    for (int j = 0; j < kPK / 2; ++j) {
        const int k = kof(j);
        const bool k_ok = k < a.p;
        if constexpr (WV) {
            if (j % G == 0) {
#pragma unroll
                for (int g = 0; g < G; g += 2) {
                    const float4 v = *reinterpret_cast<const float4*>(wp + (r_ok && k_ok ? k + g : g));
                    wv[j + g] = make_float2(v.x, v.y);
                    wv[j + g + 1] = make_float2(v.z, v.w);
                }
            }
        } else {
            wv[j] = wp[r_ok && k_ok ? (long)k * a.w_ks : 0];
        }
        if constexpr (BV) {
            if (j % G == 0) {
#pragma unroll
                for (int g = 0; g < G; g += 2) {
                    const float4 v = *reinterpret_cast<const float4*>(bp + (l_ok && k_ok ? k + g : g));
                    bv[j + g] = make_float2(v.x, v.y);
                    bv[j + g + 1] = make_float2(v.z, v.w);
                }
            }
        } else {
            bv[j] = bp[l_ok && k_ok ? (long)k * a.b_ks : 0];
        }
    }
#pragma unroll
    for (int j = 0; j < kPK / 2; ++j) {
        const bool k_ok = kof(j) < a.p;
        if (!(r_ok && k_ok)) wv[j] = E{};
        if (!(l_ok && k_ok)) bv[j] = E{};
    }

    // acc(row = lane & 31, line = mfma32_row(lane, i)) over this wave's 64-deep slice
    f32x16 ar = f32x16{}, ai = f32x16{};
#pragma unroll
    for (int j = 0; j < kPK / 2; ++j) {
        if constexpr (CPLX) {
            const float wr = wv[j].x, wi = wv[j].y, br = bv[j].x, bi = bv[j].y;
            ar = icl::mfma_32x32x2_f32(br, wr, ar);
            ai = icl::mfma_32x32x2_f32(bi, wr, ai);
            ar = icl::mfma_32x32x2_f32(-bi, wi, ar);
            ai = icl::mfma_32x32x2_f32(br, wi, ai);
        } else {
            ar = icl::mfma_32x32x2_f32(bv[j], wv[j], ar);
        }
    }
    auto Pt = [&](int w, int pl, int rr, int ll) -> float& {
        return sm[w * Cfg::PART_FLOATS + (pl * kPR + rr) * kPLd + ll];
    };
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int ll = icl::mfma32_row(lane, i);
        Pt(wave, 0, c, ll) = ar[i];
        if constexpr (CPLX) Pt(wave, 1, c, ll) = ai[i];
    }
    __syncthreads();
    const bool x_rfast = a.x_rs == 1;
    E* X = reinterpret_cast<E*>(a.X);
#pragma unroll
    for (int q = 0; q < kPR * kPL / 256; ++q) {
        const int o = tid + 256 * q;
        const int rr = x_rfast ? o % kPR : o / kPL, ll = x_rfast ? o / kPR : o % kPL;
        if (r0 + rr >= a.p || l0 + ll >= a.nlines) continue;
        float sr = Pt(0, 0, rr, ll) + Pt(1, 0, rr, ll) + Pt(2, 0, rr, ll) + Pt(3, 0, rr, ll);
        E* dst = X + (long)(r0 + rr) * a.x_rs + (long)(l0 + ll) * a.x_ls;
        if constexpr (CPLX) {
            const float si = Pt(0, 1, rr, ll) + Pt(1, 1, rr, ll) + Pt(2, 1, rr, ll) + Pt(3, 1, rr, ll);
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
        // Groups of G consecutive k in 16-byte loads: k contiguous, p a multiple
        // of G (groups never straddle the end), 16-byte aligned rows / lines.
        auto a16 = [](const void* q) { return (reinterpret_cast<uintptr_t>(q) & 15u) == 0; };
        const int G = a->p % 4 == 0 ? 4 : a->p % 2 == 0 ? 2 : 1;
        const bool wv = G > 1 && a->w_ks == 1 && a->w_rs % 2 == 0 && a16(a->W);
        const bool bv = G > 1 && a->b_ks == 1 && a->b_ls % 2 == 0 && a16(a->B);
        using K = void (*)(icl_panel_apply_args);
        static const K table[2][2][2] = { // [G == 4][wv][bv]
            {{panel_apply_kernel<true>, panel_apply_kernel<true, 2, false, true>},
             {panel_apply_kernel<true, 2, true, false>, panel_apply_kernel<true, 2, true, true>}},
            {{panel_apply_kernel<true>, panel_apply_kernel<true, 4, false, true>},
             {panel_apply_kernel<true, 4, true, false>, panel_apply_kernel<true, 4, true, true>}}};
        const void* k = reinterpret_cast<const void*>(table[G == 4][wv][bv]);
        icl_panel_apply_args args = *a;
        void* params[] = {&args};
        return hipLaunchKernel(k, grid, dim3(256), params, PanelCfg<true>::LDS, stream);
    } else {
        hipLaunchKernelGGL(panel_apply_kernel<false>, grid, dim3(256), PanelCfg<false>::LDS, stream, *a);
    }
    return hipGetLastError();
}
