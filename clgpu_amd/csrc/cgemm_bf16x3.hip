// CGEMM variant table: split-BF16 (x3) MFMA, twelve BF16 MFMAs per complex fragment product.
#include "gemm_launch.hpp"

using namespace icl;

namespace {

template <int OA, int OB, int BM, int BN, int BK, int WVM, int WVN, bool EDGE>
hipError_t run_f32(const GemmArgs* g, hipStream_t s)
{
    using Cfg = CgemmF32<OA, OB, BM, BN, BK, WVM, WVN, EDGE>;
    return launch_gemm<BM, BN, Cfg::NT, Cfg::LDS_BYTES>(cgemm_f32_kernel<OA, OB, BM, BN, BK, WVM, WVN, EDGE>, g, s);
}

template <int OA, int OB, int BM, int BN, int BK, int WVM, int WVN, bool EDGE>
hipError_t run_bf16(const GemmArgs* g, hipStream_t s)
{
    using Cfg = CgemmBf16x3<OA, OB, BM, BN, BK, WVM, WVN, EDGE>;
    return launch_gemm<BM, BN, Cfg::NT, Cfg::LDS_BYTES>(cgemm_bf16x3_kernel<OA, OB, BM, BN, BK, WVM, WVN, EDGE>, g,
                                                        s);
}

template <bool BF16, int OA, int OB, int BM, int BN, int BK, int WVM, int WVN>
hipError_t run_edge(const GemmArgs* g, int edge, hipStream_t s)
{
    if constexpr (BF16) {
        return edge ? run_bf16<OA, OB, BM, BN, BK, WVM, WVN, true>(g, s)
                    : run_bf16<OA, OB, BM, BN, BK, WVM, WVN, false>(g, s);
    } else {
        return edge ? run_f32<OA, OB, BM, BN, BK, WVM, WVN, true>(g, s)
                    : run_f32<OA, OB, BM, BN, BK, WVM, WVN, false>(g, s);
    }
}

template <bool BF16, int BM, int BN, int BK, int WVM, int WVN>
hipError_t dispatch(const icl_gemm_args* g, int opa, int opb, int edge, hipStream_t s)
{
    const int oa = opa < 0 || opa > 2 ? 1 : opa;
    const int ob = opb < 0 || opb > 2 ? 1 : opb;
    switch (oa * 3 + ob) {
    case 0: return run_edge<BF16, 0, 0, BM, BN, BK, WVM, WVN>(g, edge, s);
    case 1: return run_edge<BF16, 0, 1, BM, BN, BK, WVM, WVN>(g, edge, s);
    case 2: return run_edge<BF16, 0, 2, BM, BN, BK, WVM, WVN>(g, edge, s);
    case 3: return run_edge<BF16, 1, 0, BM, BN, BK, WVM, WVN>(g, edge, s);
    case 4: return run_edge<BF16, 1, 1, BM, BN, BK, WVM, WVN>(g, edge, s);
    case 5: return run_edge<BF16, 1, 2, BM, BN, BK, WVM, WVN>(g, edge, s);
    case 6: return run_edge<BF16, 2, 0, BM, BN, BK, WVM, WVN>(g, edge, s);
    case 7: return run_edge<BF16, 2, 1, BM, BN, BK, WVM, WVN>(g, edge, s);
    default: return run_edge<BF16, 2, 2, BM, BN, BK, WVM, WVN>(g, edge, s);
    }
}

const icl_gemm_variant kVariants[] = {
    {"cgemm_bf16x3_mfma_64x64", ICL_KIND_CGEMM, ICL_MATH_BF16X3, 64, 64, 16, 256, dispatch<true, 64, 64, 16, 2, 2>},
    {"cgemm_bf16x3_mfma_128x128", ICL_KIND_CGEMM, ICL_MATH_BF16X3, 128, 128, 16, 256,
     dispatch<true, 128, 128, 16, 2, 2>},
};

} // namespace

extern "C" const icl_gemm_variant* icl_cgemm_bf16x3_variants(int* count)
{
    *count = (int)(sizeof(kVariants) / sizeof(kVariants[0]));
    return kVariants;
}
