// FP32-MFMA SGEMM variant table (v_mfma_f32_32x32x2_f32).
#include "gemm_launch.hpp"

using namespace icl;

namespace {

template <int TA, int TB, int BM, int BN, int BK, int WVM, int WVN, bool EDGE>
hipError_t run(const GemmArgs* g, hipStream_t s)
{
    using Cfg = SgemmF32<TA, TB, BM, BN, BK, WVM, WVN, EDGE>;
    return launch_gemm<BM, BN, Cfg::NT, Cfg::LDS_BYTES>(sgemm_f32_kernel<TA, TB, BM, BN, BK, WVM, WVN, EDGE>, g, s);
}

// For real data ICLBLAS_OP_C is ICLBLAS_OP_T.
template <int BM, int BN, int BK, int WVM, int WVN>
hipError_t dispatch(const icl_gemm_args* g, int opa, int opb, int edge, hipStream_t s)
{
    const int ta = opa != 0, tb = opb != 0;
    const int key = ta * 4 + tb * 2 + (edge ? 1 : 0);
    switch (key) {
    case 0: return run<0, 0, BM, BN, BK, WVM, WVN, false>(g, s);
    case 1: return run<0, 0, BM, BN, BK, WVM, WVN, true>(g, s);
    case 2: return run<0, 1, BM, BN, BK, WVM, WVN, false>(g, s);
    case 3: return run<0, 1, BM, BN, BK, WVM, WVN, true>(g, s);
    case 4: return run<1, 0, BM, BN, BK, WVM, WVN, false>(g, s);
    case 5: return run<1, 0, BM, BN, BK, WVM, WVN, true>(g, s);
    case 6: return run<1, 1, BM, BN, BK, WVM, WVN, false>(g, s);
    default: return run<1, 1, BM, BN, BK, WVM, WVN, true>(g, s);
    }
}

const icl_gemm_variant kVariants[] = {
    {"sgemm_f32_mfma_128x128", ICL_KIND_SGEMM, ICL_MATH_FP32, 128, 128, 16, 256, dispatch<128, 128, 16, 2, 2>},
    {"sgemm_f32_mfma_256x256", ICL_KIND_SGEMM, ICL_MATH_FP32, 256, 256, 8, 256, dispatch<256, 256, 8, 2, 2>},
    {"sgemm_f32_mfma_256x256w8", ICL_KIND_SGEMM, ICL_MATH_FP32, 256, 256, 8, 512, dispatch<256, 256, 8, 4, 2>},
};

} // namespace

extern "C" const icl_gemm_variant* icl_sgemm_f32_variants(int* count)
{
    *count = (int)(sizeof(kVariants) / sizeof(kVariants[0]));
    return kVariants;
}
