// Variant tables: each tile configuration of a GEMM kernel family becomes one
// icl_gemm_variant (name, tile sizes, launcher covering every op/edge
// instantiation). Translation units register their tables at load time, the
// way the reference's generated implementation lists feed its dispatcher
// (reference include/iclgpu/functions_base.hpp:193-223); the host layer's
// accept/score selection runs over the registered set.
#pragma once
#include "gemm_launch.hpp"

namespace icl {

struct VariantRegistrar {
    VariantRegistrar(const icl_gemm_variant* table, int count) { icl_register_gemm_variants(table, count); }
};

// ---- SGEMM, FP32 MFMA ------------------------------------------------------
// Problems that use the level-3 fields of icl_gemm_args run the kernel
// instantiation that reads them; plain GEMMs run one compiled without them.
inline bool uses_level3(const GemmArgs* g)
{
    return g->write_tri != ICL_TRI_NONE || g->kband != ICL_KBAND_NONE || g->cflags != 0;
}

template <int TA, int TB, int BM, int BN, int BK, int WVM, int WVN, bool EDGE>
hipError_t run_sgemm_f32(const GemmArgs* g, hipStream_t s)
{
    using Cfg = SgemmF32<TA, TB, BM, BN, BK, WVM, WVN, EDGE>;
    if (uses_level3(g))
        return launch_gemm<BM, BN, Cfg::NT, Cfg::LDS_BYTES>(sgemm_f32_kernel<TA, TB, BM, BN, BK, WVM, WVN, EDGE, true>,
                                                            g, s);
    return launch_gemm<BM, BN, Cfg::NT, Cfg::LDS_BYTES>(sgemm_f32_kernel<TA, TB, BM, BN, BK, WVM, WVN, EDGE, false>, g,
                                                        s);
}

// ---- SGEMM, split-BF16 x3 ----------------------------------------------------
template <int TA, int TB, int BM, int BN, int BK, int WVM, int WVN, bool EDGE>
hipError_t run_sgemm_bf16x3(const GemmArgs* g, hipStream_t s)
{
    using Cfg = SgemmBf16x3<TA, TB, BM, BN, BK, WVM, WVN, EDGE>;
    return launch_gemm<BM, BN, Cfg::NT, Cfg::LDS_BYTES>(sgemm_bf16x3_kernel<TA, TB, BM, BN, BK, WVM, WVN, EDGE>, g,
                                                        s);
}

// For real data ICLBLAS_OP_C is ICLBLAS_OP_T.
template <bool SPLIT, int BM, int BN, int BK, int WVM, int WVN>
hipError_t sgemm_dispatch(const icl_gemm_args* g, int opa, int opb, int edge, void*, hipStream_t s)
{
    const int key = (opa != 0) * 4 + (opb != 0) * 2 + (edge ? 1 : 0);
#define ICL_SGEMM_CASE(K, TA, TB, E)                                                                                   \
    case K:                                                                                                            \
        if constexpr (SPLIT) return run_sgemm_bf16x3<TA, TB, BM, BN, BK, WVM, WVN, E>(g, s);                           \
        else return run_sgemm_f32<TA, TB, BM, BN, BK, WVM, WVN, E>(g, s);
    switch (key) {
        ICL_SGEMM_CASE(0, 0, 0, false)
        ICL_SGEMM_CASE(1, 0, 0, true)
        ICL_SGEMM_CASE(2, 0, 1, false)
        ICL_SGEMM_CASE(3, 0, 1, true)
        ICL_SGEMM_CASE(4, 1, 0, false)
        ICL_SGEMM_CASE(5, 1, 0, true)
        ICL_SGEMM_CASE(6, 1, 1, false)
        ICL_SGEMM_CASE(7, 1, 1, true)
    default: return hipErrorInvalidValue;
    }
#undef ICL_SGEMM_CASE
}

// ---- CGEMM -------------------------------------------------------------------
template <bool SPLIT, int OA, int OB, int BM, int BN, int BK, int WVM, int WVN>
hipError_t run_cgemm(const GemmArgs* g, int edge, hipStream_t s)
{
    if constexpr (SPLIT) {
        using C0 = CgemmBf16x3<OA, OB, BM, BN, BK, WVM, WVN, false>;
        using C1 = CgemmBf16x3<OA, OB, BM, BN, BK, WVM, WVN, true>;
        return edge ? launch_gemm<BM, BN, C1::NT, C1::LDS_BYTES>(
                          cgemm_bf16x3_kernel<OA, OB, BM, BN, BK, WVM, WVN, true>, g, s)
                    : launch_gemm<BM, BN, C0::NT, C0::LDS_BYTES>(
                          cgemm_bf16x3_kernel<OA, OB, BM, BN, BK, WVM, WVN, false>, g, s);
    } else {
        using C0 = CgemmF32<OA, OB, BM, BN, BK, WVM, WVN, false>;
        using C1 = CgemmF32<OA, OB, BM, BN, BK, WVM, WVN, true>;
        if (uses_level3(g))
            return edge ? launch_gemm<BM, BN, C1::NT, C1::LDS_BYTES>(
                              cgemm_f32_kernel<OA, OB, BM, BN, BK, WVM, WVN, true, true>, g, s)
                        : launch_gemm<BM, BN, C0::NT, C0::LDS_BYTES>(
                              cgemm_f32_kernel<OA, OB, BM, BN, BK, WVM, WVN, false, true>, g, s);
        return edge ? launch_gemm<BM, BN, C1::NT, C1::LDS_BYTES>(
                          cgemm_f32_kernel<OA, OB, BM, BN, BK, WVM, WVN, true, false>, g, s)
                    : launch_gemm<BM, BN, C0::NT, C0::LDS_BYTES>(
                          cgemm_f32_kernel<OA, OB, BM, BN, BK, WVM, WVN, false, false>, g, s);
    }
}

template <bool SPLIT, int BM, int BN, int BK, int WVM, int WVN>
hipError_t cgemm_dispatch(const icl_gemm_args* g, int opa, int opb, int edge, void*, hipStream_t s)
{
    const int oa = opa < 0 || opa > 2 ? 1 : opa;
    const int ob = opb < 0 || opb > 2 ? 1 : opb;
    switch (oa * 3 + ob) {
    case 0: return run_cgemm<SPLIT, 0, 0, BM, BN, BK, WVM, WVN>(g, edge, s);
    case 1: return run_cgemm<SPLIT, 0, 1, BM, BN, BK, WVM, WVN>(g, edge, s);
    case 2: return run_cgemm<SPLIT, 0, 2, BM, BN, BK, WVM, WVN>(g, edge, s);
    case 3: return run_cgemm<SPLIT, 1, 0, BM, BN, BK, WVM, WVN>(g, edge, s);
    case 4: return run_cgemm<SPLIT, 1, 1, BM, BN, BK, WVM, WVN>(g, edge, s);
    case 5: return run_cgemm<SPLIT, 1, 2, BM, BN, BK, WVM, WVN>(g, edge, s);
    case 6: return run_cgemm<SPLIT, 2, 0, BM, BN, BK, WVM, WVN>(g, edge, s);
    case 7: return run_cgemm<SPLIT, 2, 1, BM, BN, BK, WVM, WVN>(g, edge, s);
    default: return run_cgemm<SPLIT, 2, 2, BM, BN, BK, WVM, WVN>(g, edge, s);
    }
}

} // namespace icl

// One variant entry: name, kind, math, BM, BN, BK, waves along m, waves along n.
// TF: measured useful TFLOP/s at m = n = k = 8192 (NN) on one MI355X; OCC: workgroups per CU.
#define ICL_SGEMM_F32(name, BM, BN, BK, WVM, WVN, TF, OCC)                                                             \
    {name, ICL_KIND_SGEMM, ICL_MATH_FP32, BM, BN, BK, 64 * WVM * WVN,                                                  \
     icl::sgemm_dispatch<false, BM, BN, BK, WVM, WVN>, nullptr, TF, OCC, 1}
#define ICL_SGEMM_BF16X3(name, BM, BN, BK, WVM, WVN, TF, OCC)                                                          \
    {name, ICL_KIND_SGEMM, ICL_MATH_BF16X3, BM, BN, BK, 64 * WVM * WVN,                                                \
     icl::sgemm_dispatch<true, BM, BN, BK, WVM, WVN>, nullptr, TF, OCC, 0}
#define ICL_CGEMM_F32(name, BM, BN, BK, WVM, WVN, TF, OCC)                                                             \
    {name, ICL_KIND_CGEMM, ICL_MATH_FP32, BM, BN, BK, 64 * WVM * WVN,                                                  \
     icl::cgemm_dispatch<false, BM, BN, BK, WVM, WVN>, nullptr, TF, OCC, 1}
#define ICL_CGEMM_BF16X3(name, BM, BN, BK, WVM, WVN, TF, OCC)                                                          \
    {name, ICL_KIND_CGEMM, ICL_MATH_BF16X3, BM, BN, BK, 64 * WVM * WVN,                                                \
     icl::cgemm_dispatch<true, BM, BN, BK, WVM, WVN>, nullptr, TF, OCC, 0}

#define ICL_REGISTER_VARIANTS(table)                                                                                   \
    namespace {                                                                                                        \
    const icl::VariantRegistrar registrar_##table(table, (int)(sizeof(table) / sizeof(table[0])));                     \
    }
