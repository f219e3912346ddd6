// GEMM variant table (see gemm_variants.hpp).
#include "gemm_variants.hpp"

static const icl_gemm_variant kTable[] = {
    ICL_SGEMM_F32("sgemm_f32_mfma_64x64", 64, 64, 16, 2, 2, 100.0f, 4),
};
ICL_REGISTER_VARIANTS(kTable)
