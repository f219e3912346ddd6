// GEMM variant table (see gemm_variants.hpp).
#include "gemm_variants.hpp"

static const icl_gemm_variant kTable[] = {
    ICL_CGEMM_F32("cgemm_f32_mfma_64x64", 64, 64, 16, 2, 2, 131.0f, 2),
};
ICL_REGISTER_VARIANTS(kTable)
