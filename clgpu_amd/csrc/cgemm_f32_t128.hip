// GEMM variant table (see gemm_variants.hpp).
#include "gemm_variants.hpp"

static const icl_gemm_variant kTable[] = {
    ICL_CGEMM_F32("cgemm_f32_mfma_128x128", 128, 128, 8, 2, 2, 135.5f, 2),
};
ICL_REGISTER_VARIANTS(kTable)
