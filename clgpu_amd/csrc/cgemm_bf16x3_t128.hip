// GEMM variant table (see gemm_variants.hpp).
#include "gemm_variants.hpp"

static const icl_gemm_variant kTable[] = {
    ICL_CGEMM_BF16X3("cgemm_bf16x3_mfma_128x128", 128, 128, 16, 2, 2, 296.0f, 1),
};
ICL_REGISTER_VARIANTS(kTable)
