// GEMM variant table (see gemm_variants.hpp).
#include "gemm_variants.hpp"

static const icl_gemm_variant kTable[] = {
    ICL_SGEMM_BF16X3("sgemm_bf16x3_mfma_256x256w8", 256, 256, 16, 4, 2, 339.3f, 1),
};
ICL_REGISTER_VARIANTS(kTable)
