// GEMM variant table (see gemm_variants.hpp).
#include "gemm_variants.hpp"

static const icl_gemm_variant kTable[] = {
    ICL_SGEMM_BF16X3("sgemm_bf16x3_mfma_128x128k32", 128, 128, 32, 2, 2, 260.5f, 1),
};
ICL_REGISTER_VARIANTS(kTable)
