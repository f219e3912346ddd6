// GEMM variant table (see gemm_variants.hpp).
#include "gemm_variants.hpp"

static const icl_gemm_variant kTable[] = {
    ICL_SGEMM_BF16X3("sgemm_bf16x3_mfma_256x128", 256, 128, 16, 2, 2),
};
ICL_REGISTER_VARIANTS(kTable)
