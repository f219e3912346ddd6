// GEMM variant table (see gemm_variants.hpp).
#include "gemm_variants.hpp"

static const icl_gemm_variant kTable[] = {
    ICL_SGEMM_F32("sgemm_f32_mfma_128x128", 128, 128, 16, 2, 2, 128.7f, 2),
};
ICL_REGISTER_VARIANTS(kTable)
