// GEMM variant table (see gemm_variants.hpp).
#include "gemm_variants.hpp"

static const icl_gemm_variant kTable[] = {
    ICL_SGEMM_F32("sgemm_f32_mfma_256x256w8", 256, 256, 8, 4, 2, 133.2f, 1),
};
ICL_REGISTER_VARIANTS(kTable)
