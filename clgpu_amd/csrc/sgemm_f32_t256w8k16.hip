// GEMM variant table (see gemm_variants.hpp).
#include "gemm_variants.hpp"

static const icl_gemm_variant kTable[] = {
    ICL_SGEMM_F32("sgemm_f32_mfma_256x256w8k16", 256, 256, 16, 4, 2, 136.8f, 1),
};
ICL_REGISTER_VARIANTS(kTable)
