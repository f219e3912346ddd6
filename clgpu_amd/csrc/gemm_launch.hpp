// Launch plumbing shared by the GEMM variant tables.
#pragma once
#include "gemm_kernels.hpp"

namespace icl {

// Launches `kern(arg)` over ceil(M/BM) * ceil(N/BN) workgroups (M, N from g),
// or over the tiles that meet the written triangle when g->write_tri is set.
// Kernels that need more than 64 KiB of LDS get the dynamic-LDS limit raised
// once per instantiation.
template <int BM, int BN, int NT, size_t LDS, typename Kern, typename Arg>
inline hipError_t launch_gemm(Kern kern, const GemmArgs* g, hipStream_t s, const Arg& arg)
{
    if constexpr (LDS > 65536) {
        static const hipError_t attr =
            hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                static_cast<int>(LDS));
        if (attr != hipSuccess) return attr;
    }
    const long tiles = g->write_tri != ICL_TRI_NONE ? tri_tile_count(g->M, g->N, g->write_tri, BM, BN)
                                                    : (long)((g->M + BM - 1) / BM) * ((g->N + BN - 1) / BN);
    if (tiles <= 0) return hipSuccess;
    if (tiles > 0x7fffffffL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(kern, dim3((unsigned)tiles), dim3(NT), LDS, s, arg);
    return hipGetLastError();
}

template <int BM, int BN, int NT, size_t LDS, typename Kern>
inline hipError_t launch_gemm(Kern kern, const GemmArgs* g, hipStream_t s)
{
    return launch_gemm<BM, BN, NT, LDS>(kern, g, s, *g);
}

} // namespace icl
