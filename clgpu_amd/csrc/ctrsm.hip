// Complex triangular solve of one diagonal block (at most 64 x 64) for CTRSM.
//
// The arithmetic is the reference kernel's (libs/blas/funcs/Ctrsm/Ctrsm_naive.cl:16-183
// with the helpers of core/src/ocl/complex.h.cl:32-39): per pivot i,
// x_i = cdiv(b_i, a_ii) for a non-unit diagonal, then b_k -= cmul(a_ki, x_i) for
// every later k, and the solved value is stored as cmul(x_i, alpha). OP_C
// conjugates every element of A that is read, the diagonal included.
#include "mfma.hpp"
#include "kernel_abi.h"

namespace {

constexpr int kNB = 64; // largest diagonal block; one lane owns one right-hand side

__device__ __forceinline__ float2 c_mul(float2 a, float2 b)
{
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}

__device__ __forceinline__ float2 c_div(float2 a, float2 b)
{
    const float d = fmaf(b.x, b.x, b.y * b.y);
    return make_float2(fmaf(a.x, b.x, a.y * b.y) / d, fmaf(a.y, b.x, -a.x * b.y) / d);
}

// The block staged in LDS transposed, Lt[i][k] = L(k, i) for k > i (conjugated
// for OP_C), so the update of pivot i reads one broadcast value per k; entries
// outside the block are 0 so the fully unrolled update never reads garbage.
struct CDiagBlock {
    float2 Lt[kNB][kNB + 1];
    float2 diag[kNB];
};

__global__ void __launch_bounds__(64) ctrsm_diag_kernel(icl_ctrsm_diag_args a)
{
    __shared__ CDiagBlock s;
    const int lane = threadIdx.x;
    const int nb = a.nb;
    const float2* L = reinterpret_cast<const float2*>(a.L);
    const float sg = a.conj ? -1.0f : 1.0f;
    for (int i = 0; i < kNB; ++i) {
        float2 v = make_float2(0.0f, 0.0f);
        if (i < nb && lane < nb && lane > i) {
            v = L[(long)lane * a.l_si + (long)i * a.l_sk];
            v.y *= sg;
        }
        s.Lt[i][lane] = v;
    }
    if (lane < nb) {
        float2 d = make_float2(1.0f, 0.0f);
        if (!a.unit) {
            d = L[(long)lane * (a.l_si + a.l_sk)];
            d.y *= sg;
        }
        s.diag[lane] = d;
    }
    __syncthreads();

    const long j = (long)blockIdx.x * kNB + lane;
    const bool active = j < a.nlines;
    float2* xcol = reinterpret_cast<float2*>(a.X) + j * a.x_sj;
    float2 x[kNB];
#pragma unroll
    for (int i = 0; i < kNB; ++i) x[i] = (active && i < nb) ? xcol[(long)i * a.x_si] : make_float2(0.0f, 0.0f);
#pragma unroll
    for (int i = 0; i < kNB; ++i) {
        if (i < nb) {
            float2 xi = x[i];
            if (!a.unit) xi = c_div(xi, s.diag[i]);
            x[i] = xi;
#pragma unroll
            for (int k = i + 1; k < kNB; ++k) {
                const float2 t = c_mul(s.Lt[i][k], xi);
                x[k].x -= t.x;
                x[k].y -= t.y;
            }
        }
    }
    if (active) {
        const float2 sp = make_float2(a.s_post_re, a.s_post_im);
#pragma unroll
        for (int i = 0; i < kNB; ++i)
            if (i < nb) xcol[(long)i * a.x_si] = c_mul(x[i], sp);
    }
}

} // namespace

extern "C" hipError_t icl_ctrsm_diag(const icl_ctrsm_diag_args* a, hipStream_t stream)
{
    if (a->nb <= 0 || a->nlines <= 0) return hipSuccess;
    if (a->nb > kNB) return hipErrorInvalidValue;
    const unsigned blocks = (unsigned)((a->nlines + kNB - 1) / kNB);
    hipLaunchKernelGGL(ctrsm_diag_kernel, dim3(blocks), dim3(64), 0, stream, *a);
    return hipGetLastError();
}
