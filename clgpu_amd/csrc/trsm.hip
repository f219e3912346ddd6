// STRSM building blocks for gfx950: the diagonal-block solver, the FP64
// residual used by iterative refinement, and 2-D elementwise updates.
//
// The substitution follows the reference kernel's per-column recurrence
// (libs/blas/funcs/Strsm/Strsm_naive.cl:28-37): for pivot i, x_i = b_i / a_ii
// (non-unit), then b_k -= a_ki * x_i for every later k, scaling by alpha last.
#include "mfma.hpp"
#include "kernel_abi.h"

namespace {

constexpr int kNB = 64; // largest diagonal block; one lane owns one right-hand side

// One wavefront solves L * x = s_pre * b for 64 independent right-hand sides,
// one per lane, with x held in 64 registers. L is staged in LDS transposed
// (Lt[i][k] = L(k, i)) so that the update of pivot i reads a contiguous run of
// the column below the diagonal; every lane reads the same address (broadcast).
__global__ void __launch_bounds__(64) trsm_diag_kernel(icl_trsm_diag_args a)
{
    __shared__ float Lt[kNB][kNB + 4];
    __shared__ float dinv_src[kNB];
    const int lane = threadIdx.x;
    const int nb = a.nb;

    // Stage L: column i of L (rows k > i) into Lt[i][k]; unused entries are 0 so
    // that the fully unrolled update below never reads garbage.
    for (int i = 0; i < kNB; ++i) {
        const int k = lane;
        float v = 0.0f;
        if (i < nb && k < nb && k > i) v = a.L[(long)k * a.l_si + (long)i * a.l_sk];
        Lt[i][k] = v;
    }
    if (lane < nb) dinv_src[lane] = a.unit ? 1.0f : a.L[(long)lane * a.l_si + (long)lane * a.l_sk];
    __syncthreads();

    const long j = (long)blockIdx.x * kNB + lane;
    const bool active = j < a.nlines;
    float* xcol = a.X + j * a.x_sj;

    float x[kNB];
#pragma unroll
    for (int i = 0; i < kNB; ++i) x[i] = (active && i < nb) ? a.s_pre * xcol[(long)i * a.x_si] : 0.0f;

#pragma unroll
    for (int i = 0; i < kNB; ++i) {
        if (i < nb) {
            float xi = x[i];
            if (!a.unit) xi = xi / dinv_src[i];
            x[i] = xi;
#pragma unroll
            for (int k = i + 1; k < kNB; ++k) x[k] -= Lt[i][k] * xi;
        }
    }

    if (active) {
#pragma unroll
        for (int i = 0; i < kNB; ++i)
            if (i < nb) xcol[(long)i * a.x_si] = a.s_post * x[i];
    }
}

// ---------------------------------------------------------------------------
// FP64 residual R = fp32(alpha * B0 - P * Q) on v_mfma_f64_16x16x4_f64.
// 64 x 64 output tile per workgroup (4 waves, each 32 x 32 = 2 x 2 MFMA blocks),
// K slabs of 16 staged in LDS as FP64; slabs outside the triangle are skipped.
// ---------------------------------------------------------------------------
constexpr int RT = 64, RK = 16;

__device__ __forceinline__ float tri_load(const float* X, long ld, int trans, int r, int c, int rows, int cols,
                                          bool tri, int lower, int unit)
{
    // element (r, c) of op(X) (rows x cols), with the triangle rule applied
    if (r >= rows || c >= cols) return 0.0f;
    if (tri) {
        if (r == c && unit) return 1.0f;
        if (lower ? (c > r) : (c < r)) return 0.0f;
    }
    return trans ? X[(long)r * ld + c] : X[(long)c * ld + r];
}

__global__ void __launch_bounds__(256) residual_f64_kernel(icl_residual_args a)
{
    __shared__ double Ps[RK][RT + 1];
    __shared__ double Qs[RK][RT + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tiles_m = (a.M + RT - 1) / RT;
    const int m0 = (blockIdx.x % tiles_m) * RT;
    const int n0 = (blockIdx.x / tiles_m) * RT;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;

    // K range that can be non-zero for this tile.
    int k_lo = 0, k_hi = a.K;
    if (a.tri_on == 1) { // P = op(A) is M x K triangular; rows m0..m0+63
        if (a.tri_lower) k_hi = min(a.K, m0 + RT);
        else k_lo = m0;
    } else if (a.tri_on == 2) { // Q = op(A) is K x N triangular; cols n0..n0+63
        if (a.tri_lower) k_lo = n0;
        else k_hi = min(a.K, n0 + RT);
    }
    k_lo = (k_lo / RK) * RK;

    f64x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

    for (int k0 = k_lo; k0 < k_hi; k0 += RK) {
        // P slab: RT rows x RK k; Q slab: RK k x RT cols. 4 elements per thread each.
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int e = tid + 256 * s;
            int r, k;
            if (a.p_trans) { k = e % RK; r = e / RK; } else { r = e % RT; k = e / RT; }
            Ps[k][r] = (double)tri_load(a.P, a.ldp, a.p_trans, m0 + r, k0 + k, a.M, a.K, a.tri_on == 1,
                                        a.tri_lower, a.tri_unit);
            int c;
            if (a.q_trans) { c = e % RT; k = e / RT; } else { k = e % RK; c = e / RK; }
            Qs[k][c] = (double)tri_load(a.Q, a.ldq, a.q_trans, k0 + k, n0 + c, a.K, a.N, a.tri_on == 2,
                                        a.tri_lower, a.tri_unit);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < RK; kk += 4) {
            const int kr = kk + (lane >> 4);
            double p[2], q[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) p[i] = Ps[kr][wm + 16 * i + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 2; ++j) q[j] = Qs[kr][wn + 16 * j + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = icl::mfma_16x16x4_f64(q[j], p[i], acc[i][j]);
        }
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm + 16 * i + (lane & 15);
                const int n = n0 + wn + 16 * j + 4 * r + (lane >> 4);
                if (m < a.M && n < a.N) {
                    const double b = (double)a.B0[(long)n * a.ldb0 + m];
                    a.R[(long)n * a.ldr + m] = (float)(a.alpha * b - acc[i][j][r]);
                }
            }
}

// ---------------------------------------------------------------------------
// Elementwise 2-D updates.
// ---------------------------------------------------------------------------
__global__ void axpby_kernel(int m, int n, float alpha, const float* X, long ldx, float beta, float* Y, long ldy)
{
    const long total = (long)m * n;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const long c = e / m, r = e - c * m;
        float v = (alpha != 0.0f) ? alpha * X[c * ldx + r] : 0.0f;
        if (beta != 0.0f) v = fmaf(beta, Y[c * ldy + r], v);
        Y[c * ldy + r] = v;
    }
}

__global__ void caxpby_kernel(int m, int n, float ar, float ai, const float* X, long ldx, float br, float bi,
                              float* Y, long ldy)
{
    const long total = (long)m * n;
    const bool a_zero = ar == 0.0f && ai == 0.0f;
    const bool b_zero = br == 0.0f && bi == 0.0f;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const long c = e / m, r = e - c * m;
        float vr = 0.0f, vi = 0.0f;
        if (!a_zero) {
            const float xr = X[2 * (c * ldx + r)], xi = X[2 * (c * ldx + r) + 1];
            vr = fmaf(ar, xr, -ai * xi);
            vi = fmaf(ar, xi, ai * xr);
        }
        if (!b_zero) {
            const float yr = Y[2 * (c * ldy + r)], yi = Y[2 * (c * ldy + r) + 1];
            vr += fmaf(br, yr, -bi * yi);
            vi += fmaf(br, yi, bi * yr);
        }
        Y[2 * (c * ldy + r)] = vr;
        Y[2 * (c * ldy + r) + 1] = vi;
    }
}

unsigned elementwise_grid(long total)
{
    long blocks = (total + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    return (unsigned)blocks;
}

} // namespace

extern "C" hipError_t icl_trsm_diag(const icl_trsm_diag_args* a, hipStream_t stream)
{
    if (a->nb <= 0 || a->nlines <= 0) return hipSuccess;
    if (a->nb > kNB) return hipErrorInvalidValue;
    const unsigned blocks = (unsigned)((a->nlines + kNB - 1) / kNB);
    hipLaunchKernelGGL(trsm_diag_kernel, dim3(blocks), dim3(64), 0, stream, *a);
    return hipGetLastError();
}

extern "C" hipError_t icl_residual_f64(const icl_residual_args* a, hipStream_t stream)
{
    if (a->M <= 0 || a->N <= 0) return hipSuccess;
    const long tiles = (long)((a->M + RT - 1) / RT) * ((a->N + RT - 1) / RT);
    hipLaunchKernelGGL(residual_f64_kernel, dim3((unsigned)tiles), dim3(256), 0, stream, *a);
    return hipGetLastError();
}

extern "C" hipError_t icl_axpby_2d(int m, int n, float alpha, const float* X, long ldx, float beta, float* Y,
                                   long ldy, hipStream_t stream)
{
    if (m <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(axpby_kernel, dim3(elementwise_grid((long)m * n)), dim3(256), 0, stream, m, n, alpha, X, ldx,
                       beta, Y, ldy);
    return hipGetLastError();
}

extern "C" hipError_t icl_caxpby_2d(int m, int n, float alpha_re, float alpha_im, const float* X, long ldx,
                                    float beta_re, float beta_im, float* Y, long ldy, hipStream_t stream)
{
    if (m <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(caxpby_kernel, dim3(elementwise_grid((long)m * n)), dim3(256), 0, stream, m, n, alpha_re,
                       alpha_im, X, ldx, beta_re, beta_im, Y, ldy);
    return hipGetLastError();
}
