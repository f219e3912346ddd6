// SGEMV / CGEMV: y = alpha * op(A) * x + beta * y, HBM-bound.
//
// A (m x n, column-major, leading dimension lda) is read exactly once, with
// 16-byte loads when the columns are 16-byte aligned. The reduction dimension is
// split over the grid when the other dimension alone gives too few workgroups
// to fill the 256 CUs; split partial sums go to a workspace and a second kernel
// adds them in a fixed order, so results do not depend on timing (no atomics).
//
// op N: a workgroup owns 64 * VW rows (VW = 4 floats or 2 complex per lane, one
//       16-byte load) and a range of columns; its four waves take interleaved
//       columns, sixteen in flight per lane, and are summed through LDS.
// op T / C: a wave owns two columns and walks their rows (or a split's share of
//       them) two steps of 64 * VW at a time, reusing each x load for both
//       columns, then reduces across its lanes.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernel_abi.h"

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kColsPerWave = 2;  // op T / C
constexpr int kUnroll = 16;      // independent columns per lane in flight (op N)
constexpr int kStepsT = 2;       // row steps per iteration (op T / C)

template <typename T>
struct Num;

template <>
struct Num<float> {
    static constexpr int VW = 4; // elements per 16-byte load
    static __device__ __forceinline__ float zero() { return 0.0f; }
    static __device__ __forceinline__ float fma(float a, float b, float c) { return fmaf(a, b, c); }
    static __device__ __forceinline__ float conj(float a) { return a; }
    static __device__ __forceinline__ float add(float a, float b) { return a + b; }
    static __device__ __forceinline__ float scale(float re, float, float v) { return re * v; }
    static __device__ __forceinline__ float shfl_xor(float v, int m) { return __shfl_xor(v, m); }
    static __device__ __forceinline__ bool is_zero(float re, float) { return re == 0.0f; }
};

template <>
struct Num<float2> {
    static constexpr int VW = 2;
    static __device__ __forceinline__ float2 zero() { return make_float2(0.0f, 0.0f); }
    // c + a * b, the products fused as in the reference's cmul (complex.h.cl:32-34)
    static __device__ __forceinline__ float2 fma(float2 a, float2 b, float2 c)
    {
        return make_float2(c.x + fmaf(a.x, b.x, -a.y * b.y), c.y + fmaf(a.x, b.y, a.y * b.x));
    }
    static __device__ __forceinline__ float2 conj(float2 a) { return make_float2(a.x, -a.y); }
    static __device__ __forceinline__ float2 add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
    static __device__ __forceinline__ float2 scale(float re, float im, float2 v)
    {
        return make_float2(fmaf(re, v.x, -im * v.y), fmaf(re, v.y, im * v.x));
    }
    static __device__ __forceinline__ float2 shfl_xor(float2 v, int m)
    {
        return make_float2(__shfl_xor(v.x, m), __shfl_xor(v.y, m));
    }
    static __device__ __forceinline__ bool is_zero(float re, float im) { return re == 0.0f && im == 0.0f; }
};

// VW consecutive elements starting at p: one 16-byte load when VEC (p aligned)
// and all in range, element loads otherwise (count = elements in range).
template <typename T, bool VEC>
__device__ __forceinline__ void load_run(const T* p, int count, T (&v)[Num<T>::VW])
{
    constexpr int VW = Num<T>::VW;
    if (VEC && count >= VW) {
        const float4 q = *reinterpret_cast<const float4*>(p);
        if constexpr (VW == 4) {
            v[0] = q.x;
            v[1] = q.y;
            v[2] = q.z;
            v[3] = q.w;
        } else {
            v[0] = make_float2(q.x, q.y);
            v[1] = make_float2(q.z, q.w);
        }
    } else {
#pragma unroll
        for (int e = 0; e < VW; ++e) v[e] = e < count ? p[e] : Num<T>::zero();
    }
}

// y(i) = alpha * s + beta * y(i); y not read when beta == 0.
template <typename T>
__device__ __forceinline__ void store_y(const icl_gemv_args& a, long i, T s)
{
    T* y = reinterpret_cast<T*>(a.y);
    T v = Num<T>::scale(a.alpha_re, a.alpha_im, s);
    if (!Num<T>::is_zero(a.beta_re, a.beta_im)) v = Num<T>::add(v, Num<T>::scale(a.beta_re, a.beta_im, y[i * a.incy]));
    y[i * a.incy] = v;
}

// ---------------------------------------------------------------- op N
template <typename T, bool VEC>
__global__ void __launch_bounds__(kThreads) gemv_n_kernel(icl_gemv_args a)
{
    constexpr int VW = Num<T>::VW;
    constexpr int ROWS = 64 * VW;
    __shared__ T red[kWaves - 1][ROWS];
    const T* A = reinterpret_cast<const T*>(a.A);
    const T* x = reinterpret_cast<const T*>(a.x);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = blockIdx.x * ROWS + lane * VW;
    const int count = min(VW, a.m - r); // may be <= 0 for lanes past the end
    const int cs = (a.n + a.splits - 1) / a.splits;
    const int c0 = blockIdx.y * cs, c1 = min(a.n, c0 + cs);

    T acc[VW];
#pragma unroll
    for (int e = 0; e < VW; ++e) acc[e] = Num<T>::zero();
    const T* col = A + r;
    int j = c0 + wave;
    // kUnroll columns of this wave in flight
    for (; j + (kUnroll - 1) * kWaves < c1; j += kUnroll * kWaves) {
        T v[kUnroll][VW];
        T xv[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const long jj = j + u * kWaves;
            if (count > 0) load_run<T, VEC>(col + jj * a.lda, count, v[u]);
            xv[u] = x[jj * a.incx];
        }
        if (count > 0) {
#pragma unroll
            for (int u = 0; u < kUnroll; ++u)
#pragma unroll
                for (int e = 0; e < VW; ++e) acc[e] = Num<T>::fma(v[u][e], xv[u], acc[e]);
        }
    }
    for (; j < c1; j += kWaves) {
        if (count > 0) {
            T v[VW];
            load_run<T, VEC>(col + (long)j * a.lda, count, v);
            const T xj = x[(long)j * a.incx];
#pragma unroll
            for (int e = 0; e < VW; ++e) acc[e] = Num<T>::fma(v[e], xj, acc[e]);
        }
    }

    if (wave > 0) {
#pragma unroll
        for (int e = 0; e < VW; ++e) red[wave - 1][lane * VW + e] = acc[e];
    }
    __syncthreads();
    if (wave != 0) return;
#pragma unroll
    for (int w = 0; w < kWaves - 1; ++w)
#pragma unroll
        for (int e = 0; e < VW; ++e) acc[e] = Num<T>::add(acc[e], red[w][lane * VW + e]);
    T* part = reinterpret_cast<T*>(a.partial);
#pragma unroll
    for (int e = 0; e < VW; ++e) {
        if (e >= count) break;
        if (a.splits == 1) store_y<T>(a, r + e, acc[e]);
        else part[(long)blockIdx.y * a.m + r + e] = acc[e];
    }
}

// ---------------------------------------------------------------- op T / C
template <typename T, bool VEC, bool XVEC, bool CONJ>
__global__ void __launch_bounds__(kThreads) gemv_t_kernel(icl_gemv_args a)
{
    constexpr int VW = Num<T>::VW;
    constexpr int STEP = 64 * VW;
    const T* A = reinterpret_cast<const T*>(a.A);
    const T* x = reinterpret_cast<const T*>(a.x);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j0 = (blockIdx.x * kWaves + wave) * kColsPerWave;
    if (j0 >= a.n) return; // whole wave (no barrier below)
    // Split lengths are whole steps, so every split starts on a 16-byte boundary
    // of the column (a trailing split may be empty and then writes zeros).
    const int rs = ((a.m + a.splits - 1) / a.splits + STEP - 1) / STEP * STEP;
    const int r0 = blockIdx.y * rs, r1 = min(a.m, r0 + rs);
    const bool two = j0 + 1 < a.n;
    const T* c0p = A + (long)j0 * a.lda;
    const T* c1p = A + (long)(two ? j0 + 1 : j0) * a.lda;

    T acc0 = Num<T>::zero(), acc1 = Num<T>::zero();
    auto load_x = [&](int i, int count, T (&xv)[VW]) {
        if constexpr (XVEC) {
            load_run<T, true>(x + i, count, xv);
        } else {
#pragma unroll
            for (int e = 0; e < VW; ++e) xv[e] = e < count ? x[(long)(i + e) * a.incx] : Num<T>::zero();
        }
    };
    int i = r0 + lane * VW;
    // two full row steps per iteration, all their loads in flight together (the
    // same products in the same order as one step at a time)
    for (; i + (kStepsT - 1) * STEP + VW <= r1; i += kStepsT * STEP) {
        T v0[kStepsT][VW], v1[kStepsT][VW], xv[kStepsT][VW];
#pragma unroll
        for (int u = 0; u < kStepsT; ++u) {
            load_run<T, VEC>(c0p + i + u * STEP, VW, v0[u]);
            load_run<T, VEC>(c1p + i + u * STEP, VW, v1[u]);
            load_x(i + u * STEP, VW, xv[u]);
        }
#pragma unroll
        for (int u = 0; u < kStepsT; ++u)
#pragma unroll
            for (int e = 0; e < VW; ++e) {
                const T p0 = CONJ ? Num<T>::conj(v0[u][e]) : v0[u][e];
                const T p1 = CONJ ? Num<T>::conj(v1[u][e]) : v1[u][e];
                acc0 = Num<T>::fma(p0, xv[u][e], acc0);
                acc1 = Num<T>::fma(p1, xv[u][e], acc1);
            }
    }
    for (; i < r1; i += STEP) {
        const int count = min(VW, r1 - i);
        T v0[VW], v1[VW], xv[VW];
        load_run<T, VEC>(c0p + i, count, v0);
        load_run<T, VEC>(c1p + i, count, v1);
        load_x(i, count, xv);
#pragma unroll
        for (int e = 0; e < VW; ++e) {
            const T p0 = CONJ ? Num<T>::conj(v0[e]) : v0[e];
            const T p1 = CONJ ? Num<T>::conj(v1[e]) : v1[e];
            acc0 = Num<T>::fma(p0, xv[e], acc0);
            acc1 = Num<T>::fma(p1, xv[e], acc1);
        }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        acc0 = Num<T>::add(acc0, Num<T>::shfl_xor(acc0, m));
        acc1 = Num<T>::add(acc1, Num<T>::shfl_xor(acc1, m));
    }
    if (lane != 0) return;
    T* part = reinterpret_cast<T*>(a.partial);
    if (a.splits == 1) {
        store_y<T>(a, j0, acc0);
        if (two) store_y<T>(a, j0 + 1, acc1);
    } else {
        part[(long)blockIdx.y * a.n + j0] = acc0;
        if (two) part[(long)blockIdx.y * a.n + j0 + 1] = acc1;
    }
}

// ---------------------------------------------------------------- split sums
template <typename T>
__global__ void __launch_bounds__(kThreads) gemv_reduce_kernel(icl_gemv_args a, int len)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= len) return;
    const T* part = reinterpret_cast<const T*>(a.partial);
    T s = part[i];
    for (int k = 1; k < a.splits; ++k) s = Num<T>::add(s, part[(long)k * len + i]);
    store_y<T>(a, i, s);
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

template <typename T>
hipError_t launch(const icl_gemv_args* a, hipStream_t s)
{
    constexpr int VW = Num<T>::VW;
    if (a->m <= 0 || a->n <= 0) return hipSuccess;
    if (a->splits < 1 || (a->splits > 1 && !a->partial)) return hipErrorInvalidValue;
    // 16-byte loads need every column start aligned: A aligned and lda a multiple of VW.
    const bool vec = aligned16(a->A) && a->lda % VW == 0;
    const int len = a->trans == 0 ? a->m : a->n;
    if (a->trans == 0) {
        const dim3 grid((a->m + 64 * VW - 1) / (64 * VW), a->splits);
        if (vec) hipLaunchKernelGGL((gemv_n_kernel<T, true>), grid, dim3(kThreads), 0, s, *a);
        else hipLaunchKernelGGL((gemv_n_kernel<T, false>), grid, dim3(kThreads), 0, s, *a);
    } else {
        const int cols_per_wg = kWaves * kColsPerWave;
        const dim3 grid((a->n + cols_per_wg - 1) / cols_per_wg, a->splits);
        const bool xvec = a->incx == 1 && aligned16(a->x);
        const bool conj = a->trans == 2;
#define ICL_GEMV_T(V, XV, C) hipLaunchKernelGGL((gemv_t_kernel<T, V, XV, C>), grid, dim3(kThreads), 0, s, *a)
        if (conj) {
            if (vec && xvec) ICL_GEMV_T(true, true, true);
            else if (vec) ICL_GEMV_T(true, false, true);
            else ICL_GEMV_T(false, false, true);
        } else {
            if (vec && xvec) ICL_GEMV_T(true, true, false);
            else if (vec) ICL_GEMV_T(true, false, false);
            else ICL_GEMV_T(false, false, false);
        }
#undef ICL_GEMV_T
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || a->splits == 1) return e;
    hipLaunchKernelGGL(gemv_reduce_kernel<T>, dim3((len + kThreads - 1) / kThreads), dim3(kThreads), 0, s, *a, len);
    return hipGetLastError();
}

} // namespace

extern "C" int icl_gemv_splits(int cplx, int trans, int m, int n)
{
    // Enough workgroups for 8 per CU on 256 CUs, each split still long enough
    // to stream (op N: 64 columns; op T: 1024 rows).
    constexpr int kTarget = 2048;
    const int vw = cplx ? Num<float2>::VW : Num<float>::VW;
    long blocks, max_splits;
    if (trans == 0) {
        blocks = (m + 64L * vw - 1) / (64L * vw);
        max_splits = (n + 63) / 64;
    } else {
        blocks = (n + kWaves * kColsPerWave - 1) / (kWaves * kColsPerWave);
        max_splits = (m + 1023) / 1024;
    }
    long s = (kTarget + blocks - 1) / blocks;
    if (s > max_splits) s = max_splits;
    if (s > 64) s = 64;
    return s < 1 ? 1 : (int)s;
}

extern "C" hipError_t icl_sgemv(const icl_gemv_args* a, hipStream_t stream) { return launch<float>(a, stream); }

extern "C" hipError_t icl_cgemv(const icl_gemv_args* a, hipStream_t stream) { return launch<float2>(a, stream); }
