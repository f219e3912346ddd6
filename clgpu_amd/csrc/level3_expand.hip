// Dense expansion of triangular and symmetric / hermitian operands for the
// GEMM-based level-3 routines (TRMM, SYMM, HEMM).
//
// One 256-thread workgroup writes one 64 x 64 tile of W. Every element of the
// tile comes from a single 64 x 64 tile of A (the same tile or the mirrored
// one), which is read once, coalesced down the columns, into LDS with a
// 65-element column pitch; W is written coalesced as well. Element (i, j) of W
// then reads LDS at column pitch 65 either along a column (direct) or along a
// row (mirrored), both free of bank conflicts. Tiles of a triangular operand
// that lie wholly outside the triangle are written as zeros without reading A.
#include <hip/hip_runtime.h>

#include "kernel_abi.h"

namespace {

constexpr int kTile = 64;
constexpr int kPitch = kTile + 1;
constexpr int kThreads = 256;

template <typename T>
struct Elem;
template <>
struct Elem<float> {
    static __device__ __forceinline__ float zero() { return 0.0f; }
    static __device__ __forceinline__ float one() { return 1.0f; }
    static __device__ __forceinline__ float conj(float v) { return v; }
    static __device__ __forceinline__ float real_part(float v) { return v; }
};
template <>
struct Elem<float2> {
    static __device__ __forceinline__ float2 zero() { return make_float2(0.0f, 0.0f); }
    static __device__ __forceinline__ float2 one() { return make_float2(1.0f, 0.0f); }
    static __device__ __forceinline__ float2 conj(float2 v) { return make_float2(v.x, -v.y); }
    static __device__ __forceinline__ float2 real_part(float2 v) { return make_float2(v.x, 0.0f); }
};

// Source of W(i, j): the stored element A(p, q) and whether it is conjugated,
// or a constant (0 outside a triangle, 1 on a unit diagonal).
struct Pick {
    int p, q;
    bool conj;
    int constant; // -1: read A(p, q); 0 or 1: that constant
};

__device__ __forceinline__ bool stored(bool lower, int p, int q) { return lower ? p >= q : p <= q; }

__device__ __forceinline__ Pick pick(const icl_expand_args& a, int i, int j)
{
    Pick s{i, j, false, -1};
    if (a.mode == ICL_EXPAND_TRI) {
        if (a.trans != 0) {
            s.p = j;
            s.q = i;
            s.conj = a.trans == 2;
        }
        if (!stored(a.lower != 0, s.p, s.q)) s.constant = 0;
        else if (a.unit && s.p == s.q) s.constant = 1;
    } else if (!stored(a.lower != 0, i, j)) {
        s.p = j;
        s.q = i;
        s.conj = a.herm != 0;
    }
    return s;
}

template <typename T>
__global__ void __launch_bounds__(kThreads) expand_kernel(icl_expand_args a)
{
    __shared__ T t[kTile * kPitch];
    const T* A = reinterpret_cast<const T*>(a.A);
    T* W = reinterpret_cast<T*>(a.W);
    const int bi = blockIdx.x, bj = blockIdx.y;
    const int i0 = bi * kTile, j0 = bj * kTile;

    // The A tile that holds every source element of this W tile: the mirrored
    // one when the tile's (0, 0) corner is read transposed. For a symmetric
    // diagonal tile both directions fall inside the same tile.
    const Pick corner = pick(a, i0, j0);
    const bool mirrored = corner.p != i0 || corner.q != j0;
    const int r0 = mirrored ? j0 : i0, c0 = mirrored ? i0 : j0;

    // Wholly outside the triangle of a triangular operand: zeros, A not read.
    bool zero_tile = false;
    if (a.mode == ICL_EXPAND_TRI) {
        const bool lower = a.lower != 0;
        zero_tile = lower ? (r0 + kTile - 1 < c0) : (r0 > c0 + kTile - 1);
    }
    if (!zero_tile) {
        const bool skip_diag = a.mode == ICL_EXPAND_TRI && a.unit;
        for (int e = threadIdx.x; e < kTile * kTile; e += kThreads) {
            const int r = e & (kTile - 1), c = e >> 6;
            const int p = r0 + r, q = c0 + c;
            // Read only the stored triangle (the other one may hold anything),
            // and not a unit diagonal.
            if (p < a.order && q < a.order && stored(a.lower != 0, p, q) && !(skip_diag && p == q))
                t[c * kPitch + r] = A[(long)q * a.lda + p];
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < kTile * kTile; e += kThreads) {
        const int li = e & (kTile - 1), lj = e >> 6;
        const int i = i0 + li, j = j0 + lj;
        if (i >= a.order || j >= a.order) continue;
        T v;
        if (zero_tile) {
            v = Elem<T>::zero();
        } else {
            const Pick s = pick(a, i, j);
            if (s.constant == 0) v = Elem<T>::zero();
            else if (s.constant == 1) v = Elem<T>::one();
            else {
                v = t[(s.q - c0) * kPitch + (s.p - r0)];
                if (s.conj) v = Elem<T>::conj(v);
                if (a.mode == ICL_EXPAND_SYM && a.herm && i == j) v = Elem<T>::real_part(v);
            }
        }
        W[(long)j * a.ldw + i] = v;
    }
}

template <typename T>
hipError_t launch_expand(const icl_expand_args* a, hipStream_t stream)
{
    if (a->order <= 0) return hipSuccess;
    const int nt = (a->order + kTile - 1) / kTile;
    hipLaunchKernelGGL(expand_kernel<T>, dim3(nt, nt), dim3(kThreads), 0, stream, *a);
    return hipGetLastError();
}

} // namespace

extern "C" hipError_t icl_expand_real(const icl_expand_args* a, hipStream_t stream)
{
    return launch_expand<float>(a, stream);
}

extern "C" hipError_t icl_expand_complex(const icl_expand_args* a, hipStream_t stream)
{
    return launch_expand<float2>(a, stream);
}
