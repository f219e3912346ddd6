// Helpers shared by the C entry points (capi.cpp, capi_level3.cpp): status
// conversion at the extern "C" boundary, per-handle locking, and the staging
// of host matrices through the handle's device scratch buffers in the
// synchronous host-pointer memory mode.
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>

#include "blas3.hpp"
#include "handle.hpp"

namespace icl {
namespace capi {

using icl::status_error;

enum Slot { kSlotA = 0, kSlotB = 1, kSlotC = 2, kSlotW0 = 3, kSlotW1 = 4, kSlotW2 = 5 };

template <typename F>
iclblasStatus_t guarded(F&& f)
{
    try {
        f();
        return ICLBLAS_STATUS_SUCCESS;
    } catch (const status_error& e) {
        return e.status;
    } catch (const std::bad_alloc&) {
        return ICLBLAS_STATUS_ALLOC_FAILED;
    } catch (...) {
        return ICLBLAS_STATUS_ERROR;
    }
}

// The int a caller passed for an enum parameter of the C API. A C caller may
// pass any int there (the reference treats every non-zero op as transposed);
// loading a value outside the enumerators' range through the C++ enum type is
// undefined, so entry points read the parameter's representation instead.
template <typename E>
inline int enum_arg(const E& e)
{
    static_assert(sizeof(E) == sizeof(int), "C API enums are int-sized");
    int v;
    std::memcpy(&v, &e, sizeof v);
    return v;
}

inline bool op_ok(int op) { return op == ICLBLAS_OP_N || op == ICLBLAS_OP_T || op == ICLBLAS_OP_C; }

// STRSM / CTRSM validation in the reference's order (iclBLAS_S.cpp:500-510:
// m == 0 or n == 0 -> SUCCESS first; then iclBLASImpl.hpp:108-233), followed by
// the checks the reference does not make; *work is set when the call has work.
inline iclblasStatus_t trsm_checks(iclblasHandle_t handle, int side, int uplo, int trans, int diag, int m, int n,
                                   const void* alpha, const void* A, int lda, const void* B, int ldb, bool* work)
{
    *work = false;
    if (m == 0 || n == 0) return ICLBLAS_STATUS_SUCCESS;
    if (!alpha) return ICLBLAS_STATUS_INVALID_VALUE;
    if (n < 0 || m < 0) return ICLBLAS_STATUS_INVALID_VALUE;
    if (!A || lda <= 0 || !B || ldb <= 0) return ICLBLAS_STATUS_INVALID_VALUE;
    if ((side != ICLBLAS_SIDE_LEFT && side != ICLBLAS_SIDE_RIGHT) ||
        (uplo != ICLBLAS_FILL_MODE_UPPER && uplo != ICLBLAS_FILL_MODE_LOWER) || !op_ok(trans) ||
        (diag != ICLBLAS_DIAG_NON_UNIT && diag != ICLBLAS_DIAG_UNIT))
        return ICLBLAS_STATUS_INVALID_VALUE;
    const int order = side == ICLBLAS_SIDE_LEFT ? m : n;
    if (lda < order || ldb < m) return ICLBLAS_STATUS_INVALID_VALUE;
    if (!iclblasContext::valid(handle)) return ICLBLAS_STATUS_INVALID_VALUE;
    *work = true;
    return ICLBLAS_STATUS_SUCCESS;
}

// Number of elements a column-major rows x cols matrix with leading dimension ld spans.
inline size_t extent(int rows, int cols, int ld) { return rows == 0 || cols == 0 ? 0 : (size_t)(cols - 1) * ld + rows; }

struct HandleLock {
    iclblasContext* h;
    explicit HandleLock(iclblasContext* hh) : h(hh)
    {
        h->mu.lock();
        if (hipSetDevice(h->device) != hipSuccess) {
            h->mu.unlock();
            throw status_error(ICLBLAS_STATUS_EXECUTION_FAILED, "hipSetDevice");
        }
        icl::t_active_timer = h->timer.on ? &h->timer : nullptr;
    }
    ~HandleLock()
    {
        icl::t_active_timer = nullptr;
        h->mu.unlock();
    }
};

// Host-memory staging for the synchronous host-pointer mode, through the
// handle's pinned buffers. In: the whole extent of a rows x cols matrix with
// leading dimension ld (padding included; it lies inside the caller's
// allocation) goes to a device scratch buffer in one pass. Out: only the rows
// of each column are written back, so padding rows of the caller's buffer are
// never written.
template <typename T>
T* stage_in(iclblasContext* h, int slot, const T* host, int rows, int cols, int ld, bool copy)
{
    const size_t n = extent(rows, cols, ld);
    T* d = static_cast<T*>(h->scratch[slot].get(n * sizeof(T), h->stream));
    if (copy && n) h->stager.h2d(d, host, n * sizeof(T), h->stream);
    return d;
}

template <typename T>
void stage_out(iclblasContext* h, T* host, const T* dev, int rows, int cols, int ld)
{
    if (rows == 0 || cols == 0) return;
    h->stager.d2h(host, dev, rows, cols, ld, sizeof(T), h->stream);
}

// Host-pointer GEMM with the transfers overlapped with the compute, for large
// problems. C is done in column panels: A (and op(B) when its columns are not
// contiguous in B, i.e. op T / C) goes up first on the handle's copy lane, then
// panel j's columns of B (op N) and of C (beta != 0) go up while panel j - 1 is
// computed on the handle's stream, and panel j - 1 comes back while panel j is
// computed. One event per direction and panel orders the two streams.
// compute(dA, dB, dC, c0, w) launches the GEMM of C's columns [c0, c0 + w)
// (dB and dC already point at the panel's first column of op(B) and C).
template <typename T, typename Compute>
void host_gemm_panels(iclblasContext* h, int opb, int m, int n, int k, bool use_ab, bool use_c, const T* A, int ra,
                      int ca, int lda, const T* B, int rb, int cb, int ldb, T* C, int ldc, int width, Compute compute)
{
    h->copy.ensure();
    const hipStream_t cs = h->copy.stream;
    T* dA = use_ab ? static_cast<T*>(h->scratch[kSlotA].get(extent(ra, ca, lda) * sizeof(T), h->stream)) : nullptr;
    T* dB = use_ab ? static_cast<T*>(h->scratch[kSlotB].get(extent(rb, cb, ldb) * sizeof(T), h->stream)) : nullptr;
    T* dC = static_cast<T*>(h->scratch[kSlotC].get(extent(m, n, ldc) * sizeof(T), h->stream));
    const bool b_cols = opb == ICLBLAS_OP_N; // op(B)'s columns are B's columns
    if (use_ab) h->stager.h2d(dA, A, extent(ra, ca, lda) * sizeof(T), cs);
    if (use_ab && !b_cols) h->stager.h2d(dB, B, extent(rb, cb, ldb) * sizeof(T), cs);
    constexpr int E = CopyLane::kEvents;
    int prev_c0 = 0, prev_w = 0, j = 0;
    for (int c0 = 0; c0 < n; c0 += width, ++j) {
        const int w = n - c0 < width ? n - c0 : width;
        if (use_ab && b_cols) h->stager.h2d(dB + (size_t)c0 * ldb, B + (size_t)c0 * ldb, extent(k, w, ldb) * sizeof(T), cs);
        if (use_c) h->stager.h2d(dC + (size_t)c0 * ldc, C + (size_t)c0 * ldc, extent(m, w, ldc) * sizeof(T), cs);
        hip_check(hipEventRecord(h->copy.in[j % E], cs), "hipEventRecord");
        hip_check(hipStreamWaitEvent(h->stream, h->copy.in[j % E], 0), "hipStreamWaitEvent");
        const T* dBp = use_ab ? (b_cols ? dB + (size_t)c0 * ldb : dB + c0) : nullptr;
        compute(dA, dBp, dC + (size_t)c0 * ldc, c0, w);
        hip_check(hipEventRecord(h->copy.out[j % E], h->stream), "hipEventRecord");
        if (j > 0) {
            hip_check(hipStreamWaitEvent(cs, h->copy.out[(j - 1) % E], 0), "hipStreamWaitEvent");
            h->stager.d2h(C + (size_t)prev_c0 * ldc, dC + (size_t)prev_c0 * ldc, m, prev_w, ldc, sizeof(T), cs);
        }
        prev_c0 = c0;
        prev_w = w;
    }
    hip_check(hipStreamWaitEvent(cs, h->copy.out[(j - 1) % E], 0), "hipStreamWaitEvent");
    h->stager.d2h(C + (size_t)prev_c0 * ldc, dC + (size_t)prev_c0 * ldc, m, prev_w, ldc, sizeof(T), cs);
}

// Column-panel width for host_gemm_panels, or 0 when the overlap does not pay
// and one transfer per matrix is cheaper: small problems, and the split-BF16
// math mode, whose compute is short next to the transfers (measured at 8192^3:
// 20.8 ms pipelined against 19.7 ms, where the FP32 path gains 25.9 -> 21.3 ms).
inline int host_gemm_panel_width(const iclblasContext* h, int m, int n, int k)
{
    if (h->math_mode == ICLBLAS_MATH_SPLIT_BF16X3) return 0;
    if ((double)m * n * k < 4294967296.0 || n < 2048) return 0;
    int w = (n + 7) / 8;         // about eight panels
    w = (w + 255) / 256 * 256;   // whole 256-column tiles
    return w;
}

inline void finish(iclblasContext* h)
{
    if (h->memory_mode == ICLBLAS_MEMORY_HOST) icl::hip_check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
    else icl::hip_check(hipGetLastError(), "launch");
}

inline int math_of(iclblasContext* h) { return h->math_mode == ICLBLAS_MATH_SPLIT_BF16X3 ? ICL_MATH_BF16X3 : ICL_MATH_FP32; }

} // namespace capi
} // namespace icl
