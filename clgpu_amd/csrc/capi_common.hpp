// Helpers shared by the C entry points (capi.cpp, capi_level3.cpp): status
// conversion at the extern "C" boundary, per-handle locking, and the staging
// of host matrices through the handle's device scratch buffers in the
// synchronous host-pointer memory mode.
#pragma once
#include <hip/hip_runtime.h>

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

inline bool op_ok(int op) { return op == ICLBLAS_OP_N || op == ICLBLAS_OP_T || op == ICLBLAS_OP_C; }

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
    }
    ~HandleLock() { h->mu.unlock(); }
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

inline void finish(iclblasContext* h)
{
    if (h->memory_mode == ICLBLAS_MEMORY_HOST) icl::hip_check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
    else icl::hip_check(hipGetLastError(), "launch");
}

inline int math_of(iclblasContext* h) { return h->math_mode == ICLBLAS_MATH_SPLIT_BF16X3 ? ICL_MATH_BF16X3 : ICL_MATH_FP32; }

} // namespace capi
} // namespace icl
