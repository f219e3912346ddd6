// C entry points iclblasSgemv and iclblasCgemv.
//
// Checks in the reference's order (libs/blas/shlib/iclBLAS_S.cpp:475-485,
// iclBLAS_C.cpp:236-248, then iclBLASImpl.hpp:108-233):
//   m == 0 or n == 0 or (alpha == 0 and beta == 1) -> SUCCESS;
//   m < 0, n < 0, incx == 0 or incy == 0 -> INVALID_VALUE;
//   null x -> INVALID_VALUE; incx < 0 -> NOT_SUPPORTED; null y -> INVALID_VALUE;
//   incy < 0 -> NOT_SUPPORTED; null A or lda <= 0 -> INVALID_VALUE;
//   then, as the other entry points, the checks the reference does not make
//   (null alpha / beta, trans out of range, lda < m -> INVALID_VALUE) and the handle.
#include <cstring>

#include "capi_common.hpp"

using namespace icl::capi;

namespace {

typedef float cf[2];

iclblasStatus_t gemv_common(iclblasHandle_t handle, bool cplx, int trans, int m, int n, const float* alpha,
                            const float* A, int lda, const float* x, int incx, const float* beta, float* y, int incy)
{
    if (m == 0 || n == 0) return ICLBLAS_STATUS_SUCCESS;
    if (!alpha || !beta) return ICLBLAS_STATUS_INVALID_VALUE;
    const float a[2] = {alpha[0], cplx ? alpha[1] : 0.0f};
    const float b[2] = {beta[0], cplx ? beta[1] : 0.0f};
    const bool a_zero = a[0] == 0.0f && a[1] == 0.0f;
    if (a_zero && b[0] == 1.0f && b[1] == 0.0f) return ICLBLAS_STATUS_SUCCESS;
    if (m < 0 || n < 0 || incx == 0 || incy == 0) return ICLBLAS_STATUS_INVALID_VALUE;
    if (!x) return ICLBLAS_STATUS_INVALID_VALUE;
    if (incx < 0) return ICLBLAS_STATUS_NOT_SUPPORTED;
    if (!y) return ICLBLAS_STATUS_INVALID_VALUE;
    if (incy < 0) return ICLBLAS_STATUS_NOT_SUPPORTED;
    if (!A || lda <= 0) return ICLBLAS_STATUS_INVALID_VALUE;
    if (!op_ok(trans) || lda < m) return ICLBLAS_STATUS_INVALID_VALUE;
    if (!iclblasContext::valid(handle)) return ICLBLAS_STATUS_INVALID_VALUE;
    const int lenx = trans == ICLBLAS_OP_N ? n : m, leny = trans == ICLBLAS_OP_N ? m : n;
    return guarded([&] {
        HandleLock lock(handle);
        iclblasContext* h = handle;
        if (h->memory_mode == ICLBLAS_MEMORY_DEVICE) {
            icl::gemv_dev(h, cplx, trans, m, n, a, A, lda, x, incx, b, y, incy);
        } else {
            // Vectors are staged as 1 x len matrices with the increment as
            // leading dimension, so the gaps between elements are not written.
            const bool use_ax = !a_zero;
            const bool use_y = b[0] != 0.0f || b[1] != 0.0f;
            const float *dA = nullptr, *dx = nullptr;
            float* dy;
            if (cplx) {
                dA = use_ax ? reinterpret_cast<const float*>(stage_in<cf>(h, kSlotA, reinterpret_cast<const cf*>(A), m, n, lda, true)) : nullptr;
                dx = use_ax ? reinterpret_cast<const float*>(stage_in<cf>(h, kSlotB, reinterpret_cast<const cf*>(x), 1, lenx, incx, true)) : nullptr;
                dy = reinterpret_cast<float*>(stage_in<cf>(h, kSlotC, reinterpret_cast<const cf*>(y), 1, leny, incy, use_y));
            } else {
                dA = use_ax ? stage_in<float>(h, kSlotA, A, m, n, lda, true) : nullptr;
                dx = use_ax ? stage_in<float>(h, kSlotB, x, 1, lenx, incx, true) : nullptr;
                dy = stage_in<float>(h, kSlotC, y, 1, leny, incy, use_y);
            }
            icl::gemv_dev(h, cplx, trans, m, n, a, dA, lda, dx, incx, b, dy, incy);
            if (cplx) stage_out<cf>(h, reinterpret_cast<cf*>(y), reinterpret_cast<const cf*>(dy), 1, leny, incy);
            else stage_out<float>(h, y, dy, 1, leny, incy);
        }
        finish(h);
    });
}

} // namespace

extern "C" {

// reference: libs/blas/shlib/iclBLAS_S.cpp:474-485
ICLBLAS_API iclblasStatus_t iclblasSgemv(iclblasHandle_t handle, iclblasOperation_t trans_, int m, int n,
                                         const float* alpha, float* A, int lda, float* x, int incx, const float* beta,
                                         float* y, int incy)
{
    const int trans = enum_arg(trans_);
    return gemv_common(handle, false, trans, m, n, alpha, A, lda, x, incx, beta, y, incy);
}

// reference: libs/blas/shlib/iclBLAS_C.cpp:235-248
ICLBLAS_API iclblasStatus_t iclblasCgemv(iclblasHandle_t handle, iclblasOperation_t trans_, int m, int n,
                                         const oclComplex_t* alpha, oclComplex_t* A, int lda, oclComplex_t* x, int incx,
                                         const oclComplex_t* beta, oclComplex_t* y, int incy)
{
    const int trans = enum_arg(trans_);
    return gemv_common(handle, true, trans, m, n, reinterpret_cast<const float*>(alpha),
                       reinterpret_cast<const float*>(A), lda, reinterpret_cast<const float*>(x), incx,
                       reinterpret_cast<const float*>(beta), reinterpret_cast<float*>(y), incy);
}

} // extern "C"
