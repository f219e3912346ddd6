// Entry points of the reference C API that are not on the level-3 hot path.
// Each is exported with the reference signature (include/iclBLAS.h) so that
// programs linked against the reference link against this library, and each
// reports ICLBLAS_STATUS_NOT_SUPPORTED, the reference's status for an
// unimplemented function (libs/blas/shlib/iclBLASImpl.hpp:53-56).
#include "iclBLAS.h"

extern "C" {
ICLBLAS_API iclblasStatus_t iclblasScopy(iclblasHandle_t handle, int n, float *x, int incx, float *y, int incy)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasSscal(iclblasHandle_t handle, int n, const float* alpha, float *x, int incx)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasSaxpy(iclblasHandle_t handle, int n, const float* alpha, float *x, int incx, float *y, int incy)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasSnrm2(iclblasHandle_t handle, int n, float *x, int incx, float *result)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasSrotmg(iclblasHandle_t handle, float *d1, float *d2, float *x1, const float *y1, float* params)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasIsamax(iclblasHandle_t handle, int n, float* x, int incx, int* result)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasIsamin(iclblasHandle_t handle, int n, float* x, int incx, int* result)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasSswap(iclblasHandle_t handle, int n, float* x, int incx, float* y, int incy)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasSrot(iclblasHandle_t handle, int n, float* x, int incx, float* y, int incy, float c, float s)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasSrotm(iclblasHandle_t handle, int n, float* x, int incx, float* y, int incy, float* param)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasSasum(iclblasHandle_t handle, int n, float* x, int incx, float* result)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasSrotg(iclblasHandle_t handle, float* a, float* b, float* c, float* s)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasSdot(iclblasHandle_t handle, int n, float* x, int incx, float* y, int incy, float* result)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasCcopy(iclblasHandle_t handle, int n, oclComplex_t *x, int incx, oclComplex_t *y, int incy)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasCscal(iclblasHandle_t handle, int n, const oclComplex_t* alpha, oclComplex_t *x, int incx)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasCsscal(iclblasHandle_t handle, int n, const float* alpha, oclComplex_t *x, int incx)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasCswap(iclblasHandle_t handle, int n, oclComplex_t* x, int incx, oclComplex_t* y, int incy)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasCdotu(iclblasHandle_t handle, int n, oclComplex_t* x, int incx, oclComplex_t* y, int incy, oclComplex_t* result)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasScnrm2(iclblasHandle_t handle, int n, oclComplex_t* x, int incx, float* result)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasIcamax(iclblasHandle_t handle, int n, oclComplex_t* x, int incx, int* result)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasIcamin(iclblasHandle_t handle, int n, oclComplex_t* x, int incx, int* result)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasCaxpy(iclblasHandle_t handle, int n, const oclComplex_t* alpha, oclComplex_t* x, int incx, oclComplex_t* y, int incy)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasCrotg(iclblasHandle_t handle, oclComplex_t* a, oclComplex_t* b, float* c, oclComplex_t* s)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasCdotc(iclblasHandle_t handle, int n, oclComplex_t* x, int incx, oclComplex_t* y, int incy, oclComplex_t* result)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasCrot(iclblasHandle_t handle, int n, oclComplex_t* x, int incx, oclComplex_t* y, int incy, const float* c, const oclComplex_t* s)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasCsrot(iclblasHandle_t handle, int n, oclComplex_t* x, int incx, oclComplex_t* y, int incy, const float* c, const float* s)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasScasum(iclblasHandle_t handle, int n, oclComplex_t *x, int incx, float* result)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasStrsv(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int n, float* A, int lda, float* x, int incx)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasStbsv(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int n, int k, float* A, int lda, float* x, int incx)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasStpsv(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int n, float* AP, float* x, int incx)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasSger(iclblasHandle_t handle, int m, int n, const float* alpha, float* x, int incx, float* y, int incy, float* A, int lda)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasSsyr2(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const float* alpha, float *x, int incx, float* y, int incy, float* A, int lda)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasSsyr(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const float* alpha, float*x, int incx, float* A, int lda)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasStrmv(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int n, float* A, int lda, float* x, int incx)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasStbmv(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int n, int k, float* A, int lda, float* x, int incx)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasSgbmv(iclblasHandle_t handle, iclblasOperation_t trans, int m, int n, int kl, int ku, const float* alpha, float* A, int lda, float* x, int incx, const float* beta, float* y, int incy)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasStpmv(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int n, float* AP, float* x, int incx)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasSsbmv(iclblasHandle_t handle, iclblasFillMode_t uplo, char n, char k, const float* alpha, float* A, int lda, float* x, int incx, const float* beta, float* y, int incy)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasSspmv(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const float* alpha, float* AP, float* x, int incx, const float* beta, float* y, int incy)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasSspr2(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const float* alpha, float *x, int incx, float* y, int incy, float* AP)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasSspr(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const float* alpha, float *x, int incx, float* AP)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasSsymv(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const float* alpha, float *A, int lda, float *x, int incx, const float* beta, float *y, int incy)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasCgbmv(iclblasHandle_t handle, iclblasOperation_t trans, int m, int n, int kl, int ku, const oclComplex_t* alpha, oclComplex_t* A, int lda, oclComplex_t* x, int incx, const oclComplex_t* beta, oclComplex_t* y, int incy)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasCgeru(iclblasHandle_t handle, int m, int n, const oclComplex_t* alpha, oclComplex_t* x, int incx, oclComplex_t* y, int incy, oclComplex_t* A, int lda)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasCgerc(iclblasHandle_t handle, int m, int n, const oclComplex_t* alpha, oclComplex_t* x, int incx, oclComplex_t* y, int incy, oclComplex_t* A, int lda)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasCher(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const float* alpha, oclComplex_t* x, int incx, oclComplex_t* A, int lda)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasChemv(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const oclComplex_t* alpha, oclComplex_t* A, int lda, oclComplex_t* x, int incx, const oclComplex_t* beta, oclComplex_t* y, int incy)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasCtrmv(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int n, oclComplex_t* A, int lda, oclComplex_t* x, int incx)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasCher2(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const oclComplex_t* alpha, oclComplex_t* x, int incx, oclComplex_t* y, int incy, oclComplex_t* A, int lda)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasChpmv(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const oclComplex_t* alpha, oclComplex_t* AP, oclComplex_t* x, int incx, const oclComplex_t* beta, oclComplex_t* y, int incy)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasChpr(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const float* alpha, oclComplex_t*x, int incx, oclComplex_t* AP)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasChpr2(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const oclComplex_t* alpha, oclComplex_t* x, int incx, oclComplex_t* y, int incy, oclComplex_t* AP)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasCtpmv(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int n, oclComplex_t* AP, oclComplex_t* x, int incx)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasCtbmv(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int n, int k, oclComplex_t* A, int lda, oclComplex_t* x, int incx)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasCtbsv(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int n, int k, oclComplex_t* A, int lda, oclComplex_t* x, int incx)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasCtrsv(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int n, oclComplex_t* A, int lda, oclComplex_t* x, int incx)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasCtpsv(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int n, oclComplex_t* AP, oclComplex_t* x, int incx)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasChbmv(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, int k, const oclComplex_t* alpha, oclComplex_t* A, int lda, oclComplex_t* x, int incx, const oclComplex_t* beta, oclComplex_t* y, int incy)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

ICLBLAS_API iclblasStatus_t iclblasCsyr(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const oclComplex_t* alpha, oclComplex_t* x, int incx, oclComplex_t* A, int lda)
{
    return ICLBLAS_STATUS_NOT_SUPPORTED;
}

} // extern "C"
