/*
 * iclBLAS.h -- C API of the MI355X (gfx950) level-3 BLAS library.
 *
 * Drop-in replacement for the public header of intel/clGPU
 * (include/iclBLAS/iclBLAS.h): identical type names, enum values, status
 * values and function signatures, so code written against the reference
 * compiles and links unchanged against libiclBLAS.so.
 *
 * Implemented on the GPU: iclblasCreate, iclblasDestroy (iclBLAS.h:172,179),
 * iclblasSgemm (:1685), iclblasStrsm (:1719), iclblasCgemm (:1799),
 * iclblasCtrsm (:1925), and the GEMM-based level-3 routines iclblasSsymm
 * (:1579), Ssyrk (:1608), Ssyr2k (:1645), Strmm (:1753), Csymm (:1825), Csyr2k
 * (:1862), Csyrk (:1891), Cherk (:1954), Cher2k (:1991), Ctrmm (:2025), Chemm
 * (:2051), and the level-2 iclblasSgemv (:1049), iclblasCgemv (:1164).
 * Every other entry point is exported for link compatibility and returns
 * ICLBLAS_STATUS_NOT_SUPPORTED, the status the reference reports for an
 * unimplemented function (libs/blas/shlib/iclBLASImpl.hpp:53-56).
 *
 * Matrices are column-major; element (r, c) of a matrix with leading
 * dimension ld lives at index c*ld + r (libs/blas/funcs/iclblas_common.h.cl:47-61).
 * In the default memory mode all matrix pointers are host pointers and every
 * call returns after the result is back in host memory (iclBLASImpl.hpp:236-245).
 * iclBLAS_ext.h adds a device-pointer mode, a split-BF16 math mode, refined
 * STRSM and adjoint entry points.
 */
#pragma once

#ifdef EXPORT_ICLBLAS_SYMBOLS
#  define ICLBLAS_API __attribute__((visibility("default")))
#else
#  define ICLBLAS_API
#endif

#if defined(__GNUC__) || defined(__clang__)
#  define ICLBLAS_ALWAYS_INLINE static inline __attribute__((always_inline))
#else
#  define ICLBLAS_ALWAYS_INLINE static inline
#endif

/* Complex single precision: two floats {re, im}, 8 bytes, same layout in C and C++
 * (iclBLAS.h:71-87). */
#ifdef __cplusplus
#  include <complex>
typedef std::complex<float> oclComplex_t;
inline float Creal(const oclComplex_t& a) { return a.real(); }
inline float Cimag(const oclComplex_t& a) { return a.imag(); }
inline void Csetreal(oclComplex_t* a, float r) { a->real(r); }
inline void Csetimag(oclComplex_t* a, float i) { a->imag(i); }
#else
typedef struct _oclComplex_t {
    float val[2]; /* val[0] = real part, val[1] = imaginary part */
} oclComplex_t;
ICLBLAS_ALWAYS_INLINE float Creal(struct _oclComplex_t a) { return a.val[0]; }
ICLBLAS_ALWAYS_INLINE float Cimag(struct _oclComplex_t a) { return a.val[1]; }
ICLBLAS_ALWAYS_INLINE void Csetreal(struct _oclComplex_t* a, float r) { a->val[0] = r; }
ICLBLAS_ALWAYS_INLINE void Csetimag(struct _oclComplex_t* a, float i) { a->val[1] = i; }
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (iclBLAS.h:104-117). */
typedef enum {
    ICLBLAS_STATUS_SUCCESS = 0,
    ICLBLAS_STATUS_NOT_INITIALIZED = 1,
    ICLBLAS_STATUS_ALLOC_FAILED = 3,
    ICLBLAS_STATUS_INVALID_VALUE = 7,
    ICLBLAS_STATUS_ARCH_MISMATCH = 8,
    ICLBLAS_STATUS_MAPPING_ERROR = 11,
    ICLBLAS_STATUS_EXECUTION_FAILED = 13,
    ICLBLAS_STATUS_INTERNAL_ERROR = 14,
    ICLBLAS_STATUS_NOT_SUPPORTED = 15,
    ICLBLAS_STATUS_LICENSE_ERROR = 16,
    ICLBLAS_STATUS_ERROR /* = 17 */
} iclblasStatus_t;

/* Opaque library context (one HIP device + one HIP stream). */
typedef struct iclblasContext* iclblasHandle_t;

/* op(X) selector (iclBLAS.h:128-132). For real data OP_C means OP_T. */
typedef enum {
    ICLBLAS_OP_N = 0,
    ICLBLAS_OP_T = 1,
    ICLBLAS_OP_C = 2
} iclblasOperation_t;

/* Which triangle of a triangular matrix is stored (iclBLAS.h:137-140). */
typedef enum {
    ICLBLAS_FILL_MODE_UPPER = 0,
    ICLBLAS_FILL_MODE_LOWER = 1
} iclblasFillMode_t;

/* Unit diagonal: the diagonal is taken as 1 and never read (iclBLAS.h:147-150). */
typedef enum {
    ICLBLAS_DIAG_NON_UNIT = 0,
    ICLBLAS_DIAG_UNIT = 1
} iclblasDiagType_t;

/* Side of the triangular operand (iclBLAS.h:155-158). */
typedef enum {
    ICLBLAS_SIDE_LEFT = 0,
    ICLBLAS_SIDE_RIGHT = 1
} iclblasSideMode_t;

/* ---- hot path: implemented on the GPU ---- */
/* Creates a context bound to the current HIP device (iclBLAS.h:172). */
ICLBLAS_API iclblasStatus_t iclblasCreate(iclblasHandle_t* handle);

/* Releases a context created by iclblasCreate (iclBLAS.h:179). */
ICLBLAS_API iclblasStatus_t iclblasDestroy(iclblasHandle_t handle);

/* C = alpha*op(A)*op(B) + beta*C, op(A) m x k, op(B) k x n, C m x n (iclBLAS.h:1648-1685).
 * beta == 0: C is not read. */
ICLBLAS_API iclblasStatus_t iclblasSgemm(iclblasHandle_t handle, iclblasOperation_t transa, iclblasOperation_t transb, int m, int n, int k, const float* alpha, float* A, int lda, float* B, int ldb, const float* beta, float* C, int ldc);

/* Solves op(A)*X = alpha*B (left) or X*op(A) = alpha*B (right); X overwrites B
 * (iclBLAS.h:1687-1719). alpha == 0: B is set to 0 and neither A nor B is read. */
ICLBLAS_API iclblasStatus_t iclblasStrsm(iclblasHandle_t handle, iclblasSideMode_t side, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int m, int n, const float* alpha, float* A, int lda, float* B, int ldb);

/* Complex C = alpha*op(A)*op(B) + beta*C, op in {N, T, C=conjugate transpose}
 * (iclBLAS.h:1762-1799). beta == 0: C is not read. */
ICLBLAS_API iclblasStatus_t iclblasCgemm(iclblasHandle_t handle, iclblasOperation_t transa, iclblasOperation_t transb, int m, int n, int k, const oclComplex_t* alpha, oclComplex_t* A, int lda, oclComplex_t* B, int ldb, const oclComplex_t* beta, oclComplex_t* C, int ldc);

/* Complex solve op(A)*X = alpha*B (left) or X*op(A) = alpha*B (right), op in
 * {N, T, C=conjugate transpose}; X overwrites B (iclBLAS.h:1889-1925).
 * alpha == 0: B is set to 0 and neither A nor B is read. */
ICLBLAS_API iclblasStatus_t iclblasCtrsm(iclblasHandle_t handle, iclblasSideMode_t side, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int m, int n, const oclComplex_t* alpha, oclComplex_t* A, int lda, oclComplex_t* B, int ldb);

/* ---- GEMM-based level-3 routines ---- */
/* C = alpha*A*B + beta*C (left) or alpha*B*A + beta*C (right), A symmetric with
 * only its uplo triangle read, B and C m x n (iclBLAS.h:1555-1579). beta == 0: C
 * is not read. alpha == 0 and beta == 1: returns at once. */
ICLBLAS_API iclblasStatus_t iclblasSsymm(iclblasHandle_t handle, iclblasSideMode_t side, iclblasFillMode_t uplo, int m, int n, const float* alpha, float* A, int lda, float* B, int ldb, const float* beta, float* C, int ldc);

/* C = alpha*op(A)*op(A)^T + beta*C on the uplo triangle of C; op(A) n x k, op in
 * {N, T} (C means T) (iclBLAS.h:1581-1608). The other triangle is not touched.
 * n == 0, k == 0, or alpha == 0 and beta == 1: returns at once. */
ICLBLAS_API iclblasStatus_t iclblasSsyrk(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, int n, int k, const float* alpha, float* A, int lda, const float* beta, float* C, int ldc);

/* C = alpha*(op(A)*op(B)^T + op(B)*op(A)^T) + beta*C on the uplo triangle of C
 * (iclBLAS.h:1610-1645); otherwise as iclblasSsyrk. */
ICLBLAS_API iclblasStatus_t iclblasSsyr2k(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, int n, int k, const float* alpha, float* A, int lda, float* B, int ldb, const float* beta, float* C, int ldc);

/* C = alpha*op(A)*B (left) or alpha*B*op(A) (right), A triangular (uplo, diag),
 * B and C m x n, C written and not read; C may be the same matrix as B
 * (iclBLAS.h:1721-1753). alpha == 0: C = 0, A and B not read. */
ICLBLAS_API iclblasStatus_t iclblasStrmm(iclblasHandle_t handle, iclblasSideMode_t side, iclblasFillMode_t uplo, iclblasOperation_t transa, iclblasDiagType_t diag, int m, int n, const float* alpha, float* A, int lda, float* B, int ldb, float* C, int ldc);

/* Complex symmetric (not hermitian) version of iclblasSsymm (iclBLAS.h:1801-1825). */
ICLBLAS_API iclblasStatus_t iclblasCsymm(iclblasHandle_t handle, iclblasSideMode_t side, iclblasFillMode_t uplo, int m, int n, const oclComplex_t* alpha, oclComplex_t* A, int lda, oclComplex_t* B, int ldb, const oclComplex_t* beta, oclComplex_t* C, int ldc);

/* Complex symmetric rank-2k update C = alpha*(op(A)*op(B)^T + op(B)*op(A)^T) +
 * beta*C on the uplo triangle; op C means op(X) = X^H, as the reference kernel
 * computes it (iclBLAS.h:1827-1862). */
ICLBLAS_API iclblasStatus_t iclblasCsyr2k(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, int n, int k, const oclComplex_t* alpha, oclComplex_t* A, int lda, oclComplex_t* B, int ldb, const oclComplex_t* beta, oclComplex_t* C, int ldc);

/* Complex symmetric rank-k update C = alpha*op(A)*op(A)^T + beta*C on the uplo
 * triangle; op C means op(A) = A^H, as the reference kernel computes it
 * (iclBLAS.h:1864-1891). */
ICLBLAS_API iclblasStatus_t iclblasCsyrk(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, int n, int k, const oclComplex_t* alpha, oclComplex_t* A, int lda, const oclComplex_t* beta, oclComplex_t* C, int ldc);

/* C = alpha*op(A)*op(A)^H + beta*C on the uplo triangle, op in {N, C} (T means C),
 * alpha and beta real; the diagonal's imaginary parts are stored as 0
 * (iclBLAS.h:1927-1954). k == 0: C = beta*C on the triangle. */
ICLBLAS_API iclblasStatus_t iclblasCherk(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, int n, int k, const float* alpha, oclComplex_t* A, int lda, const float* beta, oclComplex_t* C, int ldc);

/* C = alpha*op(A)*op(B)^H + conj(alpha)*op(B)*op(A)^H + beta*C on the uplo
 * triangle, beta real; the diagonal's imaginary parts are stored as 0
 * (iclBLAS.h:1956-1991). */
ICLBLAS_API iclblasStatus_t iclblasCher2k(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, int n, int k, const oclComplex_t* alpha, oclComplex_t* A, int lda, oclComplex_t* B, int ldb, const float* beta, oclComplex_t* C, int ldc);

/* Complex version of iclblasStrmm, op in {N, T, C} (iclBLAS.h:1993-2025). m <= 0,
 * n <= 0 or uplo outside {0, 1}: returns SUCCESS at once, as the reference does. */
ICLBLAS_API iclblasStatus_t iclblasCtrmm(iclblasHandle_t handle, iclblasSideMode_t side, iclblasFillMode_t uplo, iclblasOperation_t transa, iclblasDiagType_t diag, int m, int n, const oclComplex_t* alpha, oclComplex_t* A, int lda, oclComplex_t* B, int ldb, oclComplex_t* C, int ldc);

/* C = alpha*A*B + beta*C (left) or alpha*B*A + beta*C (right), A hermitian with
 * only its uplo triangle read and the diagonal's imaginary parts taken as 0
 * (iclBLAS.h:2027-2051). */
ICLBLAS_API iclblasStatus_t iclblasChemm(iclblasHandle_t handle, iclblasSideMode_t side, iclblasFillMode_t uplo, int m, int n, const oclComplex_t* alpha, oclComplex_t* A, int lda, oclComplex_t* B, int ldb, const oclComplex_t* beta, oclComplex_t* C, int ldc);

/* ---- level 2 ---- */
/* y = alpha*op(A)*x + beta*y, A m x n, op in {N, T, C} (C means T for real data)
 * (iclBLAS.h:1020-1049). beta == 0: y is not read. m == 0, n == 0, or alpha == 0 and
 * beta == 1: returns at once; incx or incy == 0 -> INVALID_VALUE, < 0 -> NOT_SUPPORTED,
 * as in the reference (iclBLAS_S.cpp:477-479, iclBLASImpl.hpp:129-166). */
ICLBLAS_API iclblasStatus_t iclblasSgemv(iclblasHandle_t handle, iclblasOperation_t trans, int m, int n, const float* alpha, float *A, int lda, float *x, int incx, const float* beta, float *y, int incy);

/* Complex version of iclblasSgemv, op C conjugating A (iclBLAS.h:1135-1164). */
ICLBLAS_API iclblasStatus_t iclblasCgemv(iclblasHandle_t handle, iclblasOperation_t trans, int m, int n, const oclComplex_t* alpha, oclComplex_t* A, int lda, oclComplex_t* x, int incx, const oclComplex_t* beta, oclComplex_t* y, int incy);

/* ---- exported for link compatibility; each returns ICLBLAS_STATUS_NOT_SUPPORTED ---- */
ICLBLAS_API iclblasStatus_t iclblasScopy(iclblasHandle_t handle, int n, float *x, int incx, float *y, int incy);
ICLBLAS_API iclblasStatus_t iclblasSscal(iclblasHandle_t handle, int n, const float* alpha, float *x, int incx);
ICLBLAS_API iclblasStatus_t iclblasSaxpy(iclblasHandle_t handle, int n, const float* alpha, float *x, int incx, float *y, int incy);
ICLBLAS_API iclblasStatus_t iclblasSnrm2(iclblasHandle_t handle, int n, float *x, int incx, float *result);
ICLBLAS_API iclblasStatus_t iclblasSrotmg(iclblasHandle_t handle, float *d1, float *d2, float *x1, const float *y1, float* params);
ICLBLAS_API iclblasStatus_t iclblasIsamax(iclblasHandle_t handle, int n, float* x, int incx, int* result);
ICLBLAS_API iclblasStatus_t iclblasIsamin(iclblasHandle_t handle, int n, float* x, int incx, int* result);
ICLBLAS_API iclblasStatus_t iclblasSswap(iclblasHandle_t handle, int n, float* x, int incx, float* y, int incy);
ICLBLAS_API iclblasStatus_t iclblasSrot(iclblasHandle_t handle, int n, float* x, int incx, float* y, int incy, float c, float s);
ICLBLAS_API iclblasStatus_t iclblasSrotm(iclblasHandle_t handle, int n, float* x, int incx, float* y, int incy, float* param);
ICLBLAS_API iclblasStatus_t iclblasSasum(iclblasHandle_t handle, int n, float* x, int incx, float* result);
ICLBLAS_API iclblasStatus_t iclblasSrotg(iclblasHandle_t handle, float* a, float* b, float* c, float* s);
ICLBLAS_API iclblasStatus_t iclblasSdot(iclblasHandle_t handle, int n, float* x, int incx, float* y, int incy, float* result);
ICLBLAS_API iclblasStatus_t iclblasCcopy(iclblasHandle_t handle, int n, oclComplex_t *x, int incx, oclComplex_t *y, int incy);
ICLBLAS_API iclblasStatus_t iclblasCscal(iclblasHandle_t handle, int n, const oclComplex_t* alpha, oclComplex_t *x, int incx);
ICLBLAS_API iclblasStatus_t iclblasCsscal(iclblasHandle_t handle, int n, const float* alpha, oclComplex_t *x, int incx);
ICLBLAS_API iclblasStatus_t iclblasCswap(iclblasHandle_t handle, int n, oclComplex_t* x, int incx, oclComplex_t* y, int incy);
ICLBLAS_API iclblasStatus_t iclblasCdotu(iclblasHandle_t handle, int n, oclComplex_t* x, int incx, oclComplex_t* y, int incy, oclComplex_t* result);
ICLBLAS_API iclblasStatus_t iclblasScnrm2(iclblasHandle_t handle, int n, oclComplex_t* x, int incx, float* result);
ICLBLAS_API iclblasStatus_t iclblasIcamax(iclblasHandle_t handle, int n, oclComplex_t* x, int incx, int* result);
ICLBLAS_API iclblasStatus_t iclblasIcamin(iclblasHandle_t handle, int n, oclComplex_t* x, int incx, int* result);
ICLBLAS_API iclblasStatus_t iclblasCaxpy(iclblasHandle_t handle, int n, const oclComplex_t* alpha, oclComplex_t* x, int incx, oclComplex_t* y, int incy);
ICLBLAS_API iclblasStatus_t iclblasCrotg(iclblasHandle_t handle, oclComplex_t* a, oclComplex_t* b, float* c, oclComplex_t* s);
ICLBLAS_API iclblasStatus_t iclblasCdotc(iclblasHandle_t handle, int n, oclComplex_t* x, int incx, oclComplex_t* y, int incy, oclComplex_t* result);
ICLBLAS_API iclblasStatus_t iclblasCrot(iclblasHandle_t handle, int n, oclComplex_t* x, int incx, oclComplex_t* y, int incy, const float* c, const oclComplex_t* s);
ICLBLAS_API iclblasStatus_t iclblasCsrot(iclblasHandle_t handle, int n, oclComplex_t* x, int incx, oclComplex_t* y, int incy, const float* c, const float* s);
ICLBLAS_API iclblasStatus_t iclblasScasum(iclblasHandle_t handle, int n, oclComplex_t *x, int incx, float* result);
ICLBLAS_API iclblasStatus_t iclblasStrsv(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int n, float* A, int lda, float* x, int incx);
ICLBLAS_API iclblasStatus_t iclblasStbsv(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int n, int k, float* A, int lda, float* x, int incx);
ICLBLAS_API iclblasStatus_t iclblasStpsv(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int n, float* AP, float* x, int incx);
ICLBLAS_API iclblasStatus_t iclblasSger(iclblasHandle_t handle, int m, int n, const float* alpha, float* x, int incx, float* y, int incy, float* A, int lda);
ICLBLAS_API iclblasStatus_t iclblasSsyr2(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const float* alpha, float *x, int incx, float* y, int incy, float* A, int lda);
ICLBLAS_API iclblasStatus_t iclblasSsyr(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const float* alpha, float*x, int incx, float* A, int lda);
ICLBLAS_API iclblasStatus_t iclblasStrmv(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int n, float* A, int lda, float* x, int incx);
ICLBLAS_API iclblasStatus_t iclblasStbmv(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int n, int k, float* A, int lda, float* x, int incx);
ICLBLAS_API iclblasStatus_t iclblasSgbmv(iclblasHandle_t handle, iclblasOperation_t trans, int m, int n, int kl, int ku, const float* alpha, float* A, int lda, float* x, int incx, const float* beta, float* y, int incy);
ICLBLAS_API iclblasStatus_t iclblasStpmv(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int n, float* AP, float* x, int incx);
ICLBLAS_API iclblasStatus_t iclblasSsbmv(iclblasHandle_t handle, iclblasFillMode_t uplo, char n, char k, const float* alpha, float* A, int lda, float* x, int incx, const float* beta, float* y, int incy);
ICLBLAS_API iclblasStatus_t iclblasSspmv(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const float* alpha, float* AP, float* x, int incx, const float* beta, float* y, int incy);
ICLBLAS_API iclblasStatus_t iclblasSspr2(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const float* alpha, float *x, int incx, float* y, int incy, float* AP);
ICLBLAS_API iclblasStatus_t iclblasSspr(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const float* alpha, float *x, int incx, float* AP);
ICLBLAS_API iclblasStatus_t iclblasSsymv(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const float* alpha, float *A, int lda, float *x, int incx, const float* beta, float *y, int incy);
ICLBLAS_API iclblasStatus_t iclblasCgbmv(iclblasHandle_t handle, iclblasOperation_t trans, int m, int n, int kl, int ku, const oclComplex_t* alpha, oclComplex_t* A, int lda, oclComplex_t* x, int incx, const oclComplex_t* beta, oclComplex_t* y, int incy);
ICLBLAS_API iclblasStatus_t iclblasCgeru(iclblasHandle_t handle, int m, int n, const oclComplex_t* alpha, oclComplex_t* x, int incx, oclComplex_t* y, int incy, oclComplex_t* A, int lda);
ICLBLAS_API iclblasStatus_t iclblasCgerc(iclblasHandle_t handle, int m, int n, const oclComplex_t* alpha, oclComplex_t* x, int incx, oclComplex_t* y, int incy, oclComplex_t* A, int lda);
ICLBLAS_API iclblasStatus_t iclblasCher(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const float* alpha, oclComplex_t* x, int incx, oclComplex_t* A, int lda);
ICLBLAS_API iclblasStatus_t iclblasChemv(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const oclComplex_t* alpha, oclComplex_t* A, int lda, oclComplex_t* x, int incx, const oclComplex_t* beta, oclComplex_t* y, int incy);
ICLBLAS_API iclblasStatus_t iclblasCtrmv(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int n, oclComplex_t* A, int lda, oclComplex_t* x, int incx);
ICLBLAS_API iclblasStatus_t iclblasCher2(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const oclComplex_t* alpha, oclComplex_t* x, int incx, oclComplex_t* y, int incy, oclComplex_t* A, int lda);
ICLBLAS_API iclblasStatus_t iclblasChpmv(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const oclComplex_t* alpha, oclComplex_t* AP, oclComplex_t* x, int incx, const oclComplex_t* beta, oclComplex_t* y, int incy);
ICLBLAS_API iclblasStatus_t iclblasChpr(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const float* alpha, oclComplex_t*x, int incx, oclComplex_t* AP);
ICLBLAS_API iclblasStatus_t iclblasChpr2(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const oclComplex_t* alpha, oclComplex_t* x, int incx, oclComplex_t* y, int incy, oclComplex_t* AP);
ICLBLAS_API iclblasStatus_t iclblasCtpmv(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int n, oclComplex_t* AP, oclComplex_t* x, int incx);
ICLBLAS_API iclblasStatus_t iclblasCtbmv(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int n, int k, oclComplex_t* A, int lda, oclComplex_t* x, int incx);
ICLBLAS_API iclblasStatus_t iclblasCtbsv(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int n, int k, oclComplex_t* A, int lda, oclComplex_t* x, int incx);
ICLBLAS_API iclblasStatus_t iclblasCtrsv(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int n, oclComplex_t* A, int lda, oclComplex_t* x, int incx);
ICLBLAS_API iclblasStatus_t iclblasCtpsv(iclblasHandle_t handle, iclblasFillMode_t uplo, iclblasOperation_t trans, iclblasDiagType_t diag, int n, oclComplex_t* AP, oclComplex_t* x, int incx);
ICLBLAS_API iclblasStatus_t iclblasChbmv(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, int k, const oclComplex_t* alpha, oclComplex_t* A, int lda, oclComplex_t* x, int incx, const oclComplex_t* beta, oclComplex_t* y, int incy);
ICLBLAS_API iclblasStatus_t iclblasCsyr(iclblasHandle_t handle, iclblasFillMode_t uplo, int n, const oclComplex_t* alpha, oclComplex_t* x, int incx, oclComplex_t* A, int lda);

#ifdef __cplusplus
} /* extern "C" */
#endif
