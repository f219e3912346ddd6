/*
 * iclBLAS_ext.h -- MI355X extensions to the iclBLAS C API.
 *
 * None of these entry points exists in intel/clGPU; they are additions that
 * the reference API has no room for. All are plain C: pointers, ints and enums.
 *
 *  - Memory mode: by default (ICLBLAS_MEMORY_HOST) matrix arguments are host
 *    pointers and each call is synchronous, as in the reference
 *    (libs/blas/shlib/iclBLASImpl.hpp:236-245). ICLBLAS_MEMORY_DEVICE takes
 *    device pointers (hipMalloc / torch CUDA tensors), launches on the
 *    handle's HIP stream and returns without waiting. Scalars (alpha, beta)
 *    are host pointers in both modes.
 *  - Math mode: ICLBLAS_MATH_FP32 runs SGEMM/CGEMM and the level-3 routines on
 *    FP32 MFMA. ICLBLAS_MATH_SPLIT_BF16X3 splits every FP32 operand x into
 *    hi = bf16(x), lo = bf16(x - hi) and forms hi*hi + hi*lo + lo*hi on BF16
 *    MFMA with FP32 accumulation, for products at least 256 deep (shallower
 *    ones stay on FP32 MFMA); the error stays within the FP32 bound
 *    4 k 2^-24 (|A||B|).
 *  - Refined STRSM: FP32 blocked solve followed by sweeps of iterative
 *    refinement whose residual R = alpha*B - op(A)*X is formed in FP64.
 *  - Adjoints of SGEMM and STRSM.
 *  - Implementation override: run a named kernel variant, bypassing the
 *    accept/score selection (the reference's tests do the same through
 *    tests/blas_funcs/test_helpers.hpp:78-102).
 */
#pragma once
#include <stddef.h>

#include "iclBLAS.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    ICLBLAS_MEMORY_HOST = 0,
    ICLBLAS_MEMORY_DEVICE = 1
} iclblasMemoryMode_t;

typedef enum {
    ICLBLAS_MATH_FP32 = 0,
    ICLBLAS_MATH_SPLIT_BF16X3 = 1
} iclblasMathMode_t;

ICLBLAS_API iclblasStatus_t iclblasSetMemoryMode(iclblasHandle_t handle, iclblasMemoryMode_t mode);
ICLBLAS_API iclblasStatus_t iclblasGetMemoryMode(iclblasHandle_t handle, iclblasMemoryMode_t* mode);
ICLBLAS_API iclblasStatus_t iclblasSetMathMode(iclblasHandle_t handle, iclblasMathMode_t mode);
ICLBLAS_API iclblasStatus_t iclblasGetMathMode(iclblasHandle_t handle, iclblasMathMode_t* mode);

/* Stream used for every launch of this handle; `stream` is a hipStream_t
 * (NULL = the null stream). The handle does not own a stream set this way. */
ICLBLAS_API iclblasStatus_t iclblasSetStream(iclblasHandle_t handle, void* stream);
ICLBLAS_API iclblasStatus_t iclblasGetStream(iclblasHandle_t handle, void** stream);
/* Waits for all work queued on the handle's stream. Returns
 * ICLBLAS_STATUS_EXECUTION_FAILED when a solve queued since the last
 * synchronising call failed on the device (the dataflow STRSM aborting a wait;
 * host-pointer calls report that themselves). */
ICLBLAS_API iclblasStatus_t iclblasSynchronize(iclblasHandle_t handle);

/* Forces the named implementation for subsequent calls of the function it
 * belongs to (NULL or "" restores automatic selection). A call whose
 * parameters the forced implementation does not accept returns
 * ICLBLAS_STATUS_NOT_SUPPORTED. */
ICLBLAS_API iclblasStatus_t iclblasSetImplementation(iclblasHandle_t handle, const char* name);
/* Copies the name of the implementation that ran last into buf (NUL-terminated). */
ICLBLAS_API iclblasStatus_t iclblasGetLastImplementation(iclblasHandle_t handle, char* buf, int len);
/* Writes the names of all GEMM and STRSM implementations, newline-separated and
 * NUL-terminated, into buf; INVALID_VALUE if len is too small. Needs no handle. */
ICLBLAS_API iclblasStatus_t iclblasGetImplementations(char* buf, int len);

/* STRSM with mixed-precision iterative refinement: X0 = FP32 solve; then for
 * each sweep R = alpha*B - op(A)*X in FP64 on the GPU, D = FP32 solve of R,
 * X += D. Same arguments and validation as iclblasStrsm plus
 * 0 <= max_sweeps <= 8. max_sweeps == 0 is the plain solve. The sweeps stop
 * early once a correction no longer changes X at FP32 resolution
 * (max |D| <= 2^-24 max |X|); that test waits for the handle's stream once per
 * sweep, also in the device-pointer memory mode. */
ICLBLAS_API iclblasStatus_t iclblasStrsmRefined(iclblasHandle_t handle, iclblasSideMode_t side,
                                                iclblasFillMode_t uplo, iclblasOperation_t trans,
                                                iclblasDiagType_t diag, int m, int n, const float* alpha,
                                                float* A, int lda, float* B, int ldb, int max_sweeps);
/* Number of refinement sweeps the handle's last STRSM call ran (0 after
 * iclblasStrsm, max_sweeps == 0 or alpha == 0). */
ICLBLAS_API iclblasStatus_t iclblasGetLastRefinementSweeps(iclblasHandle_t handle, int* sweeps);

/* Adjoint of C = alpha*op(A)*op(B) + beta*C for an upstream gradient dC (m x n):
 *   dA (shape of A) = alpha * dC * op(B)^T      (transa == N)
 *                   = alpha * op(B) * dC^T      (transa != N)
 *   dB (shape of B) = alpha * op(A)^T * dC      (transb == N)
 *                   = alpha * dC^T * op(A)      (transb != N)
 * dA or dB may be NULL to skip it. Both are overwritten. */
ICLBLAS_API iclblasStatus_t iclblasSgemmBackward(iclblasHandle_t handle, iclblasOperation_t transa,
                                                 iclblasOperation_t transb, int m, int n, int k,
                                                 const float* alpha, float* A, int lda, float* B, int ldb,
                                                 float* dC, int lddc, float* dA, int ldda, float* dB,
                                                 int lddb);

/* Adjoint of X = alpha * op(A)^-1 * B (left) or X = alpha * B * op(A)^-1 (right),
 * given the forward output X (m x n) and the upstream gradient dX (m x n):
 *   G  = op(A)^-T * dX   (left)   or   dX * op(A)^-T   (right)
 *   dB = alpha * G
 *   dA = tri(-G * X^T) (left)  or  tri(-X^T * G) (right), transposed when
 *        trans != N; tri() keeps the stored triangle of A (uplo) and drops
 *        the diagonal when diag == UNIT. Entries of dA outside the triangle
 *        are set to 0.
 * dA or dB may be NULL to skip it. */
ICLBLAS_API iclblasStatus_t iclblasStrsmBackward(iclblasHandle_t handle, iclblasSideMode_t side,
                                                 iclblasFillMode_t uplo, iclblasOperation_t trans,
                                                 iclblasDiagType_t diag, int m, int n, const float* alpha,
                                                 float* A, int lda, float* X, int ldx, float* dX, int lddx,
                                                 float* dA, int ldda, float* dB, int lddb);

/* Kernel timing (a measurement aid). While enabled, every GEMM the handle runs
 * records a pair of HIP events on the handle's stream around its main kernel
 * alone. The operand pack of the split-BF16 kernels runs outside the pair,
 * except, for products of at least 1024 tiles, the half of it that runs on the
 * handle's second stream beside the first partial product (the pair spans the
 * three partial products and the wait for that half).
 * iclblasSetKernelTiming waits for the handle's stream and starts a new sum.
 * iclblasGetKernelTiming waits for the stream, returns the summed duration of
 * the recorded pairs and their number, and starts a new sum; at most 4096
 * launches are timed between two reads. The events carry no system-scope fence
 * (no cache write-back or invalidation between the kernels they bracket), so
 * enabling the timing leaves the timed calls' duration unchanged. */
ICLBLAS_API iclblasStatus_t iclblasSetKernelTiming(iclblasHandle_t handle, int enable);
ICLBLAS_API iclblasStatus_t iclblasGetKernelTiming(iclblasHandle_t handle, float* total_ms, int* launches);

/* Caps the device memory held by the handle's own buffers (the grow-only
 * scratch that stages host matrices in the host-pointer mode and holds the
 * workspaces: packed split-BF16 operands, TRSM and refinement temporaries,
 * level-3 expansions, GEMV partial sums), in bytes; 0 (the default) removes
 * the cap. A call that would grow them beyond it returns
 * ICLBLAS_STATUS_ALLOC_FAILED, as when the device runs out of memory; buffers
 * already held are kept. Lowering the cap frees nothing. */
ICLBLAS_API iclblasStatus_t iclblasSetScratchLimit(iclblasHandle_t handle, size_t bytes);
/* Bytes the handle's buffers hold now, and the cap (0 = none). */
ICLBLAS_API iclblasStatus_t iclblasGetScratchUsage(iclblasHandle_t handle, size_t* held, size_t* limit);

/* Bytes of the library's statically allocated device scratch, per GPU. */
#define ICLBLAS_NOALLOC_SCRATCH_BYTES (64u << 20)

/* SGEMM that allocates no memory. No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Same arguments, validation and
 * result as iclblasSgemm, except that A, B and C are always device pointers
 * (whatever the memory mode) and the call is queued on the handle's stream and
 * returns without waiting. It calls no hipMalloc, hipFree or other allocating
 * routine: in the split-BF16 math mode the packed operands go to a device
 * array of ICLBLAS_NOALLOC_SCRATCH_BYTES that the library holds on each GPU
 * from its first iclblasCreate there until the process ends. A call whose
 * packed operands need more than that, (M + N) * K * 4 bytes with M and N
 * rounded up to the kernel tile (256) and K to 32, returns
 * ICLBLAS_STATUS_INVALID_VALUE and queues nothing; the FP32 math mode and
 * products shallower than 256 use no scratch and have no such limit.
 * ICLBLAS_STATUS_NOT_INITIALIZED: the scratch could not be set up on the
 * handle's GPU. Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (other streams, other handles) race and must be ordered
 * by the caller. */
ICLBLAS_API iclblasStatus_t iclblasSgemmNoAlloc(iclblasHandle_t handle, iclblasOperation_t transa,
                                                iclblasOperation_t transb, int m, int n, int k,
                                                const float* alpha, float* A, int lda, float* B, int ldb,
                                                const float* beta, float* C, int ldc);

#ifdef __cplusplus
} /* extern "C" */
#endif
