// Device-level level-3 algorithms of the host layer. All pointers here are device
// pointers; work is queued on the handle's stream; errors are thrown as
// icl::status_error and turned into iclblasStatus_t at the C boundary.
#pragma once
#include "handle.hpp"
#include "kernel_abi.h"

#include <vector>

namespace icl {

// ---- implementation registry (reference: include/iclgpu/functions_base.hpp:149-223,
//      include/iclgpu/dispatcher.hpp:49-61) ----
struct GemmProblem {
    int kind; // ICL_KIND_*
    int math; // ICL_MATH_*
    int opa, opb;
    const icl_gemm_args* args;
};

// Every registered GEMM implementation, in registration order.
const std::vector<const icl_gemm_variant*>& gemm_registry();

// accept(): the implementation handles this kind and math mode.
bool gemm_accept(const icl_gemm_variant& v, const GemmProblem& p);
// score(): higher is better; see the definition for the model.
float gemm_score(const icl_gemm_variant& v, const GemmProblem& p);
// Whether the problem needs the bounds-checked instantiation of `v`.
bool gemm_needs_edge(const icl_gemm_variant& v, const GemmProblem& p);

// Picks the accepted implementation with the highest score (first registered wins
// ties), honouring the handle's forced implementation when it names a GEMM of
// this kind. Throws NOT_SUPPORTED when nothing accepts.
const icl_gemm_variant* gemm_select(iclblasContext* h, const GemmProblem& p);

// ---- algorithms ----
// C = alpha*op(A)*op(B) + beta*C (real). mask: ICL_MASK_* applied to the stored C.
void sgemm_dev(iclblasContext* h, int math, int opa, int opb, int m, int n, int k, float alpha, const float* A,
               long lda, const float* B, long ldb, float beta, float* C, long ldc, int mask = ICL_MASK_NONE,
               bool record = true);

void cgemm_dev(iclblasContext* h, int math, int opa, int opb, int m, int n, int k, const float alpha[2],
               const float* A, long lda, const float* B, long ldb, const float beta[2], float* C, long ldc);

// True when the handle runs a GEMM of depth k on the split-BF16 kernels: math
// mode SPLIT_BF16X3 and k at least the depth from which the split meets the
// FP32 error bound.
bool split_depth(iclblasContext* h, int k);

// One GEMM launch with the level-3 options of icl_gemm_args (write_tri, kband,
// cflags), on the split-BF16 kernels when split_depth(h, k) and on the FP32 MFMA
// kernels otherwise. kind: ICL_KIND_SGEMM (alpha[1], beta[1] unused) or
// ICL_KIND_CGEMM. k may be 0: C's selected triangle becomes beta * C and A, B are
// not read.
void gemm_level3_dev(iclblasContext* h, int kind, int opa, int opb, int m, int n, int k, const float alpha[2],
                     const float* A, long lda, const float* B, long ldb, const float beta[2], float* C, long ldc,
                     int write_tri, int kband, int cflags);

// GEMM-based level-3 routines on device pointers (complex data interleaved).
// TRMM: C = alpha * op(A) * B (left) or alpha * B * op(A) (right), out of place;
// C may alias B.
void trmm_dev(iclblasContext* h, bool cplx, int side, int uplo, int trans, int diag, int m, int n,
              const float alpha[2], const float* A, long lda, const float* B, long ldb, float* C, long ldc);
// SYMM / HEMM: C = alpha * A * B + beta * C (left) or alpha * B * A + beta * C.
void symm_dev(iclblasContext* h, bool cplx, bool herm, int side, int uplo, int m, int n, const float alpha[2],
              const float* A, long lda, const float* B, long ldb, const float beta[2], float* C, long ldc);
// SYRK / HERK (herm: op is the conjugate transpose, alpha and beta real):
// C = alpha * op(A) * op(A)^T|H + beta * C on the uplo triangle.
void syrk_dev(iclblasContext* h, bool cplx, bool herm, int uplo, int trans, int n, int k, const float alpha[2],
              const float* A, long lda, const float beta[2], float* C, long ldc);
// SYR2K / HER2K: C = alpha op(A) op(B)^T + alpha op(B) op(A)^T + beta C, or for
// herm alpha op(A) op(B)^H + conj(alpha) op(B) op(A)^H + beta C (beta real).
void syr2k_dev(iclblasContext* h, bool cplx, bool herm, int uplo, int trans, int n, int k, const float alpha[2],
               const float* A, long lda, const float* B, long ldb, const float beta[2], float* C, long ldc);

// GEMV: y = alpha * op(A) * x + beta * y on device pointers (complex data
// interleaved; incx, incy >= 1). y is not read when beta == 0; alpha == 0 scales
// y without reading A or x.
void gemv_dev(iclblasContext* h, bool cplx, int trans, int m, int n, const float alpha[2], const float* A, long lda,
              const float* x, long incx, const float beta[2], float* y, long incy);

// Blocked STRSM, in place on B. reuse_inverses: the previous STRSM on this
// handle was the same system (A, lda, side, uplo, trans, diag, m, n) and the
// handle's TRSM workspace has not been used since, so the panel solve takes
// the diagonal-block and panel inverses it left there instead of forming them
// again (the other solves ignore the flag).
void strsm_dev(iclblasContext* h, int side, int uplo, int trans, int diag, int m, int n, float alpha,
               const float* A, long lda, float* B, long ldb, bool record = true, bool reuse_inverses = false);

// Blocked CTRSM on interleaved complex data, in place on B.
void ctrsm_dev(iclblasContext* h, int side, int uplo, int trans, int diag, int m, int n, const float alpha[2],
               const float* A, long lda, float* B, long ldb);

// After the handle's stream is synchronised: when a dataflow STRSM was queued
// since the last check, reads its error word and throws EXECUTION_FAILED (and
// clears the word) when that solve aborted a wait. Host-pointer calls check at
// their end; in the device-pointer mode iclblasSynchronize reports it.
void trsm_check_error(iclblasContext* h);

// STRSM + FP64-residual iterative refinement. ws0/ws1: m*n floats each.
void strsm_refined_dev(iclblasContext* h, int side, int uplo, int trans, int diag, int m, int n, float alpha,
                       const float* A, long lda, float* B, long ldb, int sweeps, float* ws0, float* ws1);

void sgemm_backward_dev(iclblasContext* h, int transa, int transb, int m, int n, int k, float alpha, const float* A,
                        long lda, const float* B, long ldb, const float* dC, long lddc, float* dA, long ldda,
                        float* dB, long lddb);

// G workspace: m*n floats.
void strsm_backward_dev(iclblasContext* h, int side, int uplo, int trans, int diag, int m, int n, float alpha,
                        const float* A, long lda, const float* X, long ldx, const float* dX, long lddx, float* dA,
                        long ldda, float* dB, long lddb, float* G);

} // namespace icl
