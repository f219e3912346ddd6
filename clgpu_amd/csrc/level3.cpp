// GEMM-based level-3 routines: TRMM, SYMM / HEMM, SYRK / HERK, SYR2K / HER2K.
//
// Each runs in a single GEMM launch per product, on the FP32 MFMA kernels or,
// in the split-BF16 math mode when the product is at least the split's minimum
// depth, on the split-BF16 kernels (both implement the options below):
//  - SYRK-type updates write one triangle of C. Tiles wholly outside it exit
//    before loading anything and the epilogue leaves the other triangle
//    untouched (icl_gemm_args.write_tri); HERK-type updates also store the
//    diagonal's imaginary parts as 0.
//  - SYMM / HEMM first write the dense symmetric (hermitian) matrix the stored
//    triangle stands for into a handle workspace (an O(order^2) HBM-bound pass,
//    small next to the 2 order^2 n flops) and then run a plain GEMM.
//  - TRMM writes op(A) densely with zeros outside the triangle, then runs a GEMM
//    whose tiles each walk only the k range where op(A) can be non-zero
//    (icl_gemm_args.kband), so the zero half of the flops is skipped except on
//    the diagonal tiles.
//
// Semantics follow the reference's header contract (include/iclBLAS/iclBLAS.h:
// Ssymm :1555-1579, Ssyrk :1581-1608, Ssyr2k :1610-1645, Strmm :1721-1753,
// Csymm :1801-1825, Csyr2k :1827-1862, Csyrk :1864-1891, Cherk :1927-1954,
// Cher2k :1956-1991, Ctrmm :1993-2025, Chemm :2027-2051).
#include <cstring>

#include "blas3.hpp"
#include "handle.hpp"
#include "kernel_abi.h"

namespace icl {

namespace {

bool is_zero(bool cplx, const float v[2]) { return v[0] == 0.0f && (!cplx || v[1] == 0.0f); }

int kind_of(bool cplx) { return cplx ? ICL_KIND_CGEMM : ICL_KIND_SGEMM; }

size_t elem_bytes(bool cplx) { return cplx ? 2 * sizeof(float) : sizeof(float); }

// Dense order x order copy of a triangular or symmetric operand, in the handle's
// expansion workspace (leading dimension = order).
float* expand(iclblasContext* h, bool cplx, const icl_expand_args& base)
{
    const size_t bytes = (size_t)base.order * (size_t)base.order * elem_bytes(cplx);
    float* W = static_cast<float*>(h->scratch[iclblasContext::kSlotExpanded].get(bytes, h->stream));
    icl_expand_args a = base;
    a.W = W;
    a.ldw = base.order;
    hip_check(cplx ? icl_expand_complex(&a, h->stream) : icl_expand_real(&a, h->stream), "expand");
    return W;
}

bool overlaps(const float* x, long ldx, const float* y, long ldy, int m, int n, size_t es)
{
    const char* xb = reinterpret_cast<const char*>(x);
    const char* yb = reinterpret_cast<const char*>(y);
    const size_t xe = ((size_t)(n - 1) * ldx + m) * es, ye = ((size_t)(n - 1) * ldy + m) * es;
    return xb < yb + ye && yb < xb + xe;
}

} // namespace

void trmm_dev(iclblasContext* h, bool cplx, int side, int uplo, int trans, int diag, int m, int n,
              const float alpha[2], const float* A, long lda, const float* B, long ldb, float* C, long ldc)
{
    if (m <= 0 || n <= 0) return;
    if (is_zero(cplx, alpha)) {
        // C = 0; A and B are not read.
        hip_check(cplx ? icl_caxpby_2d(m, n, 0.0f, 0.0f, nullptr, 0, 0.0f, 0.0f, C, ldc, h->stream)
                       : icl_axpby_2d(m, n, 0.0f, nullptr, 0, 0.0f, C, ldc, h->stream),
                  "axpby");
        h->last_impl = "trmm_zero";
        return;
    }
    const bool left = side == ICLBLAS_SIDE_LEFT;
    const int order = left ? m : n;
    const bool lower = uplo == ICLBLAS_FILL_MODE_LOWER;
    const bool op_lower = lower == (trans == ICLBLAS_OP_N); // op(A) lower triangular
    icl_expand_args e{};
    e.order = order;
    e.A = A;
    e.lda = lda;
    e.mode = ICL_EXPAND_TRI;
    e.lower = lower ? 1 : 0;
    e.trans = trans;
    e.unit = diag == ICLBLAS_DIAG_UNIT ? 1 : 0;
    const float* W = expand(h, cplx, e);

    const size_t es = elem_bytes(cplx);
    float* out = C;
    long ldo = ldc;
    const bool alias = overlaps(B, ldb, C, ldc, m, n, es);
    if (alias) {
        out = static_cast<float*>(h->scratch[iclblasContext::kSlotLevel3Out].get((size_t)m * n * es, h->stream));
        ldo = m;
    }
    const float zero[2] = {0.0f, 0.0f};
    if (left)
        gemm_level3_dev(h, kind_of(cplx), 0, 0, m, n, m, alpha, W, order, B, ldb, zero, out, ldo, ICL_TRI_NONE,
                        op_lower ? ICL_KBAND_K_BELOW_M_END : ICL_KBAND_K_FROM_M0, 0);
    else
        gemm_level3_dev(h, kind_of(cplx), 0, 0, m, n, n, alpha, B, ldb, W, order, zero, out, ldo, ICL_TRI_NONE,
                        op_lower ? ICL_KBAND_K_FROM_N0 : ICL_KBAND_K_BELOW_N_END, 0);
    if (alias)
        hip_check(hipMemcpy2DAsync(C, (size_t)ldc * es, out, (size_t)m * es, (size_t)m * es, (size_t)n,
                                   hipMemcpyDeviceToDevice, h->stream),
                  "hipMemcpy2DAsync");
}

void symm_dev(iclblasContext* h, bool cplx, bool herm, int side, int uplo, int m, int n, const float alpha[2],
              const float* A, long lda, const float* B, long ldb, const float beta[2], float* C, long ldc)
{
    if (m <= 0 || n <= 0) return;
    const bool left = side == ICLBLAS_SIDE_LEFT;
    if (is_zero(cplx, alpha)) {
        // C = beta * C; A and B are not read.
        gemm_level3_dev(h, kind_of(cplx), 0, 0, m, n, 0, alpha, nullptr, 1, nullptr, 1, beta, C, ldc, ICL_TRI_NONE,
                        ICL_KBAND_NONE, 0);
        return;
    }
    const int order = left ? m : n;
    icl_expand_args e{};
    e.order = order;
    e.A = A;
    e.lda = lda;
    e.mode = ICL_EXPAND_SYM;
    e.lower = uplo == ICLBLAS_FILL_MODE_LOWER ? 1 : 0;
    e.herm = herm ? 1 : 0;
    const float* W = expand(h, cplx, e);
    if (left)
        gemm_level3_dev(h, kind_of(cplx), 0, 0, m, n, m, alpha, W, order, B, ldb, beta, C, ldc, ICL_TRI_NONE,
                        ICL_KBAND_NONE, 0);
    else
        gemm_level3_dev(h, kind_of(cplx), 0, 0, m, n, n, alpha, B, ldb, W, order, beta, C, ldc, ICL_TRI_NONE,
                        ICL_KBAND_NONE, 0);
}

namespace {

// GEMM ops of the rank-k product op(X) op(Y)^T (or ^H): with op(X) n x k,
// trans == N reads X as n x k (op N) and Y^T / Y^H through op T / C; otherwise
// X is k x n and op(X) = X^T (or X^H).
struct RankOps {
    int opa, opb, cflags;
};

RankOps rank_ops(bool cplx, bool herm, int trans)
{
    RankOps r{0, 0, 0};
    if (herm) {
        r.opa = trans == ICLBLAS_OP_N ? ICLBLAS_OP_N : ICLBLAS_OP_C;
        r.opb = trans == ICLBLAS_OP_N ? ICLBLAS_OP_C : ICLBLAS_OP_N;
        r.cflags = ICL_CFLAG_REAL_DIAG;
    } else {
        r.opa = trans == ICLBLAS_OP_N ? ICLBLAS_OP_N : ICLBLAS_OP_T;
        r.opb = trans == ICLBLAS_OP_N ? ICLBLAS_OP_T : ICLBLAS_OP_N;
        // Complex symmetric with op(A) = A^H: op(A) op(A)^T = conj(A^T A).
        if (cplx && trans == ICLBLAS_OP_C) r.cflags = ICL_CFLAG_CONJ_ACC;
    }
    return r;
}

} // namespace

void syrk_dev(iclblasContext* h, bool cplx, bool herm, int uplo, int trans, int n, int k, const float alpha[2],
              const float* A, long lda, const float beta[2], float* C, long ldc)
{
    if (n <= 0) return;
    const int tri = uplo == ICLBLAS_FILL_MODE_LOWER ? ICL_TRI_LOWER : ICL_TRI_UPPER;
    const RankOps r = rank_ops(cplx, herm, trans);
    const int kk = is_zero(cplx, alpha) ? 0 : k; // alpha == 0: A is not read
    gemm_level3_dev(h, kind_of(cplx), r.opa, r.opb, n, n, kk, alpha, A, lda, A, lda, beta, C, ldc, tri,
                    ICL_KBAND_NONE, r.cflags);
}

void syr2k_dev(iclblasContext* h, bool cplx, bool herm, int uplo, int trans, int n, int k, const float alpha[2],
               const float* A, long lda, const float* B, long ldb, const float beta[2], float* C, long ldc)
{
    if (n <= 0) return;
    const int tri = uplo == ICLBLAS_FILL_MODE_LOWER ? ICL_TRI_LOWER : ICL_TRI_UPPER;
    const RankOps r = rank_ops(cplx, herm, trans);
    if (k == 0 || is_zero(cplx, alpha)) {
        gemm_level3_dev(h, kind_of(cplx), r.opa, r.opb, n, n, 0, alpha, nullptr, 1, nullptr, 1, beta, C, ldc, tri,
                        ICL_KBAND_NONE, r.cflags);
        return;
    }
    // First product with beta, second accumulated on top (beta = 1).
    gemm_level3_dev(h, kind_of(cplx), r.opa, r.opb, n, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tri, ICL_KBAND_NONE,
                    r.cflags);
    const float one[2] = {1.0f, 0.0f};
    const float alpha2[2] = {alpha[0], herm ? -alpha[1] : alpha[1]};
    gemm_level3_dev(h, kind_of(cplx), r.opa, r.opb, n, n, k, alpha2, B, ldb, A, lda, one, C, ldc, tri, ICL_KBAND_NONE,
                    r.cflags);
}

} // namespace icl
