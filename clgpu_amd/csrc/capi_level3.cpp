// C entry points of the GEMM-based level-3 routines: iclblasSsymm, Ssyrk,
// Ssyr2k, Strmm, Csymm, Csyrk, Csyr2k, Chemm, Cherk, Cher2k, Ctrmm.
//
// Argument checks follow each reference entry point's own order, then the
// generic chain of iclBLASImpl.hpp:108-233 (n, m, k sign and zero tests, null
// matrices and non-positive leading dimensions -> INVALID_VALUE, zero sizes ->
// SUCCESS), and finally the handle:
//   Ssymm / Csymm / Chemm (iclBLAS_S.cpp:387-397, iclBLAS_C.cpp:472-483, 578-589):
//     m == 0 or n == 0 or (alpha == 0 and beta == 1) -> SUCCESS first.
//   Ssyrk / Ssyr2k / Csyrk / Csyr2k (iclBLAS_S.cpp:399-420, iclBLAS_C.cpp:485-509):
//     n == 0 or k == 0 or (alpha == 0 and beta == 1) -> SUCCESS first (C is then
//     not scaled by beta, as in the reference).
//   Cherk / Cher2k (iclBLAS_C.cpp:539-561): n < 0 or k < 0 -> INVALID_VALUE first;
//     k == 0 still runs (C = beta * C on the triangle).
//   Strmm (iclBLAS_S.cpp:514-520): the generic chain only.
//   Ctrmm (iclBLAS_C.cpp:566-575): m <= 0, n <= 0 or uplo outside {0, 1} ->
//     SUCCESS first.
// As in capi.cpp, checks the reference does not make (null alpha / beta, enum
// values out of range, leading dimensions below the rows they describe) run
// after the reference's own and return INVALID_VALUE. op C on the real routines
// means op T, and op T on Cherk / Cher2k means op C, as in the reference kernels,
// which only distinguish op N from the rest (Ssyrk_naive.cl:21, Cherk_naive.cl:18).
#include <cstring>

#include "capi_common.hpp"

using namespace icl::capi;

namespace {

typedef float cf[2];

bool side_ok(int s) { return s == ICLBLAS_SIDE_LEFT || s == ICLBLAS_SIDE_RIGHT; }
bool uplo_ok(int u) { return u == ICLBLAS_FILL_MODE_UPPER || u == ICLBLAS_FILL_MODE_LOWER; }
bool diag_ok(int d) { return d == ICLBLAS_DIAG_NON_UNIT || d == ICLBLAS_DIAG_UNIT; }

// A scalar argument as {re, im} (im = 0 for real routines).
void scalar(const float* p, bool cplx, float out[2])
{
    out[0] = p[0];
    out[1] = cplx ? p[1] : 0.0f;
}
bool is_zero(const float v[2]) { return v[0] == 0.0f && v[1] == 0.0f; }
bool is_one(const float v[2]) { return v[0] == 1.0f && v[1] == 0.0f; }

float* stage(iclblasContext* h, bool cplx, int slot, const float* host, int rows, int cols, int ld, bool copy)
{
    if (cplx)
        return reinterpret_cast<float*>(stage_in<cf>(h, slot, reinterpret_cast<const cf*>(host), rows, cols, ld, copy));
    return stage_in<float>(h, slot, host, rows, cols, ld, copy);
}

void unstage(iclblasContext* h, bool cplx, float* host, const float* dev, int rows, int cols, int ld)
{
    if (cplx) stage_out<cf>(h, reinterpret_cast<cf*>(host), reinterpret_cast<const cf*>(dev), rows, cols, ld);
    else stage_out<float>(h, host, dev, rows, cols, ld);
}

int at_least_1(int v) { return v > 1 ? v : 1; }

// ---- SYMM / HEMM ----
iclblasStatus_t symm_common(iclblasHandle_t handle, bool cplx, bool herm, int side, int uplo, int m, int n,
                            const float* alpha, const float* A, int lda, const float* B, int ldb, const float* beta,
                            float* C, int ldc)
{
    if (m == 0 || n == 0) return ICLBLAS_STATUS_SUCCESS;
    if (!alpha || !beta) return ICLBLAS_STATUS_INVALID_VALUE;
    float a[2], b[2];
    scalar(alpha, cplx, a);
    scalar(beta, cplx, b);
    if (is_zero(a) && is_one(b)) return ICLBLAS_STATUS_SUCCESS;
    if (n < 0 || m < 0) return ICLBLAS_STATUS_INVALID_VALUE;
    if (!A || lda <= 0 || !B || ldb <= 0 || !C || ldc <= 0) return ICLBLAS_STATUS_INVALID_VALUE;
    if (!side_ok(side) || !uplo_ok(uplo)) return ICLBLAS_STATUS_INVALID_VALUE;
    const int order = side == ICLBLAS_SIDE_LEFT ? m : n;
    if (lda < order || ldb < m || ldc < m) return ICLBLAS_STATUS_INVALID_VALUE;
    if (!iclblasContext::valid(handle)) return ICLBLAS_STATUS_INVALID_VALUE;
    return guarded([&] {
        HandleLock lock(handle);
        iclblasContext* h = handle;
        if (h->memory_mode == ICLBLAS_MEMORY_DEVICE) {
            icl::symm_dev(h, cplx, herm, side, uplo, m, n, a, A, lda, B, ldb, b, C, ldc);
        } else {
            const bool use_ab = !is_zero(a);
            const float* dA = use_ab ? stage(h, cplx, kSlotA, A, order, order, lda, true) : nullptr;
            const float* dB = use_ab ? stage(h, cplx, kSlotB, B, m, n, ldb, true) : nullptr;
            float* dC = stage(h, cplx, kSlotC, C, m, n, ldc, !is_zero(b));
            icl::symm_dev(h, cplx, herm, side, uplo, m, n, a, dA, lda, dB, ldb, b, dC, ldc);
            unstage(h, cplx, C, dC, m, n, ldc);
        }
        finish(h);
    });
}

// ---- SYRK / SYR2K / HERK / HER2K ----
// The reference-specific early exits have run; this is the generic chain and the
// call. B == nullptr with two == false for the rank-k forms.
iclblasStatus_t rank_common(iclblasHandle_t handle, bool cplx, bool herm, bool two, int uplo, int trans, int n, int k,
                            const float a[2], const float* A, int lda, const float* B, int ldb, const float b[2],
                            float* C, int ldc)
{
    if (n < 0 || k < 0) return ICLBLAS_STATUS_INVALID_VALUE;
    if (n == 0) return ICLBLAS_STATUS_SUCCESS;
    if (!A || lda <= 0 || (two && (!B || ldb <= 0)) || !C || ldc <= 0) return ICLBLAS_STATUS_INVALID_VALUE;
    if (!uplo_ok(uplo) || !op_ok(trans)) return ICLBLAS_STATUS_INVALID_VALUE;
    const int rows = trans == ICLBLAS_OP_N ? n : k, cols = trans == ICLBLAS_OP_N ? k : n;
    if (lda < at_least_1(rows) || (two && ldb < at_least_1(rows)) || ldc < n) return ICLBLAS_STATUS_INVALID_VALUE;
    if (!iclblasContext::valid(handle)) return ICLBLAS_STATUS_INVALID_VALUE;
    return guarded([&] {
        HandleLock lock(handle);
        iclblasContext* h = handle;
        const float* dA = A;
        const float* dB = B;
        float* dC = C;
        const bool host = h->memory_mode == ICLBLAS_MEMORY_HOST;
        if (host) {
            const bool use_ab = k > 0 && !is_zero(a);
            dA = use_ab ? stage(h, cplx, kSlotA, A, rows, cols, lda, true) : nullptr;
            dB = use_ab && two ? stage(h, cplx, kSlotB, B, rows, cols, ldb, true) : nullptr;
            // The whole of C goes back to the caller, so the triangle that is
            // not updated is staged in as well.
            dC = stage(h, cplx, kSlotC, C, n, n, ldc, true);
        }
        if (two) icl::syr2k_dev(h, cplx, herm, uplo, trans, n, k, a, dA, lda, dB, ldb, b, dC, ldc);
        else icl::syrk_dev(h, cplx, herm, uplo, trans, n, k, a, dA, lda, b, dC, ldc);
        if (host) unstage(h, cplx, C, dC, n, n, ldc);
        finish(h);
    });
}

// Ssyrk / Ssyr2k / Csyrk / Csyr2k front: n == 0 or k == 0 or (alpha == 0 and
// beta == 1) -> SUCCESS before anything else.
iclblasStatus_t sym_rank_front(iclblasHandle_t handle, bool cplx, bool two, int uplo, int trans, int n, int k,
                               const float* alpha, const float* A, int lda, const float* B, int ldb,
                               const float* beta, float* C, int ldc)
{
    if (n == 0 || k == 0) return ICLBLAS_STATUS_SUCCESS;
    if (!alpha || !beta) return ICLBLAS_STATUS_INVALID_VALUE;
    float a[2], b[2];
    scalar(alpha, cplx, a);
    scalar(beta, cplx, b);
    if (is_zero(a) && is_one(b)) return ICLBLAS_STATUS_SUCCESS;
    return rank_common(handle, cplx, false, two, uplo, trans, n, k, a, A, lda, B, ldb, b, C, ldc);
}

// ---- TRMM ----
iclblasStatus_t trmm_common(iclblasHandle_t handle, bool cplx, int side, int uplo, int trans, int diag, int m, int n,
                            const float* alpha, const float* A, int lda, const float* B, int ldb, float* C, int ldc)
{
    if (n < 0) return ICLBLAS_STATUS_INVALID_VALUE;
    if (n == 0) return ICLBLAS_STATUS_SUCCESS;
    if (m < 0) return ICLBLAS_STATUS_INVALID_VALUE;
    if (m == 0) return ICLBLAS_STATUS_SUCCESS;
    if (!A || lda <= 0 || !B || ldb <= 0 || !C || ldc <= 0) return ICLBLAS_STATUS_INVALID_VALUE;
    if (!alpha) return ICLBLAS_STATUS_INVALID_VALUE;
    if (!side_ok(side) || !uplo_ok(uplo) || !op_ok(trans) || !diag_ok(diag)) return ICLBLAS_STATUS_INVALID_VALUE;
    const int order = side == ICLBLAS_SIDE_LEFT ? m : n;
    if (lda < order || ldb < m || ldc < m) return ICLBLAS_STATUS_INVALID_VALUE;
    if (!iclblasContext::valid(handle)) return ICLBLAS_STATUS_INVALID_VALUE;
    float a[2];
    scalar(alpha, cplx, a);
    return guarded([&] {
        HandleLock lock(handle);
        iclblasContext* h = handle;
        if (h->memory_mode == ICLBLAS_MEMORY_DEVICE) {
            icl::trmm_dev(h, cplx, side, uplo, trans, diag, m, n, a, A, lda, B, ldb, C, ldc);
        } else {
            // A and B are staged into their own buffers, so C may be the same
            // host matrix as B (the in-place use of the reference's tests).
            const bool use_ab = !is_zero(a);
            const float* dA = use_ab ? stage(h, cplx, kSlotA, A, order, order, lda, true) : nullptr;
            const float* dB = use_ab ? stage(h, cplx, kSlotB, B, m, n, ldb, true) : nullptr;
            float* dC = stage(h, cplx, kSlotC, C, m, n, ldc, false);
            icl::trmm_dev(h, cplx, side, uplo, trans, diag, m, n, a, dA, lda, dB, ldb, dC, ldc);
            unstage(h, cplx, C, dC, m, n, ldc);
        }
        finish(h);
    });
}

const float* F(const void* p) { return static_cast<const float*>(p); }
float* F(void* p) { return static_cast<float*>(p); }

} // namespace

extern "C" {

// reference: libs/blas/shlib/iclBLAS_S.cpp:387-397
ICLBLAS_API iclblasStatus_t iclblasSsymm(iclblasHandle_t handle, iclblasSideMode_t side_, iclblasFillMode_t uplo_, int m,
                                         int n, const float* alpha, float* A, int lda, float* B, int ldb,
                                         const float* beta, float* C, int ldc)
{
    const int side = enum_arg(side_);
    const int uplo = enum_arg(uplo_);
    return symm_common(handle, false, false, side, uplo, m, n, alpha, A, lda, B, ldb, beta, C, ldc);
}

// reference: libs/blas/shlib/iclBLAS_S.cpp:399-408
ICLBLAS_API iclblasStatus_t iclblasSsyrk(iclblasHandle_t handle, iclblasFillMode_t uplo_, iclblasOperation_t trans_,
                                         int n, int k, const float* alpha, float* A, int lda, const float* beta,
                                         float* C, int ldc)
{
    const int trans = enum_arg(trans_);
    const int uplo = enum_arg(uplo_);
    return sym_rank_front(handle, false, false, uplo, trans, n, k, alpha, A, lda, nullptr, 0, beta, C, ldc);
}

// reference: libs/blas/shlib/iclBLAS_S.cpp:410-420
ICLBLAS_API iclblasStatus_t iclblasSsyr2k(iclblasHandle_t handle, iclblasFillMode_t uplo_, iclblasOperation_t trans_,
                                          int n, int k, const float* alpha, float* A, int lda, float* B, int ldb,
                                          const float* beta, float* C, int ldc)
{
    const int trans = enum_arg(trans_);
    const int uplo = enum_arg(uplo_);
    return sym_rank_front(handle, false, true, uplo, trans, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

// reference: libs/blas/shlib/iclBLAS_S.cpp:512-520
ICLBLAS_API iclblasStatus_t iclblasStrmm(iclblasHandle_t handle, iclblasSideMode_t side_, iclblasFillMode_t uplo_,
                                         iclblasOperation_t transa_, iclblasDiagType_t diag_, int m, int n,
                                         const float* alpha, float* A, int lda, float* B, int ldb, float* C, int ldc)
{
    const int transa = enum_arg(transa_);
    const int side = enum_arg(side_);
    const int uplo = enum_arg(uplo_);
    const int diag = enum_arg(diag_);
    return trmm_common(handle, false, side, uplo, transa, diag, m, n, alpha, A, lda, B, ldb, C, ldc);
}

// reference: libs/blas/shlib/iclBLAS_C.cpp:472-483
ICLBLAS_API iclblasStatus_t iclblasCsymm(iclblasHandle_t handle, iclblasSideMode_t side_, iclblasFillMode_t uplo_, int m,
                                         int n, const oclComplex_t* alpha, oclComplex_t* A, int lda, oclComplex_t* B,
                                         int ldb, const oclComplex_t* beta, oclComplex_t* C, int ldc)
{
    const int side = enum_arg(side_);
    const int uplo = enum_arg(uplo_);
    return symm_common(handle, true, false, side, uplo, m, n, F(alpha), F(A), lda, F(B), ldb, F(beta), F(C), ldc);
}

// reference: libs/blas/shlib/iclBLAS_C.cpp:578-589
ICLBLAS_API iclblasStatus_t iclblasChemm(iclblasHandle_t handle, iclblasSideMode_t side_, iclblasFillMode_t uplo_, int m,
                                         int n, const oclComplex_t* alpha, oclComplex_t* A, int lda, oclComplex_t* B,
                                         int ldb, const oclComplex_t* beta, oclComplex_t* C, int ldc)
{
    const int side = enum_arg(side_);
    const int uplo = enum_arg(uplo_);
    return symm_common(handle, true, true, side, uplo, m, n, F(alpha), F(A), lda, F(B), ldb, F(beta), F(C), ldc);
}

// reference: libs/blas/shlib/iclBLAS_C.cpp:498-509
ICLBLAS_API iclblasStatus_t iclblasCsyrk(iclblasHandle_t handle, iclblasFillMode_t uplo_, iclblasOperation_t trans_,
                                         int n, int k, const oclComplex_t* alpha, oclComplex_t* A, int lda,
                                         const oclComplex_t* beta, oclComplex_t* C, int ldc)
{
    const int trans = enum_arg(trans_);
    const int uplo = enum_arg(uplo_);
    return sym_rank_front(handle, true, false, uplo, trans, n, k, F(alpha), F(A), lda, nullptr, 0, F(beta), F(C), ldc);
}

// reference: libs/blas/shlib/iclBLAS_C.cpp:485-496
ICLBLAS_API iclblasStatus_t iclblasCsyr2k(iclblasHandle_t handle, iclblasFillMode_t uplo_, iclblasOperation_t trans_,
                                          int n, int k, const oclComplex_t* alpha, oclComplex_t* A, int lda,
                                          oclComplex_t* B, int ldb, const oclComplex_t* beta, oclComplex_t* C, int ldc)
{
    const int trans = enum_arg(trans_);
    const int uplo = enum_arg(uplo_);
    return sym_rank_front(handle, true, true, uplo, trans, n, k, F(alpha), F(A), lda, F(B), ldb, F(beta), F(C), ldc);
}

// reference: libs/blas/shlib/iclBLAS_C.cpp:537-548 (alpha and beta are real)
ICLBLAS_API iclblasStatus_t iclblasCherk(iclblasHandle_t handle, iclblasFillMode_t uplo_, iclblasOperation_t trans_,
                                         int n, int k, const float* alpha, oclComplex_t* A, int lda, const float* beta,
                                         oclComplex_t* C, int ldc)
{
    const int trans = enum_arg(trans_);
    const int uplo = enum_arg(uplo_);
    if (n < 0 || k < 0) return ICLBLAS_STATUS_INVALID_VALUE;
    if (n == 0) return ICLBLAS_STATUS_SUCCESS;
    if (!alpha || !beta) return ICLBLAS_STATUS_INVALID_VALUE;
    const float a[2] = {*alpha, 0.0f}, b[2] = {*beta, 0.0f};
    return rank_common(handle, true, true, false, uplo, trans, n, k, a, F(A), lda, nullptr, 0, b, F(C), ldc);
}

// reference: libs/blas/shlib/iclBLAS_C.cpp:550-561 (beta is real)
ICLBLAS_API iclblasStatus_t iclblasCher2k(iclblasHandle_t handle, iclblasFillMode_t uplo_, iclblasOperation_t trans_,
                                          int n, int k, const oclComplex_t* alpha, oclComplex_t* A, int lda,
                                          oclComplex_t* B, int ldb, const float* beta, oclComplex_t* C, int ldc)
{
    const int trans = enum_arg(trans_);
    const int uplo = enum_arg(uplo_);
    if (n < 0 || k < 0) return ICLBLAS_STATUS_INVALID_VALUE;
    if (n == 0) return ICLBLAS_STATUS_SUCCESS;
    if (!alpha || !beta) return ICLBLAS_STATUS_INVALID_VALUE;
    float a[2];
    scalar(F(alpha), true, a);
    const float b[2] = {*beta, 0.0f};
    return rank_common(handle, true, true, true, uplo, trans, n, k, a, F(A), lda, F(B), ldb, b, F(C), ldc);
}

// reference: libs/blas/shlib/iclBLAS_C.cpp:564-575
ICLBLAS_API iclblasStatus_t iclblasCtrmm(iclblasHandle_t handle, iclblasSideMode_t side_, iclblasFillMode_t uplo_,
                                         iclblasOperation_t transa_, iclblasDiagType_t diag_, int m, int n,
                                         const oclComplex_t* alpha, oclComplex_t* A, int lda, oclComplex_t* B, int ldb,
                                         oclComplex_t* C, int ldc)
{
    const int transa = enum_arg(transa_);
    const int side = enum_arg(side_);
    const int uplo = enum_arg(uplo_);
    const int diag = enum_arg(diag_);
    if (n <= 0 || m <= 0 || !uplo_ok(uplo)) return ICLBLAS_STATUS_SUCCESS;
    return trmm_common(handle, true, side, uplo, transa, diag, m, n, F(alpha), F(A), lda, F(B), ldb, F(C), ldc);
}

} // extern "C"
