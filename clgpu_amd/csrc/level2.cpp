// GEMV (SGEMV, CGEMV) on the HBM-bound kernels of gemv.hip; semantics of the
// reference's header (include/iclBLAS/iclBLAS.h: Sgemv :1020-1049, Cgemv
// :1135-1164): y = alpha * op(A) * x + beta * y, op in {N, T, C}, C = T for real.
#include "blas3.hpp"
#include "handle.hpp"
#include "kernel_abi.h"

namespace icl {

void gemv_dev(iclblasContext* h, bool cplx, int trans, int m, int n, const float alpha[2], const float* A, long lda,
              const float* x, long incx, const float beta[2], float* y, long incy)
{
    if (m <= 0 || n <= 0) return;
    const int leny = trans == ICLBLAS_OP_N ? m : n;
    if (alpha[0] == 0.0f && (!cplx || alpha[1] == 0.0f)) {
        // y = beta * y; A and x are not read (y not read when beta == 0).
        hip_check(cplx ? icl_caxpby_2d(1, leny, 0.0f, 0.0f, nullptr, 0, beta[0], beta[1], y, incy, h->stream)
                       : icl_axpby_2d(1, leny, 0.0f, nullptr, 0, beta[0], y, incy, h->stream),
                  "axpby");
        h->last_impl = "scale_y";
        return;
    }
    icl_gemv_args a{};
    a.trans = cplx ? trans : (trans == ICLBLAS_OP_N ? 0 : 1);
    a.m = m;
    a.n = n;
    a.A = A;
    a.lda = lda;
    a.x = x;
    a.incx = incx;
    a.y = y;
    a.incy = incy;
    a.alpha_re = alpha[0];
    a.alpha_im = cplx ? alpha[1] : 0.0f;
    a.beta_re = beta[0];
    a.beta_im = cplx ? beta[1] : 0.0f;
    a.splits = icl_gemv_splits(cplx ? 1 : 0, a.trans, m, n);
    if (a.splits > 1) {
        const size_t bytes = (size_t)a.splits * (size_t)leny * (cplx ? 8 : 4);
        a.partial = static_cast<float*>(h->scratch[iclblasContext::kSlotGemvPartial].get(bytes, h->stream));
    }
    hip_check(cplx ? icl_cgemv(&a, h->stream) : icl_sgemv(&a, h->stream), cplx ? "cgemv" : "sgemv");
    h->last_impl = a.splits > 1 ? (cplx ? "cgemv_hbm_split" : "sgemv_hbm_split") : (cplx ? "cgemv_hbm" : "sgemv_hbm");
}

} // namespace icl
