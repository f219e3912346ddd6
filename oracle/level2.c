/*
 * level2.c -- FP64 oracle for GEMV. TEST INFRASTRUCTURE ONLY (see oracle.c).
 *
 * orc_zgemv restates Sgemv_naive.cl:22-59 and Cgemv_naive.cl:16-50 in complex
 * double (a real call is the case of zero imaginary parts): for each element of
 * y, in the kernels' loop order, v = sum over the reduction index of
 * op(A) element times x element (op C conjugates A for complex data), then
 * y = alpha * v + beta * y; y is not read when beta == 0 (the header's
 * contract, iclBLAS.h:1043 "if beta == 0, y does not have to be initialized").
 * mag (optional) receives sum |op(A)||x| per element of y for the error bound.
 * Vectors are strided by incx / incy (>= 1); y and out are interleaved complex.
 */
#include <math.h>
#include <stddef.h>

void orc_zgemv(int trans, int m, int n, const double* alpha, const double* A, int lda, const double* x, int incx,
               const double* beta, const double* y, int incy, double* out, double* mag)
{
    const int leny = trans == 0 ? m : n, lenr = trans == 0 ? n : m;
    for (int i = 0; i < leny; ++i) {
        double vr = 0.0, vi = 0.0, g = 0.0;
        for (int l = 0; l < lenr; ++l) {
            const long ai = trans == 0 ? (long)l * lda + i : (long)i * lda + l;
            double ar = A[2 * ai], aim = A[2 * ai + 1];
            if (trans == 2) aim = -aim;
            const double xr = x[2 * (long)l * incx], xi = x[2 * (long)l * incx + 1];
            vr += ar * xr - aim * xi;
            vi += ar * xi + aim * xr;
            g += hypot(ar, aim) * hypot(xr, xi);
        }
        double o_r = alpha[0] * vr - alpha[1] * vi, o_i = alpha[0] * vi + alpha[1] * vr;
        if (beta[0] != 0.0 || beta[1] != 0.0) {
            const double yr = y[2 * (long)i * incy], yi = y[2 * (long)i * incy + 1];
            o_r += beta[0] * yr - beta[1] * yi;
            o_i += beta[0] * yi + beta[1] * yr;
        }
        out[2 * i] = o_r;
        out[2 * i + 1] = o_i;
        if (mag) mag[i] = g;
    }
}
