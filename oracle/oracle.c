/*
 * oracle.c -- CPU oracle for the level-3 hot path. TEST INFRASTRUCTURE ONLY:
 * it is used by tests/, by __graft_entry__.smoke() as the checker and by
 * bench.py's CPU-baseline leg; the library never calls it.
 *
 * Two layers:
 *  1. Restatements of the reference's own CPU semantics, arithmetic order kept:
 *     - orc_sgemm / orc_cgemm: cpu_gemm<float> / cpu_gemm<std::complex<float>>
 *       (reference tests/blas_funcs/test_gemm.hpp:83-201): loops mi -> ni -> ki,
 *       accumulator in double (complex double) (test_helpers.hpp:186-205),
 *       conj for OP_C (test_gemm.hpp:33-54), result
 *       (float)(acc*alpha + C*beta) computed in the accumulator type, or
 *       (float)(acc*alpha) without reading C when beta == 0 (test_gemm.hpp:56-81).
 *       Index forms (test_gemm.hpp:20-30): N -> c*ld + r, otherwise r*ld + c.
 *     - orc_strsm: the substitution of libs/blas/funcs/Strsm/Strsm_naive.cl:18-126,
 *       all 8 side/uplo/trans variants, float arithmetic, alpha applied after
 *       the solve; a non-zero trans means transposed (OP_C == OP_T for real).
 *       The elimination `b -= a * x` is an FMA: OpenCL C contracts it by default
 *       (FP_CONTRACT ON), and the reference's own known answer
 *       3x4_right_up_trans_ndiag (tests/iclBLAS/test_Strsm.cpp:185-225) is only
 *       met within its 4-ULP tolerance with the fused form.
 *     - orc_ctrsm: the substitution of libs/blas/funcs/Ctrsm/Ctrsm_naive.cl:16-183
 *       with the fma-based cmul / cdiv of core/src/ocl/complex.h.cl:32-39,
 *       OP_C conjugating A, alpha applied after the solve.
 *  2. FP64 references used for error reporting: orc_dgemm / orc_zgemm (full
 *     double arithmetic, plus the |op(A)||op(B)| product for the elementwise
 *     bound), orc_dtrsm / orc_ztrsm (the same substitutions in double), sampled variants
 *     for large shapes, and a column-parallel orc_sgemm_mt for the CPU baseline.
 *
 * Compiled with -ffp-contract=off so that no product is fused implicitly; the
 * only FMAs are the explicit ones of the TRSM eliminations and complex helpers.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* element (r, c) of op(X) for a column-major X with leading dimension ld */
static inline long idx(int trans, int r, int c, int ld) { return trans == 0 ? (long)c * ld + r : (long)r * ld + c; }

/* ------------------------------------------------------------------------- */
/* Layer 1: restatement of cpu_gemm                                           */
/* ------------------------------------------------------------------------- */
static void sgemm_cols(int ta, int tb, int m, int n0, int n1, int k, float alpha, const float* A, int lda,
                       const float* B, int ldb, float beta, float* C, int ldc)
{
    const int beta_zero = (beta == 0.0f);
    for (int mi = 0; mi < m; ++mi) {
        for (int ni = n0; ni < n1; ++ni) {
            double acc = 0.0;
            for (int ki = 0; ki < k; ++ki)
                acc += (double)A[idx(ta, mi, ki, lda)] * (double)B[idx(tb, ki, ni, ldb)];
            float* c = &C[(long)ni * ldc + mi];
            if (beta_zero) *c = (float)(acc * (double)alpha);
            else *c = (float)(acc * (double)alpha + (double)(*c) * (double)beta);
        }
    }
}

void orc_sgemm(int transa, int transb, int m, int n, int k, float alpha, const float* A, int lda, const float* B,
               int ldb, float beta, float* C, int ldc)
{
    sgemm_cols(transa, transb, m, 0, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

typedef struct {
    int ta, tb, m, n0, n1, k;
    float alpha, beta;
    const float *A, *B;
    float* C;
    int lda, ldb, ldc;
} sgemm_job;

static void* sgemm_worker(void* p)
{
    sgemm_job* j = (sgemm_job*)p;
    sgemm_cols(j->ta, j->tb, j->m, j->n0, j->n1, j->k, j->alpha, j->A, j->lda, j->B, j->ldb, j->beta, j->C, j->ldc);
    return NULL;
}

/* orc_sgemm with the columns of C split over `threads` threads (identical results). */
int orc_sgemm_mt(int threads, int transa, int transb, int m, int n, int k, float alpha, const float* A, int lda,
                 const float* B, int ldb, float beta, float* C, int ldc)
{
    if (threads < 1) threads = 1;
    if (threads > n) threads = n > 0 ? n : 1;
    pthread_t* tid = (pthread_t*)malloc(sizeof(pthread_t) * threads);
    sgemm_job* jobs = (sgemm_job*)malloc(sizeof(sgemm_job) * threads);
    if (!tid || !jobs) {
        free(tid);
        free(jobs);
        return -1;
    }
    for (int t = 0; t < threads; ++t) {
        sgemm_job j = {transa, transb, m, (int)((long)n * t / threads), (int)((long)n * (t + 1) / threads), k, alpha,
                       beta, A, B, C, lda, ldb, ldc};
        jobs[t] = j;
        pthread_create(&tid[t], NULL, sgemm_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
    free(tid);
    free(jobs);
    return threads;
}

/* complex: interleaved {re, im}; acc in complex double with the textbook product
 * (what std::complex<double>::operator* computes for finite operands). */
void orc_cgemm(int transa, int transb, int m, int n, int k, const float* alpha, const float* A, int lda,
               const float* B, int ldb, const float* beta, float* C, int ldc)
{
    const int beta_zero = (beta[0] == 0.0f && beta[1] == 0.0f);
    const double sa = transa == 2 ? -1.0 : 1.0, sb = transb == 2 ? -1.0 : 1.0;
    const double ar = alpha[0], ai = alpha[1], br = beta[0], bi = beta[1];
    for (int mi = 0; mi < m; ++mi) {
        for (int ni = 0; ni < n; ++ni) {
            double accr = 0.0, acci = 0.0;
            for (int ki = 0; ki < k; ++ki) {
                const long ia = 2 * idx(transa, mi, ki, lda), ib = 2 * idx(transb, ki, ni, ldb);
                const double xr = A[ia], xi = sa * (double)A[ia + 1];
                const double yr = B[ib], yi = sb * (double)B[ib + 1];
                accr += xr * yr - xi * yi;
                acci += xr * yi + xi * yr;
            }
            float* c = &C[2 * ((long)ni * ldc + mi)];
            double vr = accr * ar - acci * ai;
            double vi = accr * ai + acci * ar;
            if (!beta_zero) {
                const double cr = c[0], ci = c[1];
                vr += cr * br - ci * bi;
                vi += cr * bi + ci * br;
            }
            c[0] = (float)vr;
            c[1] = (float)vi;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Layer 1: restatement of Strsm_naive (Strsm_naive.cl:18-126)                */
/* ------------------------------------------------------------------------- */
#define AT(X, r, c, ld) X[(long)(c) * (ld) + (r)]

#define TRSM_BODY(T, FMA)                                                                                               \
    const int left = side == 0, lower = uplo == 1, ntrans = trans == 0, ndiag = diag == 0;                           \
    if (left) {                                                                                                      \
        if (lower == ntrans) { /* forward: (lower, N) or (upper, T) */                                               \
            for (int i = 0; i < m; i++)                                                                              \
                for (int j = 0; j < n; j++) {                                                                        \
                    T x = AT(b, i, j, ldb);                                                                          \
                    if (ndiag) x /= (T)AT(a, i, i, lda);                                                             \
                    for (int k = i + 1; k < m; k++)                                                                  \
                        AT(b, k, j, ldb) = FMA(-(T)(ntrans ? AT(a, k, i, lda) : AT(a, i, k, lda)), x, AT(b, k, j, ldb)); \
                    AT(b, i, j, ldb) = x * (T)alpha;                                                                 \
                }                                                                                                    \
        } else {                                                                                                     \
            for (int i = m - 1; i >= 0; i--)                                                                         \
                for (int j = 0; j < n; j++) {                                                                        \
                    T x = AT(b, i, j, ldb);                                                                          \
                    if (ndiag) x /= (T)AT(a, i, i, lda);                                                             \
                    for (int k = 0; k < i; k++)                                                                      \
                        AT(b, k, j, ldb) = FMA(-(T)(ntrans ? AT(a, k, i, lda) : AT(a, i, k, lda)), x, AT(b, k, j, ldb)); \
                    AT(b, i, j, ldb) = x * (T)alpha;                                                                 \
                }                                                                                                    \
        }                                                                                                            \
    } else {                                                                                                         \
        if (lower == ntrans) { /* backward over columns: (lower, N) or (upper, T) */                                 \
            for (int i = n - 1; i >= 0; i--)                                                                         \
                for (int j = 0; j < m; j++) {                                                                        \
                    T x = AT(b, j, i, ldb);                                                                          \
                    if (ndiag) x /= (T)AT(a, i, i, lda);                                                             \
                    for (int k = 0; k < i; k++)                                                                      \
                        AT(b, j, k, ldb) = FMA(-(T)(ntrans ? AT(a, i, k, lda) : AT(a, k, i, lda)), x, AT(b, j, k, ldb)); \
                    AT(b, j, i, ldb) = x * (T)alpha;                                                                 \
                }                                                                                                    \
        } else {                                                                                                     \
            for (int i = 0; i < n; i++)                                                                              \
                for (int j = 0; j < m; j++) {                                                                        \
                    T x = AT(b, j, i, ldb);                                                                          \
                    if (ndiag) x /= (T)AT(a, i, i, lda);                                                             \
                    for (int k = i + 1; k < n; k++)                                                                  \
                        AT(b, j, k, ldb) = FMA(-(T)(ntrans ? AT(a, i, k, lda) : AT(a, k, i, lda)), x, AT(b, j, k, ldb)); \
                    AT(b, j, i, ldb) = x * (T)alpha;                                                                 \
                }                                                                                                    \
        }                                                                                                            \
    }

void orc_strsm(int side, int uplo, int trans, int diag, int m, int n, float alpha, const float* a, int lda, float* b,
               int ldb)
{
    TRSM_BODY(float, fmaf)
}

/* Same substitution in double: a (float, as stored) and b (double in/out). */
void orc_dtrsm(int side, int uplo, int trans, int diag, int m, int n, double alpha, const float* a, int lda,
               double* b, int ldb)
{
    TRSM_BODY(double, fma)
}

/* ------------------------------------------------------------------------- */
/* Restatement of Ctrsm_naive (libs/blas/funcs/Ctrsm/Ctrsm_naive.cl:16-183)    */
/* with the complex helpers of core/src/ocl/complex.h.cl:32-39:                */
/*   cmul(a, b) = (fma(ax, bx, -ay*by), fma(ax, by, ay*bx))                    */
/*   cdiv(a, b) = (fma(ax, bx, ay*by) / d, fma(ay, bx, -ax*by) / d),           */
/*                d = fma(bx, bx, by*by)                                       */
/* The eight loop nests of the kernel reduce to two shapes (as for the real    */
/* kernel): left solves run forward when lower == (trans == N) and use         */
/* a(k, i) for N, a(i, k) otherwise; right solves run backward when            */
/* lower == (trans == N) and use a(i, k) for N, a(k, i) otherwise. OP_C        */
/* conjugates every element of A read, the diagonal included. B is updated by  */
/* `b -= cmul(coef, x)` and the solved entry is stored as cmul(x, alpha).      */
/* ------------------------------------------------------------------------- */
#define CTRSM_BODY(T, FMA)                                                                                              \
    const int left = side == 0, lower = uplo == 1, ntrans = trans == 0, cj = trans == 2, ndiag = diag == 0;          \
    const int order = left ? m : n, lines = left ? n : m;                                                            \
    const int fwd = left ? (lower == ntrans) : !(lower == ntrans);                                                   \
    for (int t = 0; t < order; t++) {                                                                                \
        const int i = fwd ? t : order - 1 - t;                                                                       \
        for (int j = 0; j < lines; j++) {                                                                            \
            T *xb = left ? &b[2 * ((long)j * ldb + i)] : &b[2 * ((long)i * ldb + j)];                                \
            T xr = xb[0], xi = xb[1];                                                                                \
            if (ndiag) {                                                                                             \
                const T dr = (T)a[2 * ((long)i * lda + i)], di = (cj ? -1 : 1) * (T)a[2 * ((long)i * lda + i) + 1];  \
                const T dv = FMA(dr, dr, di * di);                                                                   \
                const T qr = FMA(xr, dr, xi * di) / dv, qi = FMA(xi, dr, -xr * di) / dv;                             \
                xr = qr;                                                                                             \
                xi = qi;                                                                                             \
            }                                                                                                        \
            const int k0 = fwd ? i + 1 : 0, k1 = fwd ? order : i;                                                    \
            for (int k = k0; k < k1; k++) {                                                                          \
                /* left: coef = N ? a(k, i) : a(i, k); right: coef = N ? a(i, k) : a(k, i) */                       \
                const int use_ki = left ? ntrans : !ntrans;                                                          \
                const long ia = 2 * (use_ki ? (long)i * lda + k : (long)k * lda + i);                                \
                const T cr = (T)a[ia], ci = (cj ? -1 : 1) * (T)a[ia + 1];                                            \
                const T tr = FMA(cr, xr, -ci * xi), ti = FMA(cr, xi, ci * xr);                                       \
                T *bk = left ? &b[2 * ((long)j * ldb + k)] : &b[2 * ((long)k * ldb + j)];                            \
                bk[0] -= tr;                                                                                         \
                bk[1] -= ti;                                                                                         \
            }                                                                                                        \
            xb[0] = FMA(xr, (T)alpha[0], -xi * (T)alpha[1]);                                                         \
            xb[1] = FMA(xr, (T)alpha[1], xi * (T)alpha[0]);                                                          \
        }                                                                                                            \
    }

/* a, b: interleaved complex64; alpha {re, im} */
void orc_ctrsm(int side, int uplo, int trans, int diag, int m, int n, const float* alpha, const float* a, int lda,
               float* b, int ldb)
{
    CTRSM_BODY(float, fmaf)
}

/* Same substitution in complex double: a as stored (complex64), b complex128 in/out. */
void orc_ztrsm(int side, int uplo, int trans, int diag, int m, int n, const double* alpha, const float* a, int lda,
               double* b, int ldb)
{
    CTRSM_BODY(double, fma)
}

/* ------------------------------------------------------------------------- */
/* Layer 2: FP64 references                                                    */
/* ------------------------------------------------------------------------- */
/* C64 = alpha*op(A)*op(B) + beta*C (C read only if beta != 0) and
 * absAB = |op(A)| |op(B)| (elementwise bound denominator); either output may be NULL.
 * Outputs are dense m x n column-major doubles. */
void orc_dgemm(int transa, int transb, int m, int n, int k, double alpha, const float* A, int lda, const float* B,
               int ldb, double beta, const float* C, int ldc, double* C64, double* absAB)
{
    for (int ni = 0; ni < n; ++ni)
        for (int mi = 0; mi < m; ++mi) {
            double acc = 0.0, aacc = 0.0;
            for (int ki = 0; ki < k; ++ki) {
                const double x = A[idx(transa, mi, ki, lda)], y = B[idx(transb, ki, ni, ldb)];
                acc += x * y;
                aacc += fabs(x) * fabs(y);
            }
            double v = alpha * acc;
            if (beta != 0.0) v += beta * (double)C[(long)ni * ldc + mi];
            if (C64) C64[(long)ni * m + mi] = v;
            if (absAB) absAB[(long)ni * m + mi] = aacc;
        }
}

/* Sampled version: npts (row, col) pairs. */
void orc_dgemm_sample(int transa, int transb, int k, double alpha, const float* A, int lda, const float* B, int ldb,
                      double beta, const float* C, int ldc, int npts, const int* rows, const int* cols, double* C64,
                      double* absAB)
{
    for (int p = 0; p < npts; ++p) {
        const int mi = rows[p], ni = cols[p];
        double acc = 0.0, aacc = 0.0;
        for (int ki = 0; ki < k; ++ki) {
            const double x = A[idx(transa, mi, ki, lda)], y = B[idx(transb, ki, ni, ldb)];
            acc += x * y;
            aacc += fabs(x) * fabs(y);
        }
        double v = alpha * acc;
        if (beta != 0.0) v += beta * (double)C[(long)ni * ldc + mi];
        C64[p] = v;
        if (absAB) absAB[p] = aacc;
    }
}

/* Complex FP64 reference, interleaved output C64 (dense m x n), and
 * absAB = sum |a||b| with |z| = |re| + |im| (reference blas_abs, test_helpers.hpp:207-217). */
static void zgemm_point(int transa, int transb, int k, const double* alpha, const float* A, int lda, const float* B,
                        int ldb, const double* beta, const float* C, int ldc, int mi, int ni, double* out,
                        double* aout)
{
    const double sa = transa == 2 ? -1.0 : 1.0, sb = transb == 2 ? -1.0 : 1.0;
    double accr = 0.0, acci = 0.0, aacc = 0.0;
    for (int ki = 0; ki < k; ++ki) {
        const long ia = 2 * idx(transa, mi, ki, lda), ib = 2 * idx(transb, ki, ni, ldb);
        const double xr = A[ia], xi = sa * (double)A[ia + 1];
        const double yr = B[ib], yi = sb * (double)B[ib + 1];
        accr += xr * yr - xi * yi;
        acci += xr * yi + xi * yr;
        aacc += (fabs(xr) + fabs(xi)) * (fabs(yr) + fabs(yi));
    }
    double vr = accr * alpha[0] - acci * alpha[1];
    double vi = accr * alpha[1] + acci * alpha[0];
    if (beta[0] != 0.0 || beta[1] != 0.0) {
        const double cr = C[2 * ((long)ni * ldc + mi)], ci = C[2 * ((long)ni * ldc + mi) + 1];
        vr += cr * beta[0] - ci * beta[1];
        vi += cr * beta[1] + ci * beta[0];
    }
    out[0] = vr;
    out[1] = vi;
    if (aout) *aout = aacc;
}

void orc_zgemm(int transa, int transb, int m, int n, int k, const double* alpha, const float* A, int lda,
               const float* B, int ldb, const double* beta, const float* C, int ldc, double* C64, double* absAB)
{
    for (int ni = 0; ni < n; ++ni)
        for (int mi = 0; mi < m; ++mi)
            zgemm_point(transa, transb, k, alpha, A, lda, B, ldb, beta, C, ldc, mi, ni, &C64[2 * ((long)ni * m + mi)],
                        absAB ? &absAB[(long)ni * m + mi] : NULL);
}

void orc_zgemm_sample(int transa, int transb, int k, const double* alpha, const float* A, int lda, const float* B,
                      int ldb, const double* beta, const float* C, int ldc, int npts, const int* rows,
                      const int* cols, double* C64, double* absAB)
{
    for (int p = 0; p < npts; ++p)
        zgemm_point(transa, transb, k, alpha, A, lda, B, ldb, beta, C, ldc, rows[p], cols[p], &C64[2 * p],
                    absAB ? &absAB[p] : NULL);
}
