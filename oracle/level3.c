/*
 * level3.c -- FP64 oracle for the GEMM-based level-3 routines. TEST
 * INFRASTRUCTURE ONLY (see oracle.c).
 *
 * Every routine computes in complex double (interleaved {re, im}; a real
 * routine is the case of zero imaginary parts, which complex double arithmetic
 * reproduces exactly), element by element in the loop order of the reference
 * kernel it restates, and can also return the elementwise magnitude
 * mag(i, j) = sum_l |x_il| |y_lj| of the product it forms (the |A||B| of the
 * error bound; alpha not included).
 *
 *   orc_zsymm  <- Ssymm_naive.cl:18-75, Csymm_naive.cl:18-73, Chemm_naive.cl:18-93
 *                 (stored triangle read directly, the other one mirrored;
 *                 hermitian: mirrored entries conjugated, diagonal imaginary
 *                 parts taken as 0).
 *   orc_zsyrk  <- Ssyrk_naive.cl:18-47, Csyrk_naive.cl:18-49 (op C of the complex
 *                 symmetric form conjugates the sum, i.e. op(A) = A^H), and for
 *                 herm the header contract of Cherk (iclBLAS.h:1927-1954):
 *                 C = alpha op(A) op(A)^H + beta C, diagonal imaginary parts 0.
 *   orc_zsyr2k <- Ssyr2k_naive.cl:18-49, Csyr2k_naive.cl:18-51, and for herm the
 *                 header contract of Cher2k (iclBLAS.h:1956-1991):
 *                 alpha op(A) op(B)^H + conj(alpha) op(B) op(A)^H + beta C.
 *   orc_ztrmm  <- the header contract of Strmm / Ctrmm (iclBLAS.h:1721-1753,
 *                 1993-2025): C = alpha op(A) B or alpha B op(A), A triangular.
 *
 * Where Cherk_naive.cl, Cher2k_naive.cl, Strmm_naive.cl and Ctrmm_naive.cl
 * compute something else than their header states (they index the wrong
 * operand elements or mask the output instead of A), the oracle follows the
 * header; the tests show with the reference's own known answers where the two
 * differ.
 *
 * beta == 0 does not read C (the header: "if beta == 0, C does not have to be
 * initialized"). Updates of SYRK type touch only the uplo triangle of C.
 */
#include <math.h>
#include <stddef.h>

typedef struct {
    double re, im;
} zc;

static inline zc zget(const double* X, long i) { zc v = {X[2 * i], X[2 * i + 1]}; return v; }
static inline void zput(double* X, long i, zc v) { X[2 * i] = v.re; X[2 * i + 1] = v.im; }
static inline zc zmul(zc a, zc b) { zc v = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; return v; }
static inline zc zadd(zc a, zc b) { zc v = {a.re + b.re, a.im + b.im}; return v; }
static inline zc zconj(zc a) { zc v = {a.re, -a.im}; return v; }
static inline double zabs(zc a) { return hypot(a.re, a.im); }
static inline int zzero(zc a) { return a.re == 0.0 && a.im == 0.0; }
static const zc kZero = {0.0, 0.0};
static const zc kOne = {1.0, 0.0};

#define AT(r, c, ld) ((long)(c) * (ld) + (r))

/* C(i, j) = alpha * v + beta * C(i, j), C not read when beta == 0. */
static void finish(double* C, long ci, zc alpha, zc v, zc beta)
{
    zc out = zmul(alpha, v);
    if (!zzero(beta)) out = zadd(out, zmul(beta, zget(C, ci)));
    zput(C, ci, out);
}

/* A(p, q) of a symmetric / hermitian matrix stored in one triangle. */
static zc sym_at(const double* A, int lda, int lower, int herm, int p, int q)
{
    const int stored = lower ? p >= q : p <= q;
    zc v = stored ? zget(A, AT(p, q, lda)) : zget(A, AT(q, p, lda));
    if (herm && !stored) v = zconj(v);
    if (herm && p == q) v.im = 0.0;
    return v;
}

void orc_zsymm(int side, int uplo, int herm, int m, int n, const double* alpha, const double* A, int lda,
               const double* B, int ldb, const double* beta, double* C, int ldc, double* mag)
{
    const zc al = {alpha[0], alpha[1]}, be = {beta[0], beta[1]};
    const int lower = uplo == 1;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) {
            zc v = kZero;
            double g = 0.0;
            if (side == 0) {
                for (int l = 0; l < m; ++l) {
                    const zc a = sym_at(A, lda, lower, herm, i, l), b = zget(B, AT(l, j, ldb));
                    v = zadd(v, zmul(a, b));
                    g += zabs(a) * zabs(b);
                }
            } else {
                for (int l = 0; l < n; ++l) {
                    const zc b = zget(B, AT(i, l, ldb)), a = sym_at(A, lda, lower, herm, l, j);
                    v = zadd(v, zmul(b, a));
                    g += zabs(a) * zabs(b);
                }
            }
            finish(C, AT(i, j, ldc), al, v, be);
            if (mag) mag[AT(i, j, m)] = g;
        }
}

/* op(X)(i, l) of an n x k operand: X(i, l) for trans 0, otherwise X(l, i),
 * conjugated for conj. */
static zc op_at(const double* X, int ldx, int trans, int conj, int i, int l)
{
    zc v = trans == 0 ? zget(X, AT(i, l, ldx)) : zget(X, AT(l, i, ldx));
    return conj ? zconj(v) : v;
}

void orc_zsyrk(int uplo, int trans, int herm, int n, int k, const double* alpha, const double* A, int lda,
               const double* beta, double* C, int ldc, double* mag)
{
    const zc al = {alpha[0], alpha[1]}, be = {beta[0], beta[1]};
    /* op(A) = A, A^T, or A^H (complex symmetric op C; every non-N op of herk) */
    const int conj_op = trans != 0 && (herm || trans == 2);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            if (uplo == 1 ? i < j : i > j) continue;
            zc v = kZero;
            double g = 0.0;
            for (int l = 0; l < k; ++l) {
                const zc x = op_at(A, lda, trans, conj_op, i, l);
                zc y = op_at(A, lda, trans, conj_op, j, l);
                if (herm) y = zconj(y);
                v = zadd(v, zmul(x, y));
                g += zabs(x) * zabs(y);
            }
            finish(C, AT(i, j, ldc), al, v, be);
            if (herm && i == j) C[2 * AT(i, j, ldc) + 1] = 0.0;
            if (mag) mag[AT(i, j, n)] = g;
        }
}

void orc_zsyr2k(int uplo, int trans, int herm, int n, int k, const double* alpha, const double* A, int lda,
                const double* B, int ldb, const double* beta, double* C, int ldc, double* mag)
{
    const zc al = {alpha[0], alpha[1]}, be = {beta[0], beta[1]};
    const int conj_op = trans != 0 && (herm || trans == 2);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            if (uplo == 1 ? i < j : i > j) continue;
            zc v1 = kZero, v2 = kZero;
            double g = 0.0;
            for (int l = 0; l < k; ++l) {
                const zc ai = op_at(A, lda, trans, conj_op, i, l), bi = op_at(B, ldb, trans, conj_op, i, l);
                zc aj = op_at(A, lda, trans, conj_op, j, l), bj = op_at(B, ldb, trans, conj_op, j, l);
                if (herm) {
                    aj = zconj(aj);
                    bj = zconj(bj);
                }
                v1 = zadd(v1, zmul(ai, bj));
                v2 = zadd(v2, zmul(bi, aj));
                g += zabs(ai) * zabs(bj) + zabs(bi) * zabs(aj);
            }
            /* symmetric: alpha (v1 + v2); hermitian: alpha v1 + conj(alpha) v2 */
            zc v = herm ? zadd(zmul(al, v1), zmul(zconj(al), v2)) : zmul(al, zadd(v1, v2));
            finish(C, AT(i, j, ldc), kOne, v, be);
            if (herm && i == j) C[2 * AT(i, j, ldc) + 1] = 0.0;
            if (mag) mag[AT(i, j, n)] = g;
        }
}

/* op(A)(p, q) of a triangular A (uplo triangle stored, unit diagonal not read). */
static zc tri_at(const double* A, int lda, int lower, int trans, int unit, int p, int q)
{
    int r = p, c = q;
    if (trans != 0) {
        r = q;
        c = p;
    }
    if (lower ? r < c : r > c) return kZero;
    if (unit && r == c) return kOne;
    zc v = zget(A, AT(r, c, lda));
    return trans == 2 ? zconj(v) : v;
}

void orc_ztrmm(int side, int uplo, int trans, int diag, int m, int n, const double* alpha, const double* A, int lda,
               const double* B, int ldb, double* C, int ldc, double* mag)
{
    const zc al = {alpha[0], alpha[1]};
    const int lower = uplo == 1, unit = diag == 1;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) {
            zc v = kZero;
            double g = 0.0;
            if (!zzero(al)) {
                if (side == 0) {
                    for (int l = 0; l < m; ++l) {
                        const zc a = tri_at(A, lda, lower, trans, unit, i, l), b = zget(B, AT(l, j, ldb));
                        v = zadd(v, zmul(a, b));
                        g += zabs(a) * zabs(b);
                    }
                } else {
                    for (int l = 0; l < n; ++l) {
                        const zc b = zget(B, AT(i, l, ldb)), a = tri_at(A, lda, lower, trans, unit, l, j);
                        v = zadd(v, zmul(b, a));
                        g += zabs(a) * zabs(b);
                    }
                }
            }
            zput(C, AT(i, j, ldc), zmul(al, v));
            if (mag) mag[AT(i, j, m)] = g;
        }
}
