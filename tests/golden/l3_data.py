"""Input data of the reference's known-answer tests that are built by loops
rather than written out, rebuilt here with the same float32 arithmetic.

sgemv_arrays: tests/iclBLAS/test_Sgemv.cpp (all 8 cases).
ssymm_256: tests/iclBLAS/test_Ssymm.cpp:236-277 (256x256_left_up).
  A (lda = 256) is filled with 1.2f, then the upper triangle set to
  (float)(i*j + i + j) / (m*m + 2*m); B[i] = 1.f * i / ldb / n; C = 2.3f;
  alpha = 0.9f, beta = -1.2f; the reference compares with EXPECT_NEAR 1e-4
  against a double-precision loop over the upper triangle.
"""
import numpy as np

f32 = np.float32


def ssymm_256():
    """Scalar parameters of the case (the arrays come from ssymm_256_arrays)."""
    return dict(side=0, uplo=0, m=256, n=256, lda=256, ldb=256, ldc=256, alpha=float(f32(0.9)),
                beta=float(f32(-1.2)), compare_abs=1e-4)


def ssymm_256_arrays():
    m = n = 256
    lda = ldb = ldc = 256
    A = np.full((lda, m), f32(1.2), np.float32, order="F")
    i, j = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    up = j >= i
    # (float)(i*j + i + j) / (m*m + 2*m): int numerator converted to float, then
    # a float / int division in float32.
    vals = (i * j + i + j).astype(np.float32) / f32(m * m + 2 * m)
    A[up] = vals[up]
    # B[idx] = 1.f * idx / ldb / n, left to right in float32
    idx = np.arange(ldb * n)
    Bflat = (idx.astype(np.float32) / f32(ldb)) / f32(n)
    B = np.asfortranarray(Bflat.reshape(n, ldb).T)
    C = np.full((ldc, n), f32(2.3), np.float32, order="F")
    return A, B, C


def _ratio(count, n):
    """v[i] = 1.f * i / n / n for i < count, in float32 left to right."""
    return (np.arange(count).astype(np.float32) / f32(n)) / f32(n)


def sgemv_arrays(c):
    """test_Sgemv.cpp: A[i], x[i], y[i] = 1.f * i / n / n over A (lda * n), x
    (n * incx for op N, m * incx otherwise) and y (m * incy for op N, n * incy
    otherwise), with the test's n. Returns (A as a Fortran lda x n matrix, x, y)."""
    m, n, lda = c["m"], c["n"], c["lda"]
    lenx, leny = (n, m) if c["trans"] == 0 else (m, n)
    A = np.asfortranarray(_ratio(lda * n, n).reshape(n, lda).T)
    return A, _ratio(lenx * c["incx"], n), _ratio(leny * c["incy"], n)
