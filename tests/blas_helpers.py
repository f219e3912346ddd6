"""Comparison helpers shared by the tests (test infrastructure).

ulp_distance follows gtest's FloatingPoint<float>::AlmostEquals, the
comparison behind the reference's EXPECT_FLOAT_EQ (4 ULP) and
ComplexEqual (4 ULP per component; reference tests/gtest_utils.hpp:43-47).
adaptive_ok is the reference's ADAPTIVE_FLOAT_EQ (tests/iclBLAS/test_Strsm.cpp:19):
|expected| < 1 -> |e - r| <= 4*FLT_EPSILON, else 4 ULP.
"""
import numpy as np

FLT_EPSILON = float(np.finfo(np.float32).eps)
U32 = 2.0 ** -24


def _biased(x):
    b = np.asarray(x, np.float32).view(np.uint32).astype(np.int64)
    sign = b & 0x80000000
    mag = b & 0x7FFFFFFF
    return np.where(sign != 0, -mag, mag)


def ulp_distance(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    d = np.abs(_biased(a) - _biased(b))
    return np.where(np.isnan(a) | np.isnan(b), np.iinfo(np.int64).max, d)


def float_eq(a, b, ulps=4):
    return bool(np.all(ulp_distance(a, b) <= ulps))


def complex_eq(a, b, ulps=4):
    a = np.asarray(a, np.complex64)
    b = np.asarray(b, np.complex64)
    return float_eq(a.real, b.real, ulps) and float_eq(a.imag, b.imag, ulps)


def adaptive_ok(expected, result):
    e = np.asarray(expected, np.float64)
    r = np.asarray(result, np.float64)
    small = np.abs(e) < 1.0
    ok_small = np.abs(e - r) <= 4.0 * FLT_EPSILON
    ok_big = ulp_distance(e, r) <= 4
    return bool(np.all(np.where(small, ok_small, ok_big)))


def fmat(values, ld, cols, dtype=np.float32):
    """Column-major (ld x cols) array from a flat storage-order list."""
    v = np.asarray(values, dtype=dtype)
    return np.asfortranarray(v.reshape(cols, ld).T)


def cmat(pairs, ld, cols):
    p = np.asarray(pairs, np.float64)
    return fmat(p[:, 0] + 1j * p[:, 1], ld, cols, np.complex64)


def flat(a):
    return np.asarray(a).T.reshape(-1)


def bound_ratio(err, k, absAB, extra=None):
    """max over ij of err_ij / (k*2^-24*absAB_ij + extra_ij); the elementwise
    bound err <= 4*k*2^-24*(|A||B|)_ij holds when this is <= 4."""
    den = k * U32 * np.asarray(absAB, np.float64)
    if extra is not None:
        den = den + extra
    err = np.asarray(err, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(den > 0, err / den, np.where(err == 0, 0.0, np.inf))
    return float(r.max(initial=0.0))


def complex_near(expected, result, eps=4e-4):
    """EXPECT_COMPLEX_NEAR of the reference (tests/iclBLAS/test_Ctrsm.cpp:21-34):
    per component, |1 - result/expected| <= eps, or |result| <= eps where the
    expected component is 0."""
    e = np.asarray(expected, np.complex128)
    r = np.asarray(result, np.complex128)
    ok = True
    for ev, rv in ((e.real, r.real), (e.imag, r.imag)):
        nz = ev != 0
        ok &= bool(np.all(np.abs(1.0 - rv[nz] / ev[nz]) <= eps)) and bool(np.all(np.abs(rv[~nz]) <= eps))
    return ok
