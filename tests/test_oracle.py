"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the parity
checker for the GPU path. CPU only.

Each oracle function is checked against
  * the reference's known-answer tests (tests/golden/kats.json),
  * the reference's own cpu_gemm compiled from its headers (oracle/_ref), where
    that build exists (this container; not on the GPU box),
  * closed forms / library routines that the mathematics fixes (numpy float64
    products, scipy triangular solves, identity and unit-diagonal special cases),
chosen so that a transposed operand, a missing conjugate, a wrong substitution
order or a dropped alpha/beta term fails at least one of them.
"""
import json
from pathlib import Path

import numpy as np
import pytest
import scipy.linalg as sla

import oracle
from blas_helpers import adaptive_ok, cmat, complex_eq, complex_near, flat, float_eq, fmat
from clgpu_amd import datagen

KATS = json.loads((Path(__file__).parent / "golden" / "kats.json").read_text())


def opnp(X, trans, rows, cols):
    """op(X) of the logical rows x cols operand as float64/complex128 numpy."""
    if trans == 0:
        return X[:rows, :cols].astype(np.complex128 if np.iscomplexobj(X) else np.float64)
    Y = X[:cols, :rows].T.astype(np.complex128 if np.iscomplexobj(X) else np.float64)
    return np.conj(Y) if trans == 2 else Y


# ------------------------------------------------------------------ KATs
@pytest.mark.parametrize("case", KATS["sgemm"], ids=lambda c: c["name"])
def test_oracle_sgemm_kat(case):
    c = case
    ka = c["k"] if c["transa"] == 0 else c["m"]
    kb = c["n"] if c["transb"] == 0 else c["k"]
    A = fmat(c["A"], c["lda"], ka)
    B = fmat(c["B"], c["ldb"], kb)
    C = fmat(c["C"], c["ldc"], c["n"])
    oracle.sgemm(c["transa"], c["transb"], c["m"], c["n"], c["k"], c["alpha"], A, B, c["beta"], C)
    exp = np.asarray(c["expected"], np.float64)
    if c["compare"] == "ulp4":
        assert float_eq(flat(C), exp)
    else:  # FP64 exact expected; double accumulation makes the oracle exact to FP32 rounding
        got = flat(C).astype(np.float64)
        assert np.all(np.abs(got - exp) <= 0.5 * np.spacing(np.abs(exp).astype(np.float32)).astype(np.float64) + 0)


@pytest.mark.parametrize("case", KATS["cgemm"], ids=lambda c: c["name"])
def test_oracle_cgemm_kat(case):
    c = case
    ka = c["k"] if c["transa"] == 0 else c["m"]
    kb = c["n"] if c["transb"] == 0 else c["k"]
    A = cmat(c["A"], c["lda"], ka)
    B = cmat(c["B"], c["ldb"], kb)
    C = cmat(c["C"], c["ldc"], c["n"])
    oracle.cgemm(c["transa"], c["transb"], c["m"], c["n"], c["k"], complex(*c["alpha"]), A, B, complex(*c["beta"]), C)
    exp = np.asarray(c["expected"], np.float64)
    assert complex_eq(flat(C), exp[:, 0] + 1j * exp[:, 1])


@pytest.mark.parametrize("case", KATS["strsm"], ids=lambda c: c["name"])
def test_oracle_strsm_kat(case):
    c = case
    order = c["m"] if c["side"] == 0 else c["n"]
    A = fmat(c["A"], c["lda"], order)
    B = fmat(c["B"], c["ldb"], c["n"])
    oracle.strsm(c["side"], c["uplo"], c["trans"], c["diag"], c["m"], c["n"], c["alpha"], A, B)
    assert adaptive_ok(c["expected"], flat(B))


# ------------------------------------------------------------------ vs the reference build
needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (no reference tree here)")

TRANS_REAL = [(0, 0), (0, 1), (1, 0), (1, 1), (2, 0), (0, 2)]


@needs_ref
@pytest.mark.parametrize("ta,tb", TRANS_REAL)
@pytest.mark.parametrize("beta", [0.0, 0.5])
@pytest.mark.parametrize("pad", [0, 3])
def test_oracle_sgemm_matches_reference_build(ta, tb, beta, pad):
    m, n, k = 37, 23, 41
    ra, ca = (m, k) if ta == 0 else (k, m)
    rb, cb = (k, n) if tb == 0 else (n, k)
    A = datagen.matrix(ra, ca, ra + pad, "A", seed=3)
    B = datagen.matrix(rb, cb, rb + pad, "B", seed=3)
    C = datagen.matrix(m, n, m + pad, "C", seed=3)
    if beta == 0.0:
        C[:] = np.nan  # beta == 0 must not read C (test_gemm.hpp:75-81)
    C1, C2 = C.copy(order="F"), C.copy(order="F")
    oracle.sgemm(ta, tb, m, n, k, 1.25, A, B, beta, C1)
    oracle.sgemm(ta, tb, m, n, k, 1.25, A, B, beta, C2, reference=True)
    assert np.array_equal(C1[:m], C2[:m])


@needs_ref
@pytest.mark.parametrize("ta", [0, 1, 2])
@pytest.mark.parametrize("tb", [0, 1, 2])
@pytest.mark.parametrize("beta", [0.0, 0.25 - 0.5j])
def test_oracle_cgemm_matches_reference_build(ta, tb, beta):
    m, n, k = 19, 13, 29
    ra, ca = (m, k) if ta == 0 else (k, m)
    rb, cb = (k, n) if tb == 0 else (n, k)
    A = datagen.complex_matrix(ra, ca, ra + 2, "A", seed=5)
    B = datagen.complex_matrix(rb, cb, rb + 1, "B", seed=5)
    C = datagen.complex_matrix(m, n, m + 3, "C", seed=5)
    C1, C2 = C.copy(order="F"), C.copy(order="F")
    oracle.cgemm(ta, tb, m, n, k, 1 + 0.5j, A, B, beta, C1)
    oracle.cgemm(ta, tb, m, n, k, 1 + 0.5j, A, B, beta, C2, reference=True)
    assert np.array_equal(C1[:m], C2[:m])


# ------------------------------------------------------------------ mathematics
@pytest.mark.parametrize("ta,tb", TRANS_REAL)
def test_oracle_sgemm_quarter_integers_exact(ta, tb):
    """Quarter-integer data: every product and partial sum is exact, so cpu_gemm
    must equal the float64 product exactly (the reference's EXPECT_EQ design,
    tests/blas_funcs/test_gemm.cpp:47-53)."""
    m = n = k = 64
    ra, ca = (m, k) if ta == 0 else (k, m)
    rb, cb = (k, n) if tb == 0 else (n, k)
    A = np.asfortranarray(datagen.quarter_integers(ra * ca, 1, 0).reshape(ca, ra).T.astype(np.float32))
    B = np.asfortranarray(datagen.quarter_integers(rb * cb, 1, 1).reshape(cb, rb).T.astype(np.float32))
    C = np.asfortranarray(datagen.quarter_integers(m * n, 1, 2).reshape(n, m).T.astype(np.float32))
    exp = 0.75 * opnp(A, ta, m, k) @ opnp(B, tb, k, n) - 0.5 * C.astype(np.float64)
    oracle.sgemm(ta, tb, m, n, k, 0.75, A, B, -0.5, C)
    assert np.array_equal(C.astype(np.float64), exp)


def test_oracle_dgemm_zgemm_match_numpy():
    m, n, k = 17, 11, 9
    A = datagen.matrix(k, m, name="A")  # used transposed
    B = datagen.matrix(k, n, name="B")
    C = datagen.matrix(m, n, name="C")
    C64, ab = oracle.dgemm(1, 0, m, n, k, 2.0, A, B, 0.5, C)
    ref = 2.0 * A.astype(np.float64).T @ B + 0.5 * C
    assert np.allclose(C64, ref, rtol=1e-14, atol=1e-14)
    assert np.allclose(ab, np.abs(A.astype(np.float64)).T @ np.abs(B), rtol=1e-14)
    Az = datagen.complex_matrix(k, m, name="A")
    Bz = datagen.complex_matrix(n, k, name="B")
    Z, _ = oracle.zgemm(2, 1, m, n, k, 1 - 1j, Az, Bz)
    refz = (1 - 1j) * np.conj(Az.astype(np.complex128)).T @ Bz.astype(np.complex128).T
    assert np.allclose(Z, refz, rtol=1e-13, atol=1e-13)
    rows = np.array([0, 5, 16]); cols = np.array([0, 3, 10])
    s, _ = oracle.dgemm_sample(1, 0, k, 2.0, A, B, rows, cols, 0.5, C)
    assert np.allclose(s, ref[rows, cols], rtol=1e-14)


VARIANTS = [(s, u, t, d) for s in (0, 1) for u in (0, 1) for t in (0, 1) for d in (0, 1)]


def _tri(order, lower, seed=7):
    return datagen.triangular(order, lower=bool(lower), seed=seed, poison=True)


@pytest.mark.parametrize("side,uplo,trans,diag", VARIANTS)
def test_oracle_trsm_solves_and_reads_only_its_triangle(side, uplo, trans, diag):
    """The float restatement and the FP64 substitution solve op(A) X = alpha B
    (checked with scipy's solver) while the unused triangle -- and, for a unit
    diagonal, the diagonal -- hold NaN and are never read."""
    m, n = 23, 17
    order = m if side == 0 else n
    A = _tri(order, uplo)
    if diag == 1:
        np.fill_diagonal(A, np.nan)
    B = datagen.matrix(m, n, name="B", seed=7)
    alpha = -0.75
    # scipy on the clean matrix
    Ac = np.nan_to_num(A.astype(np.float64), nan=0.0)
    if diag == 1:
        np.fill_diagonal(Ac, 1.0)
    opA = Ac if trans == 0 else Ac.T
    lower_eff = (uplo == 1) != (trans != 0)
    if side == 0:
        X_ref = sla.solve_triangular(opA, alpha * B.astype(np.float64), lower=lower_eff)
    else:
        X_ref = sla.solve_triangular(opA.T, alpha * B.astype(np.float64).T, lower=not lower_eff).T
    X64 = oracle.dtrsm(side, uplo, trans, diag, m, n, alpha, A, B)
    assert np.all(np.isfinite(X64))
    assert np.allclose(X64, X_ref, rtol=1e-12, atol=1e-12)
    X32 = B.copy(order="F")
    oracle.strsm(side, uplo, trans, diag, m, n, alpha, A, X32)
    assert np.all(np.isfinite(X32))
    assert np.max(np.abs(X32 - X_ref)) <= 1e-5 * max(1.0, np.max(np.abs(X_ref)))


@pytest.mark.parametrize("side", [0, 1])
def test_oracle_trsm_identity_gives_alpha_b(side):
    m, n = 9, 6
    order = m if side == 0 else n
    A = np.asfortranarray(np.eye(order, dtype=np.float32))
    B = datagen.matrix(m, n, name="B")
    X = B.copy(order="F")
    oracle.strsm(side, 1, 0, 0, m, n, 1.5, A, X)
    assert np.array_equal(X, (B * np.float32(1.5)).astype(np.float32))


def test_planted_trsm_order_mistake_is_caught():
    """A plausible mistake -- forward instead of backward substitution for an
    upper-triangular left solve -- must fail the scipy check used above."""
    m, n = 12, 3
    A = _tri(m, 0)
    B = datagen.matrix(m, n, name="B")
    Af = np.nan_to_num(A.astype(np.float64))
    X_ref = sla.solve_triangular(Af, B.astype(np.float64), lower=False)
    wrong = oracle.dtrsm(0, 1, 0, 0, m, n, 1.0, np.asfortranarray(np.nan_to_num(A)), B)  # solves as lower
    assert not np.allclose(wrong, X_ref, rtol=1e-6)


# ------------------------------------------------------------------ CTRSM
@pytest.mark.parametrize("case", KATS["ctrsm"], ids=lambda c: c["name"])
def test_oracle_ctrsm_kat(case):
    c = case
    order = c["m"] if c["side"] == 0 else c["n"]
    A = cmat(c["A"], c["lda"], order)
    B = cmat(c["B"], c["ldb"], c["n"])
    oracle.ctrsm(c["side"], c["uplo"], c["trans"], c["diag"], c["m"], c["n"], complex(*c["alpha"]), A, B)
    exp = np.asarray(c["expected"], np.float64)
    assert complex_near(exp[:, 0] + 1j * exp[:, 1], flat(B))


CVARIANTS = [(s, u, t, d) for s in (0, 1) for u in (0, 1) for t in (0, 1, 2) for d in (0, 1)]


def _scipy_ctrsm(side, uplo, trans, diag, m, n, alpha, A, B):
    order = m if side == 0 else n
    Ac = np.nan_to_num(A[:order, :order].astype(np.complex128), nan=0.0)
    Ac = np.tril(Ac) if uplo == 1 else np.triu(Ac)
    if diag == 1:
        np.fill_diagonal(Ac, 1.0)
    opA = Ac if trans == 0 else (Ac.T if trans == 1 else Ac.conj().T)
    lower_eff = (uplo == 1) != (trans != 0)
    rhs = alpha * B[:m, :n].astype(np.complex128)
    if side == 0:
        return sla.solve_triangular(opA, rhs, lower=lower_eff)
    return sla.solve_triangular(opA.T, rhs.T, lower=not lower_eff).T


@pytest.mark.parametrize("side,uplo,trans,diag", CVARIANTS)
def test_oracle_ctrsm_solves_and_reads_only_its_triangle(side, uplo, trans, diag):
    """The complex float restatement and the complex FP64 substitution solve
    op(A) X = alpha B (checked with scipy's complex solver; op = A^H conjugates)
    while the unused triangle -- and, for a unit diagonal, the diagonal -- hold
    NaN and are never read."""
    m, n = 19, 13
    order = m if side == 0 else n
    A = datagen.complex_triangular(order, lower=bool(uplo), ld=order + 1, seed=11, poison=True)
    if diag == 1:
        A[np.arange(order), np.arange(order)] = np.nan
    B = datagen.complex_matrix(m + 2, n, name="B", seed=11)
    alpha = 0.8 - 0.35j
    X_ref = _scipy_ctrsm(side, uplo, trans, diag, m, n, alpha, A, B)
    Z = oracle.ztrsm(side, uplo, trans, diag, m, n, alpha, A, B)
    assert np.all(np.isfinite(Z))
    assert np.allclose(Z, X_ref, rtol=1e-12, atol=1e-12)
    X = B.copy(order="F")
    oracle.ctrsm(side, uplo, trans, diag, m, n, alpha, A, X)
    assert np.array_equal(X[m:], B[m:])  # padding rows untouched
    assert np.all(np.isfinite(X[:m]))
    assert np.max(np.abs(X[:m] - X_ref)) <= 1e-5 * max(1.0, np.max(np.abs(X_ref)))


def test_oracle_ctrsm_identity_gives_alpha_b():
    m, n = 7, 5
    A = np.asfortranarray(np.eye(m, dtype=np.complex64))
    B = datagen.complex_matrix(m, n, name="B")
    X = B.copy(order="F")
    oracle.ctrsm(0, 1, 2, 0, m, n, 1.5 - 0.5j, A, X)
    assert np.allclose(X, B.astype(np.complex128) * (1.5 - 0.5j), rtol=1e-6, atol=1e-6)


def test_planted_ctrsm_conjugation_mistake_is_caught():
    """A plausible mistake -- OP_C solved as plain transpose (conjugate dropped)
    -- must fail the scipy check used above."""
    m, n = 10, 4
    A = datagen.complex_triangular(m, lower=True, seed=5, poison=False)
    B = datagen.complex_matrix(m, n, name="B", seed=5)
    X_ref = _scipy_ctrsm(0, 1, 2, 0, m, n, 1.0, A, B)
    wrong = oracle.ztrsm(0, 1, 1, 0, m, n, 1.0, A, B)  # conjugate dropped
    assert not np.allclose(wrong, X_ref, rtol=1e-6)
    assert np.allclose(oracle.ztrsm(0, 1, 2, 0, m, n, 1.0, A, B), X_ref, rtol=1e-12, atol=1e-12)
