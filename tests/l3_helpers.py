"""Shared pieces of the level-3 tests (test infrastructure): the known answers
of tests/golden/kats_l3.json turned into arrays, calls of the library and of
the FP64 oracle with one argument convention, and independent numpy closed
forms of each routine's header contract (include/iclBLAS/iclBLAS.h).

A case dict holds the routine's scalar parameters (side, uplo, trans, diag,
m, n, k, alpha, beta, lda, ldb, ldc) and arrays A, B, C as Fortran numpy arrays
whose shape[0] is the leading dimension.
"""
import json
from pathlib import Path

import numpy as np

import oracle
from clgpu_amd import iclblas as ib

GOLDEN = Path(__file__).resolve().parent / "golden"
KATS = json.loads((GOLDEN / "kats_l3.json").read_text())["cases"]

SYMM = ("Ssymm", "Csymm", "Chemm")
RANKK = ("Ssyrk", "Csyrk", "Cherk")
RANK2K = ("Ssyr2k", "Csyr2k", "Cher2k")
TRMM = ("Strmm", "Ctrmm")
COMPLEX = {"Csymm", "Chemm", "Csyrk", "Cherk", "Csyr2k", "Cher2k", "Ctrmm"}
HERM = {"Chemm", "Cherk", "Cher2k"}


def _mat(vals, ld, cols, cplx):
    v = np.asarray(vals, np.float64)
    v = (v[:, 0] + 1j * v[:, 1]).astype(np.complex64) if cplx else v.astype(np.float32)
    return np.asfortranarray(v.reshape(cols, ld).T)


def _scalar(v):
    return complex(v[0], v[1]) if isinstance(v, list) else float(v)


def shapes(c):
    """(rows, cols) of A, B and C for the case's routine."""
    r = c["routine"]
    if r in SYMM or r in TRMM:
        order = c["m"] if c["side"] == 0 else c["n"]
        return (order, order), (c["m"], c["n"]), (c["m"], c["n"])
    rows, cols = (c["n"], c["k"]) if c["trans"] == 0 else (c["k"], c["n"])
    return (rows, cols), (rows, cols), (c["n"], c["n"])


def kat_case(c):
    """Fixture record -> case dict with numpy arrays."""
    cplx = c["routine"] in COMPLEX
    out = {k: v for k, v in c.items() if k not in ("A", "B", "C")}
    out["alpha"] = _scalar(c["alpha"])
    if "beta" in c:
        out["beta"] = _scalar(c["beta"])
    if c.get("generator") == "ssymm_256":
        from golden.l3_data import ssymm_256_arrays
        out["A"], out["B"], out["C"] = ssymm_256_arrays()
        return out
    (ra, ca), (rb, cb), (rc, cc) = shapes(c)
    out["A"] = _mat(c["A"], c["lda"], ca, cplx)
    if "B" in c:
        out["B"] = _mat(c["B"], c["ldb"], cb, cplx)
    out["C"] = _mat(c["C"], c["ldc"], cc, cplx)
    return out


def expected_matrix(c, key):
    """A fixture's flat expected list as a matrix shaped like C."""
    cplx = c["routine"] in COMPLEX
    return _mat(c[key], c["ldc"], shapes(c)[2][1], cplx)


def call_library(h, c, A, B, C):
    """Runs the case's routine through the C ABI on handle h (C is updated in place)."""
    r = c["routine"]
    fn = getattr(ib, "iclblas" + r)
    if r in SYMM:
        return fn(h, c["side"], c["uplo"], c["m"], c["n"], c["alpha"], A, c["lda"], B, c["ldb"], c["beta"], C, c["ldc"])
    if r in RANKK:
        return fn(h, c["uplo"], c["trans"], c["n"], c["k"], c["alpha"], A, c["lda"], c["beta"], C, c["ldc"])
    if r in RANK2K:
        return fn(h, c["uplo"], c["trans"], c["n"], c["k"], c["alpha"], A, c["lda"], B, c["ldb"], c["beta"], C, c["ldc"])
    return fn(h, c["side"], c["uplo"], c["trans"], c["diag"], c["m"], c["n"], c["alpha"], A, c["lda"], B, c["ldb"], C,
              c["ldc"])


def oracle_result(c, A, B, C):
    """FP64 oracle: (full C after the call as complex128 / float64 matrix shaped
    like C, magnitude m x n or n x n, k of the bound)."""
    r = c["routine"]
    cplx = r in COMPLEX
    out = np.array(C, np.complex128 if cplx else np.float64)
    if r in SYMM:
        X, mag = oracle.zsymm(c["side"], c["uplo"], r in HERM, c["m"], c["n"], c["alpha"], A, B, c["beta"], C)
        out[: c["m"], : c["n"]] = X
        k = c["m"] if c["side"] == 0 else c["n"]
    elif r in RANKK:
        X, mag = oracle.zsyrk(c["uplo"], c["trans"], r in HERM, c["n"], c["k"], c["alpha"], A, c["beta"], C)
        out[: c["n"], : c["n"]] = X
        k = c["k"]
    elif r in RANK2K:
        X, mag = oracle.zsyr2k(c["uplo"], c["trans"], r in HERM, c["n"], c["k"], c["alpha"], A, B, c["beta"], C)
        out[: c["n"], : c["n"]] = X
        k = 2 * c["k"]
    else:
        X, mag = oracle.ztrmm(c["side"], c["uplo"], c["trans"], c["diag"], c["m"], c["n"], c["alpha"], A, B)
        out[: c["m"], : c["n"]] = X
        k = c["m"] if c["side"] == 0 else c["n"]
    return out, mag, k


# --------------------------------------------------------------------------
# numpy closed forms of the header contract (float64 / complex128), written
# independently of the oracle's loops.
# --------------------------------------------------------------------------
def _op(X, trans, herm_as_conj=False):
    X = np.asarray(X, np.complex128)
    if trans == 0:
        return X
    return X.T.conj() if (trans == 2 or herm_as_conj) else X.T


def closed_form(c, A, B, C):
    r = c["routine"]
    cplx = r in COMPLEX
    al, lower = c["alpha"], c["uplo"] == 1
    out = np.array(C, np.complex128)
    if r in SYMM:
        order = c["m"] if c["side"] == 0 else c["n"]
        Az = np.asarray(A[:order, :order], np.complex128)
        T = np.tril(Az) if lower else np.triu(Az)
        strict = np.tril(Az, -1) if lower else np.triu(Az, 1)
        mirror = strict.T.conj() if r == "Chemm" else strict.T
        S = T + mirror
        if r == "Chemm":
            S[np.diag_indices(order)] = S.diagonal().real
        Bz = np.asarray(B[: c["m"], : c["n"]], np.complex128)
        P = S @ Bz if c["side"] == 0 else Bz @ S
        Cz = out[: c["m"], : c["n"]]
        out[: c["m"], : c["n"]] = al * P + (c["beta"] * Cz if c["beta"] != 0 else 0)
    elif r in RANKK or r in RANK2K:
        n, k, t = c["n"], c["k"], c["trans"]
        herm = r in HERM
        rows, cols = (n, k) if t == 0 else (k, n)
        oA = _op(A[:rows, :cols], t, herm)
        if r in RANK2K:
            oB = _op(B[:rows, :cols], t, herm)
            if herm:
                P = al * oA @ oB.conj().T + np.conj(al) * oB @ oA.conj().T
            else:
                P = al * (oA @ oB.T + oB @ oA.T)
        else:
            P = al * (oA @ oA.conj().T if herm else oA @ oA.T)
        Cz = out[:n, :n]
        full = P + (c["beta"] * Cz if c["beta"] != 0 else 0)
        mask = np.tril(np.ones((n, n), bool)) if lower else np.triu(np.ones((n, n), bool))
        Cz[mask] = full[mask]
        if herm:
            Cz[np.diag_indices(n)] = Cz.diagonal().real
    else:
        order = c["m"] if c["side"] == 0 else c["n"]
        Az = np.asarray(A[:order, :order], np.complex128)
        T = np.tril(Az) if lower else np.triu(Az)
        if c["diag"] == 1:
            T[np.diag_indices(order)] = 1.0
        oT = _op(T, c["trans"])
        Bz = np.asarray(B[: c["m"], : c["n"]], np.complex128)
        out[: c["m"], : c["n"]] = al * (oT @ Bz if c["side"] == 0 else Bz @ oT)
    return out if cplx else out.real


# --------------------------------------------------------------------------
# Random problems
# --------------------------------------------------------------------------
def random_case(routine, rng, m=None, n=None, k=None, side=0, uplo=0, trans=0, diag=0, pad=3, poison=True,
                alpha=None, beta=None):
    """A random case whose never-read parts (the other triangle of A, a unit
    diagonal, ld padding, C when beta == 0) are NaN when poison is set, and
    whose untouched parts of C hold recognisable sentinels."""
    cplx = routine in COMPLEX
    dt = np.complex64 if cplx else np.float32

    def rnd(rows, cols, ld):
        X = np.full((ld, cols), -7.0, dt, order="F")
        v = rng.uniform(-1, 1, (rows, cols))
        if cplx:
            v = v + 1j * rng.uniform(-1, 1, (rows, cols))
        X[:rows] = v.astype(dt)
        return X

    c = {"routine": routine, "side": side, "uplo": uplo, "trans": trans, "diag": diag}
    c["m"], c["n"], c["k"] = m, n, k
    if alpha is None:
        alpha = complex(0.75, -0.5) if cplx else 0.75
    if routine in ("Cherk",):
        alpha = float(np.real(alpha))
    c["alpha"] = alpha
    if beta is None:
        beta = complex(-0.5, 0.25) if cplx else -0.5
    if routine in ("Cherk", "Cher2k"):
        beta = float(np.real(beta))
    c["beta"] = beta
    (ra, ca), (rb, cb), (rc, cc) = shapes(c)
    c["lda"], c["ldb"], c["ldc"] = ra + pad, rb + pad, rc + pad
    A, B, C = rnd(ra, ca, c["lda"]), rnd(rb, cb, c["ldb"]), rnd(rc, cc, c["ldc"])
    if poison and (routine in SYMM or routine in TRMM):
        order = ra
        iu = np.triu_indices(order, 1) if uplo == 1 else np.tril_indices(order, -1)
        A[iu] = np.nan  # the triangle that is not stored
        if routine in TRMM and diag == 1:
            A[np.arange(order), np.arange(order)] = np.nan
        if routine == "Chemm":
            d = np.arange(order)
            diag = A[d, d].copy()
            diag.imag = np.nan  # imaginary parts of the diagonal are not read
            A[d, d] = diag
    if poison and c.get("beta", 1) == 0 and not (routine in TRMM):
        C[:rc] = np.nan
    if routine in TRMM and poison:
        C[:rc] = np.nan  # output only
    c["A"], c["B"], c["C"] = A, B, C
    return c
