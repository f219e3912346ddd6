"""Shared pieces of the GEMV tests (test infrastructure): the reference's known
answers (tests/golden/kats_gemv.json) as arrays, and random problems whose
unused parts are poisoned."""
import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"
KATS = json.loads((GOLDEN / "kats_gemv.json").read_text())["cases"]


def _scalar(v):
    return complex(v[0], v[1]) if isinstance(v, list) else float(v)


def kat_arrays(c):
    """(A Fortran lda x n, x, y) of a fixture case."""
    if c["routine"] == "Sgemv":
        from golden.l3_data import sgemv_arrays
        return sgemv_arrays(c)
    v = np.asarray(c["A"], np.float64)
    A = np.asfortranarray((v[:, 0] + 1j * v[:, 1]).astype(np.complex64).reshape(-1, c["lda"]).T)
    x = np.asarray([complex(*p) for p in c["x"]], np.complex64)
    y = np.asarray([complex(*p) for p in c["y"]], np.complex64)
    return A, x, y


def expected(c, key):
    v = c[key]
    if c["routine"] == "Sgemv":
        return np.asarray(v, np.float64)
    v = np.asarray(v, np.float64)
    return v[:, 0] + 1j * v[:, 1]


def scalars(c):
    return _scalar(c["alpha"]), _scalar(c["beta"])


def random_problem(cplx, trans, m, n, incx, incy, rng, pad=3, beta=None, poison=True):
    """A (lda = m + pad, padding NaN), x and y with strides (gaps NaN / sentinel);
    y NaN when beta == 0 (it must not be read)."""
    dt = np.complex64 if cplx else np.float32

    def rnd(*shape):
        v = rng.uniform(-1, 1, shape)
        if cplx:
            v = v + 1j * rng.uniform(-1, 1, shape)
        return v.astype(dt)

    lda = m + pad
    A = np.full((lda, n), np.nan if poison else 0, dt, order="F")
    A[:m] = rnd(m, n)
    lenx, leny = (n, m) if trans == 0 else (m, n)
    x = np.full(max(1, (lenx - 1) * incx + 1), np.nan if poison else 0, dt)
    x[::incx] = rnd(lenx)
    y = np.full(max(1, (leny - 1) * incy + 1), -5.0, dt)
    y[::incy] = rnd(leny)
    alpha = (0.75 - 0.5j) if cplx else 0.75
    if beta is None:
        beta = (-0.5 + 0.25j) if cplx else -0.5
    if beta == 0 and poison:
        y[::incy] = np.nan
    return A, x, y, alpha, beta
