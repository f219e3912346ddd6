"""CTRSM parity on the GPU through the C ABI: the reference's known answers
(tests/iclBLAS/test_Ctrsm.cpp, EXPECT_COMPLEX_NEAR with 4e-4), all 24
side/uplo/trans(N, T, C)/diag variants against the complex FP64 substitution of
the oracle at sizes that take one diagonal block, several, and several 256-unknown
panels with a partial last one (the default panel solve and the recursive one),
never-read triangles and unit diagonals poisoned with NaN, padded B,
alpha == 0, and the FP32 path at 1024 x 512."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from blas_helpers import cmat, complex_near, flat
from clgpu_amd import datagen
from clgpu_amd import iclblas as ib

pytestmark = pytest.mark.gpu
KATS = json.loads((Path(__file__).parent / "golden" / "kats.json").read_text())
VARIANTS = [(s, u, t, d) for s in (0, 1) for u in (0, 1) for t in (0, 1, 2) for d in (0, 1)]


def vid(v):
    s, u, t, d = v
    return f"{'LR'[s]}{'UL'[u]}{'NTC'[t]}{'NU'[d]}"


@pytest.fixture
def h(handle):
    handle.set_math_mode(ib.MATH_FP32)
    handle.set_implementation(None)
    yield handle


@pytest.mark.parametrize("case", KATS["ctrsm"], ids=lambda c: c["name"])
def test_ctrsm_kat(h, case):
    c = case
    order = c["m"] if c["side"] == 0 else c["n"]
    A = cmat(c["A"], c["lda"], order)
    B = cmat(c["B"], c["ldb"], c["n"])
    ib.check(ib.iclblasCtrsm(h.value, c["side"], c["uplo"], c["trans"], c["diag"], c["m"], c["n"],
                             complex(*c["alpha"]), A, c["lda"], B, c["ldb"]))
    exp = np.asarray(c["expected"], np.float64)
    assert complex_near(exp[:, 0] + 1j * exp[:, 1], flat(B)), flat(B)


def _problem(side, uplo, diag, m, n, seed=41, pad=3):
    order = m if side == 0 else n
    A = datagen.complex_triangular(order, lower=bool(uplo), ld=order + 2, seed=seed, poison=True)
    if diag == 1:
        A[np.arange(order), np.arange(order)] = np.nan  # unit diagonal is never read
    B = datagen.complex_matrix(m + pad, n, name="B", seed=seed)
    B[m:] = -9.0 + 4.0j
    return A, B


@pytest.mark.parametrize("impl", [None, "ctrsm_recursive_inv64"], ids=["default", "recursive"])
@pytest.mark.parametrize("v", VARIANTS, ids=vid)
@pytest.mark.parametrize("shape", [(5, 3), (64, 64), (65, 17), (200, 130), (600, 40), (70, 520), (520, 70), (700, 33)],
                         ids=lambda s: f"{s[0]}x{s[1]}")
def test_ctrsm_variants_vs_fp64(h, v, shape, impl):
    side, uplo, trans, diag = v
    m, n = shape
    h.set_implementation(impl)
    A, B = _problem(side, uplo, diag, m, n)
    alpha = -0.75 + 0.4j
    X64 = oracle.ztrsm(side, uplo, trans, diag, m, n, alpha, A, B)
    Xc = B.copy(order="F")
    oracle.ctrsm(side, uplo, trans, diag, m, n, alpha, A, Xc)  # the reference's float path
    X = B.copy(order="F")
    ib.check(ib.iclblasCtrsm(h.value, side, uplo, trans, diag, m, n, alpha, A, A.shape[0], X, X.shape[0]))
    assert np.all(np.isfinite(X[:m]))
    assert np.all(X[m:] == np.complex64(-9.0 + 4.0j))
    scale = np.abs(X64).max()
    err_gpu = np.abs(X[:m] - X64).max() / scale
    err_cpu = np.abs(Xc[:m] - X64).max() / scale
    order = m if side == 0 else n
    # within a small multiple of the reference's own float error, or of order*u
    assert err_gpu <= max(8 * err_cpu, 8 * order * 2.0 ** -24), (err_gpu, err_cpu)
    assert h.last_implementation() == ("ctrsm_diag64" if order <= 64 else impl or "ctrsm_panel256_inv")


@pytest.mark.parametrize("v", VARIANTS[::5], ids=vid)
def test_ctrsm_alpha_zero_sets_zero_without_reading(h, v):
    side, uplo, trans, diag = v
    m, n = 70, 9
    order = m if side == 0 else n
    A = np.full((order, order), np.nan, np.complex64, order="F")
    B = np.full((m + 1, n), np.nan, np.complex64, order="F")
    B[m:] = 5.0 - 1.0j
    ib.check(ib.iclblasCtrsm(h.value, side, uplo, trans, diag, m, n, 0.0, A, order, B, m + 1))
    assert np.all(B[:m] == 0) and np.all(B[m:] == np.complex64(5.0 - 1.0j))


@pytest.mark.parametrize("trans", [0, 2], ids=["N", "C"])
def test_ctrsm_1024x512(h, trans):
    """Left / lower, 1024 unknowns, 512 right-hand sides; forward error against
    the complex FP64 substitution on a column sample."""
    m, n = 1024, 512
    A = datagen.complex_triangular(m, lower=True, seed=2, poison=True)
    B = datagen.complex_matrix(m, n, name="B", seed=2)
    cols = np.random.default_rng(2).choice(n, 24, replace=False)
    Bs = np.asfortranarray(B[:, cols])
    X64 = oracle.ztrsm(0, 1, trans, 0, m, len(cols), 1.0, A, Bs)
    X = B.copy(order="F")
    ib.check(ib.iclblasCtrsm(h.value, 0, 1, trans, 0, m, n, 1.0, A, m, X, m))
    rel = np.linalg.norm(X[:, cols] - X64) / np.linalg.norm(X64)
    assert rel <= 1e-5, rel
