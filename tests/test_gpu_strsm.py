"""STRSM parity on the GPU through the C ABI: the reference's known answers
(tests/iclBLAS/test_Strsm.cpp, ADAPTIVE_FLOAT_EQ), all 16 side/uplo/trans/diag
variants against the FP64 substitution of the oracle, never-read triangles
and unit diagonals poisoned with NaN, alpha == 0, padded B, and the
refined (FP64-residual) variant at the 4096 x 1024 configuration."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from blas_helpers import adaptive_ok, flat, fmat
from clgpu_amd import datagen
from clgpu_amd import iclblas as ib

pytestmark = pytest.mark.gpu
KATS = json.loads((Path(__file__).parent / "golden" / "kats.json").read_text())
VARIANTS = [(s, u, t, d) for s in (0, 1) for u in (0, 1) for t in (0, 1) for d in (0, 1)]


def vid(v):
    s, u, t, d = v
    return f"{'LR'[s]}{'UL'[u]}{'NT'[t]}{'NU'[d]}"


@pytest.fixture
def h(handle):
    handle.set_math_mode(ib.MATH_FP32)
    handle.set_implementation(None)
    yield handle
    handle.set_implementation(None)


@pytest.mark.parametrize("case", KATS["strsm"], ids=lambda c: c["name"])
def test_strsm_kat(h, case):
    c = case
    order = c["m"] if c["side"] == 0 else c["n"]
    A = fmat(c["A"], c["lda"], order)
    B = fmat(c["B"], c["ldb"], c["n"])
    ib.check(ib.iclblasStrsm(h.value, c["side"], c["uplo"], c["trans"], c["diag"], c["m"], c["n"], c["alpha"], A,
                             c["lda"], B, c["ldb"]))
    assert adaptive_ok(c["expected"], flat(B)), (flat(B), c["expected"])


def _problem(side, uplo, diag, m, n, seed=41, pad=3):
    order = m if side == 0 else n
    A = datagen.triangular(order, lower=bool(uplo), ld=order + 2, seed=seed, poison=True)
    if diag == 1:
        A[np.arange(order), np.arange(order)] = np.nan  # unit diagonal is never read
    B = datagen.matrix(m, n, m + pad, "B", seed=seed)
    B[m:] = -9.0
    return A, B


@pytest.mark.parametrize("impl", ["strsm_blocked_inv64", "strsm_dataflow64", "strsm_panel256", "strsm_panel256_inv",
                                  "strsm_panel256_fused",
                                  "strsm_recursive"])
@pytest.mark.parametrize("v", VARIANTS, ids=vid)
@pytest.mark.parametrize("shape", [(5, 3), (64, 64), (65, 17), (200, 130), (300, 77), (600, 40), (70, 520),
                                   (512, 33)],
                         ids=lambda s: f"{s[0]}x{s[1]}")
def test_strsm_variants_vs_fp64(h, impl, v, shape):
    side, uplo, trans, diag = v
    m, n = shape
    h.set_implementation(impl)
    A, B = _problem(side, uplo, diag, m, n)
    alpha = -0.75
    X64 = oracle.dtrsm(side, uplo, trans, diag, m, n, alpha, A, B)
    Xc = B.copy(order="F")
    oracle.strsm(side, uplo, trans, diag, m, n, alpha, A, Xc)  # reference float path
    X = B.copy(order="F")
    ib.check(ib.iclblasStrsm(h.value, side, uplo, trans, diag, m, n, alpha, A, A.shape[0], X, X.shape[0]))
    assert np.all(np.isfinite(X[:m]))
    assert np.all(X[m:] == -9.0)
    scale = np.abs(X64).max()
    err_gpu = np.abs(X[:m] - X64).max() / scale
    err_cpu = np.abs(Xc[:m] - X64).max() / scale
    order = m if side == 0 else n
    # within a small multiple of the reference's own float error, or of order*u
    assert err_gpu <= max(8 * err_cpu, 4 * order * 2.0 ** -24), (err_gpu, err_cpu)
    expect = "strsm_diag64" if order <= 64 else impl
    assert h.last_implementation() == expect


@pytest.mark.parametrize("v", VARIANTS[::3], ids=vid)
def test_strsm_alpha_zero_sets_zero_without_reading(h, v):
    side, uplo, trans, diag = v
    m, n = 70, 9
    order = m if side == 0 else n
    A = np.full((order, order), np.nan, np.float32, order="F")
    B = np.full((m + 1, n), np.nan, np.float32, order="F")
    B[m:] = 5.0
    ib.check(ib.iclblasStrsm(h.value, side, uplo, trans, diag, m, n, 0.0, A, order, B, m + 1))
    assert np.all(B[:m] == 0.0) and np.all(B[m:] == 5.0)


@pytest.mark.parametrize("v", VARIANTS, ids=vid)
@pytest.mark.parametrize("shape", [(190, 45), (45, 190), (300, 130)], ids=lambda s: f"{s[0]}x{s[1]}")
def test_strsm_refined_variants(h, v, shape):
    side, uplo, trans, diag = v
    m, n = shape
    A, B = _problem(side, uplo, diag, m, n, seed=43)
    alpha = 1.25
    X64 = oracle.dtrsm(side, uplo, trans, diag, m, n, alpha, A, B)
    X = B.copy(order="F")
    ib.check(ib.iclblasStrsmRefined(h.value, side, uplo, trans, diag, m, n, alpha, A, A.shape[0], X, X.shape[0], 2))
    assert np.all(X[m:] == -9.0)
    rel = np.linalg.norm(X[:m] - X64) / np.linalg.norm(X64)
    assert rel <= 2.0 ** -23, rel


@pytest.mark.parametrize("v", VARIANTS, ids=vid)
def test_strsm_refined_aligned_variants(h, v):
    """Refinement on 320 x 192 with leading dimensions that are multiples of 4:
    the FP64 residual runs on its LDS-DMA kernel (M and N multiples of 64, K of
    32), the triangle rule applied at the fragment reads of the slabs that cross
    the diagonal; never-read parts of A are NaN."""
    side, uplo, trans, diag = v
    m, n = 320, 192
    order = m if side == 0 else n
    A = datagen.triangular(order, lower=bool(uplo), ld=order + 4, seed=44, poison=True)
    if diag == 1:
        A[np.arange(order), np.arange(order)] = np.nan
    B = datagen.matrix(m, n, m + 4, "B", seed=44)
    B[m:] = -9.0
    alpha = -0.75
    X64 = oracle.dtrsm(side, uplo, trans, diag, m, n, alpha, A, B)
    X = B.copy(order="F")
    ib.check(ib.iclblasStrsmRefined(h.value, side, uplo, trans, diag, m, n, alpha, A, A.shape[0], X, X.shape[0], 3))
    assert np.all(X[m:] == -9.0)
    rel = np.linalg.norm(X[:m] - X64) / np.linalg.norm(X64)
    assert rel <= 2.0 ** -23, rel
    assert h.last_refinement_sweeps() >= 1


def test_strsm_refined_sweep_count_resets(h):
    """A refined call with no sweeps (or alpha == 0) reports 0 sweeps, whatever
    the handle's previous call ran."""
    m, n = 190, 45
    A, B = _problem(0, 1, 0, m, n, seed=46)
    X = B.copy(order="F")
    ib.check(ib.iclblasStrsmRefined(h.value, 0, 1, 0, 0, m, n, 1.0, A, A.shape[0], X, X.shape[0], 2))
    assert h.last_refinement_sweeps() >= 1
    for sweeps, alpha in ((0, 1.0), (2, 0.0)):
        X = B.copy(order="F")
        ib.check(ib.iclblasStrsmRefined(h.value, 0, 1, 0, 0, m, n, alpha, A, A.shape[0], X, X.shape[0], sweeps))
        assert h.last_refinement_sweeps() == 0


def _to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a.T)).cuda()


def test_strsm_config4_plain_and_refined(dev_handle):
    """STRSM L/Lo/N/NonUnit m=4096 n=1024; forward error vs FP64 on a column sample."""
    import torch
    m, n = 4096, 1024
    A = datagen.triangular(m, lower=True, seed=0, poison=True)
    B = datagen.matrix(m, n, name="B", seed=0)
    cols = np.random.default_rng(0).choice(n, 48, replace=False)
    Bs = np.asfortranarray(B[:, cols])
    X64 = oracle.dtrsm(0, 1, 0, 0, m, len(cols), 1.0, A, Bs)
    h = dev_handle
    dA = _to_dev(A)
    res, ran = {}, {}
    for sweeps in (0, 1, 3):
        dB = _to_dev(B)
        ib.check(ib.iclblasStrsmRefined(h.value, 0, 1, 0, 0, m, n, 1.0, dA.data_ptr(), m, dB.data_ptr(), m, sweeps))
        torch.cuda.synchronize()
        ran[sweeps] = h.last_refinement_sweeps()
        X = dB.t().cpu().numpy()[:, cols].astype(np.float64)
        assert np.all(np.isfinite(X))
        res[sweeps] = np.linalg.norm(X - X64) / np.linalg.norm(X64)
    assert res[0] <= 1e-5, res
    assert res[3] <= 2.0 ** -23, res  # FP32 accuracy
    assert res[3] <= res[0]
    # sweeps stop once a correction no longer changes X at FP32 resolution
    assert ran[0] == 0 and ran[1] == 1 and 1 <= ran[3] <= 3, ran
