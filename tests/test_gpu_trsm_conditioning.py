"""STRSM / CTRSM on ill-conditioned triangles through the C ABI.

The default solves apply explicit inverses of the 256 x 256 diagonal panels, which
can amplify rounding where substitution does not. Each case is compared with the
FP64 substitution of the oracle (the truth) and with the oracle's restatement of
the reference's float substitution (Strsm_naive.cl:18-126, Ctrsm_naive.cl:16-183),
and must stay within 8x of that substitution's error (floor 2^-18). Families at
order 1024 (built lower, transposed for the upper triangle), their scales set so
that the float substitution's error lies between about 1e-5 and 3e-2 -- ill
conditioned, yet inside what FP32 can answer (stronger coupling drives every
FP32 method, substitution included, to errors of order 1, or overflows FP64):
  - coupled: off-diagonals c U[-1, 1]/sqrt(order), c = 10 (real), 7 (real, unit
    diagonal; complex), 5 (complex, unit diagonal), diagonal magnitudes U[1, 2];
    compatible right-hand sides B = op(A) X (left) or X op(A) (right), X ~ U[-1, 1],
    where an explicit inverse meets the most cancellation;
  - coupled_random: the same matrices with random right-hand sides;
  - graded4: diagonal magnitudes 2^-4U[0, 1], off-diagonals f U[-1, 1]/sqrt(order)
    (f = 1 real, 0.5 complex), compatible right-hand sides;
  - graded10: diagonal magnitudes 2^-10U[0, 1], off-diagonals f U[-1, 1]/order,
    compatible right-hand sides.
The refined STRSM (FP64 residual, at most 3 sweeps) must reach 2^-20 on the real
coupled and graded4 systems, whose condition numbers keep cond * 2^-24 < 1."""
import numpy as np
import pytest

import oracle
from clgpu_amd import iclblas as ib

pytestmark = pytest.mark.gpu
ORDER, LINES = 1024, 96
FAMILIES = ["coupled", "coupled_random", "graded4", "graded10"]
S_VARIANTS = [(s, u, t, d) for s in (0, 1) for u in (0, 1) for t in (0, 1) for d in (0, 1)]
C_VARIANTS = [(s, u, t, d) for s in (0, 1) for u in (0, 1) for t in (0, 1, 2) for d in (0, 1)]
FLOOR = 2.0 ** -18


def vid(v):
    s, u, t, d = v
    return f"{'LR'[s]}{'UL'[u]}{'NTC'[t]}{'NU'[d]}"


def _lower(family, n, cplx, unit, rng):
    def u(shape):
        x = rng.uniform(-1, 1, shape)
        return x + 1j * rng.uniform(-1, 1, shape) if cplx else x
    if family.startswith("coupled"):
        c = (5.0 if unit else 7.0) if cplx else (7.0 if unit else 10.0)
        scale = c / np.sqrt(n)
        mag = rng.uniform(1, 2, n)
    elif family == "graded4":
        scale = (0.5 if cplx else 1.0) / np.sqrt(n)
        mag = 2.0 ** (-4 * rng.uniform(0, 1, n))
    else:
        scale = (0.5 if cplx else 1.0) / n
        mag = 2.0 ** (-10 * rng.uniform(0, 1, n))
    L = np.tril(u((n, n)), -1) * scale
    phase = np.exp(1j * rng.uniform(0, 2 * np.pi, n)) if cplx else rng.choice([-1.0, 1.0], n)
    L[np.arange(n), np.arange(n)] = mag * phase
    return L


def _problem(family, v, cplx, seed):
    side, uplo, trans, diag = v
    rng = np.random.default_rng(seed)
    m, n = (ORDER, LINES) if side == 0 else (LINES, ORDER)
    dt = np.complex64 if cplx else np.float32
    L = _lower(family, ORDER, cplx, diag == 1, rng)
    T = L if uplo == 1 else L.T.copy()
    A = np.asfortranarray(T.astype(dt))
    A[~(np.tril(np.ones((ORDER, ORDER), bool)) if uplo == 1 else np.triu(np.ones((ORDER, ORDER), bool)))] = np.nan
    # op(A) in FP64 as the solver sees it (stored triangle, unit diagonal)
    opa = np.nan_to_num(A.astype(np.complex128 if cplx else np.float64))
    if diag == 1:
        opa[np.arange(ORDER), np.arange(ORDER)] = 1.0
    opa = opa if trans == 0 else (opa.T if trans == 1 else opa.conj().T)
    if family == "coupled_random":
        Bv = rng.uniform(-1, 1, (m, n)) + (1j * rng.uniform(-1, 1, (m, n)) if cplx else 0)
    else:
        Xt = rng.uniform(-1, 1, (m, n)) + (1j * rng.uniform(-1, 1, (m, n)) if cplx else 0)
        Bv = opa @ Xt if side == 0 else Xt @ opa
    B = np.asfortranarray(Bv.astype(dt))
    return A, B, m, n


def _errors(v, A, B, m, n, X, cplx, alpha=1.0):
    side, uplo, trans, diag = v
    X64 = (oracle.ztrsm if cplx else oracle.dtrsm)(side, uplo, trans, diag, m, n, alpha, A, B)
    Xs = B.copy(order="F")
    (oracle.ctrsm if cplx else oracle.strsm)(side, uplo, trans, diag, m, n, alpha, A, Xs)
    scale = np.abs(X64).max()
    return np.abs(X - X64).max() / scale, np.abs(Xs - X64).max() / scale


@pytest.fixture
def h(handle):
    handle.set_math_mode(ib.MATH_FP32)
    handle.set_implementation(None)
    yield handle
    handle.set_implementation(None)


@pytest.mark.parametrize("family", FAMILIES)
@pytest.mark.parametrize("v", S_VARIANTS, ids=vid)
def test_strsm_ill_conditioned_default(h, family, v):
    A, B, m, n = _problem(family, v, False, seed=3)
    X = B.copy(order="F")
    ib.check(ib.iclblasStrsm(h.value, *v, m, n, 1.0, A, ORDER, X, m))
    e_gpu, e_sub = _errors(v, A, B, m, n, X, False)
    assert e_gpu <= max(8 * e_sub, FLOOR), (h.last_implementation(), e_gpu, e_sub)


@pytest.mark.parametrize("impl", ["strsm_panel256_fused", "strsm_panel256_inv", "strsm_panel256",
                                  "strsm_blocked_inv64", "strsm_recursive"])
@pytest.mark.parametrize("family", FAMILIES)
def test_strsm_ill_conditioned_implementations(h, impl, family):
    v = (0, 1, 0, 0)
    h.set_implementation(impl)
    A, B, m, n = _problem(family, v, False, seed=5)
    X = B.copy(order="F")
    ib.check(ib.iclblasStrsm(h.value, *v, m, n, 1.0, A, ORDER, X, m))
    assert h.last_implementation() == impl
    e_gpu, e_sub = _errors(v, A, B, m, n, X, False)
    assert e_gpu <= max(8 * e_sub, FLOOR), (impl, e_gpu, e_sub)


@pytest.mark.parametrize("family", ["coupled", "graded4"])
@pytest.mark.parametrize("v", [(0, 1, 0, 0), (1, 0, 1, 0)], ids=vid)
def test_strsm_ill_conditioned_refined(h, family, v):
    # cond * 2^-24 < 1 here, so FP64-residual refinement converges
    A, B, m, n = _problem(family, v, False, seed=7)
    X = B.copy(order="F")
    ib.check(ib.iclblasStrsmRefined(h.value, *v, m, n, 1.0, A, ORDER, X, m, 3))
    e_ref, e_sub = _errors(v, A, B, m, n, X, False)
    assert e_ref <= 2.0 ** -20, (e_ref, e_sub, h.last_refinement_sweeps())


@pytest.mark.parametrize("family", FAMILIES)
@pytest.mark.parametrize("v", C_VARIANTS, ids=vid)
def test_ctrsm_ill_conditioned_default(h, family, v):
    A, B, m, n = _problem(family, v, True, seed=11)
    X = B.copy(order="F")
    ib.check(ib.iclblasCtrsm(h.value, *v, m, n, 1.0, A, ORDER, X, m))
    e_gpu, e_sub = _errors(v, A, B, m, n, X, True)
    assert e_gpu <= max(8 * e_sub, FLOOR), (h.last_implementation(), e_gpu, e_sub)
