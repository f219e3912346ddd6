"""The GEMV FP64 oracle (oracle/level2.c) pinned on the CPU: against the
reference's 7 literal CGEMV known answers (tests/iclBLAS/test_Cgemv.cpp, 4 ULP
per part), against the 8 SGEMV expectations the reference computes with its
own FP32 loop (test_Sgemv.cpp, EXPECT_NEAR 1e-5 / EXPECT_FLOAT_EQ), against
numpy closed forms for every op and stride, and with planted mistakes."""
import numpy as np
import pytest

import oracle
from blas_helpers import complex_eq, float_eq
from gemv_helpers import KATS, expected, kat_arrays, random_problem, scalars


def test_fixture_counts():
    assert sum(c["routine"] == "Sgemv" for c in KATS) == 8 and sum(c["routine"] == "Cgemv" for c in KATS) == 7


def _oracle_full(c, A, x, y):
    al, be = scalars(c)
    res, _ = oracle.zgemv(c["trans"], c["m"], c["n"], al, A, x, c["incx"], be, y, c["incy"])
    out = np.array(y, np.complex128 if np.iscomplexobj(A) else np.float64)
    out[:: c["incy"]][: len(res)] = res
    return out


@pytest.mark.parametrize("case", KATS, ids=lambda c: f"{c['routine']}-{c['name']}")
def test_oracle_meets_reference_kats(case):
    A, x, y = kat_arrays(case)
    out = _oracle_full(case, A, x, y)
    ref = expected(case, "reference_expected")
    if case["routine"] == "Cgemv":
        got = out.astype(np.complex64)
        assert all(complex_eq(e, g) for e, g in zip(ref.astype(np.complex64), got))
    elif case["compare"] == "near_1e-5":
        assert np.abs(out - ref).max() <= 1e-5
    else:
        got = out.astype(np.float32)
        assert all(float_eq(e, g) for e, g in zip(ref.astype(np.float32), got))
    np.testing.assert_array_equal(expected(case, "contract_expected"), out)


@pytest.mark.parametrize("cplx", [False, True], ids=["S", "C"])
@pytest.mark.parametrize("trans", [0, 1, 2], ids=["N", "T", "C"])
@pytest.mark.parametrize("incs", [(1, 1), (3, 2)], ids=["inc1", "inc32"])
def test_oracle_matches_closed_form(cplx, trans, incs):
    rng = np.random.default_rng(5)
    m, n = 13, 9
    A, x, y, al, be = random_problem(cplx, trans, m, n, *incs, rng, poison=False)
    res, mag = oracle.zgemv(trans, m, n, al, A, x, incs[0], be, y, incs[1])
    Am = A[:m].astype(np.complex128)
    op = Am if trans == 0 else (Am.conj().T if (trans == 2 and cplx) else Am.T)
    ref = al * (op @ x[:: incs[0]].astype(np.complex128)) + be * y[:: incs[1]].astype(np.complex128)
    np.testing.assert_allclose(res, ref if cplx else ref.real, rtol=1e-13, atol=1e-13)
    assert np.all(mag >= 0)


def test_oracle_catches_planted_mistakes():
    rng = np.random.default_rng(9)
    A, x, y, al, be = random_problem(True, 2, 8, 6, 1, 1, rng, poison=False)
    res, _ = oracle.zgemv(2, 8, 6, al, A, x, 1, be, y, 1)
    no_conj = al * (A[:8].astype(np.complex128).T @ x.astype(np.complex128)) + be * y.astype(np.complex128)
    assert not np.allclose(res, no_conj)
    res_n, _ = oracle.zgemv(0, 8, 6, al, A, np.resize(x, 6), 1, be, np.resize(y, 8), 1)
    assert res_n.shape == (8,) and res.shape == (6,)


def test_oracle_beta_zero_does_not_read_y():
    rng = np.random.default_rng(2)
    A, x, y, al, be = random_problem(False, 1, 20, 7, 2, 3, rng, beta=0.0)
    res, _ = oracle.zgemv(1, 20, 7, al, A, x, 2, 0.0, y, 3)
    assert np.all(np.isfinite(res))
