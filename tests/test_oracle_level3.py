"""The level-3 FP64 oracle (oracle/level3.c) pinned on the CPU:
  - against the reference's literal known answers (tests/iclBLAS/test_Ssymm.cpp,
    test_Ssyrk.cpp, test_Ssyr2k.cpp, transcribed into tests/golden/kats_l3.json)
    with the reference's EXPECT_FLOAT_EQ (4 ULP);
  - against the reference's loop-built 256x256 SSYMM case with its EXPECT_NEAR
    1e-4 (test_Ssymm.cpp:236-277);
  - against numpy closed forms of the header contract for every routine and
    side / uplo / trans / diag variant on random data, including the known
    answers whose expectation the reference computes with inline loops;
  - with planted mistakes (a dropped conjugation, the wrong triangle, a unit
    diagonal read) that the same checks must catch.
For the routines whose reference expectation is not the header contract
(Strmm, Ctrmm, Cherk, Cher2k) the tests restate what the reference kernels
compute and show that their known answers follow that, not the header.
"""
import numpy as np
import pytest

import oracle
from blas_helpers import complex_eq, float_eq
from l3_helpers import (COMPLEX, KATS, RANK2K, RANKK, SYMM, TRMM, closed_form, expected_matrix, kat_case,
                        oracle_result, random_case)


def _rounded(X, cplx):
    return X.astype(np.complex64) if cplx else X.astype(np.float32)


LITERAL = [c for c in KATS if c["reference_expected"] is not None and not c["deviation"]]
INLINE = [c for c in KATS if c["reference_expected"] is None and not c["deviation"] and not c.get("generator")]
DEVIATION = [c for c in KATS if c["deviation"]]


def test_fixture_counts():
    counts = {}
    for c in KATS:
        counts[c["routine"]] = counts.get(c["routine"], 0) + 1
    assert counts == {"Ssymm": 5, "Csymm": 4, "Chemm": 4, "Ssyrk": 4, "Csyrk": 6, "Cherk": 7, "Ssyr2k": 4,
                      "Csyr2k": 6, "Cher2k": 7, "Strmm": 11, "Ctrmm": 8}


@pytest.mark.parametrize("case", LITERAL, ids=lambda c: f"{c['routine']}-{c['name']}")
def test_oracle_meets_reference_literal_kats(case):
    c = kat_case(case)
    out, _, _ = oracle_result(c, c["A"], c.get("B"), c["C"])
    got = _rounded(out, False).T.reshape(-1)
    exp = np.asarray(case["reference_expected"], np.float32)
    assert all(float_eq(e, g) for e, g in zip(exp, got)), (exp, got)


def test_oracle_meets_loop_built_ssymm_256():
    case = next(c for c in KATS if c.get("generator") == "ssymm_256")
    c = kat_case(case)
    out, _, _ = oracle_result(c, c["A"], c["B"], c["C"])
    # the reference's expectation: double loop over the upper triangle (test_Ssymm.cpp:188-232)
    A = c["A"].astype(np.float64)
    S = np.triu(A) + np.triu(A, 1).T
    exp = 0.9000000357627869 * (S @ c["B"].astype(np.float64)) + (-1.2000000476837158) * c["C"].astype(np.float64)
    assert np.abs(out.astype(np.float32) - exp.astype(np.float32)).max() <= 1e-4


@pytest.mark.parametrize("case", INLINE + LITERAL, ids=lambda c: f"{c['routine']}-{c['name']}")
def test_oracle_matches_closed_form_on_kats(case):
    c = kat_case(case)
    out, _, _ = oracle_result(c, c["A"], c.get("B"), c["C"])
    ref = closed_form(c, c["A"], c.get("B"), c["C"])
    assert np.allclose(out, ref, rtol=1e-12, atol=1e-9)
    # the stored contract expectation is the same oracle result
    np.testing.assert_allclose(expected_matrix(case, "contract_expected"), _rounded(out, c["routine"] in COMPLEX),
                               rtol=0, atol=0)


def _variants(routine):
    if routine in SYMM:
        return [dict(side=s, uplo=u) for s in (0, 1) for u in (0, 1)]
    if routine in RANKK or routine in RANK2K:
        return [dict(uplo=u, trans=t) for u in (0, 1) for t in (0, 1, 2)]
    return [dict(side=s, uplo=u, trans=t, diag=d) for s in (0, 1) for u in (0, 1) for t in (0, 1, 2) for d in (0, 1)]


ALL = [(r, v) for r in SYMM + RANKK + RANK2K + TRMM for v in _variants(r)]


@pytest.mark.parametrize("routine,v", ALL, ids=lambda x: x if isinstance(x, str) else "-".join(map(str, x.values())))
def test_oracle_matches_closed_form_random(routine, v):
    rng = np.random.default_rng(7)
    c = random_case(routine, rng, m=7, n=5, k=4, **v)
    out, mag, k = oracle_result(c, c["A"], c["B"], c["C"])
    ref = closed_form(c, c["A"], c["B"], c["C"])
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12)
    assert np.all(np.isfinite(out[: out.shape[0] - 3]) | np.isnan(c["C"][: out.shape[0] - 3]))
    assert np.all(mag >= 0)


def test_oracle_catches_planted_mistakes():
    rng = np.random.default_rng(3)
    # hermitian mirror without the conjugation
    c = random_case("Chemm", rng, m=6, n=4, side=0, uplo=1, poison=False)
    out, _, _ = oracle_result(c, c["A"], c["B"], c["C"])
    wrong = dict(c, routine="Csymm")
    assert not np.allclose(out, closed_form(wrong, c["A"], c["B"], c["C"]))
    # the other triangle of C
    c = random_case("Ssyrk", rng, n=6, k=3, uplo=0, trans=0)
    out, _, _ = oracle_result(c, c["A"], None, c["C"])
    assert not np.allclose(out, closed_form(dict(c, uplo=1), c["A"], None, c["C"]), equal_nan=True)
    # a unit diagonal taken from A
    c = random_case("Strmm", rng, m=5, n=4, side=0, uplo=1, trans=0, diag=1, poison=False)
    out, _, _ = oracle_result(c, c["A"], c["B"], c["C"])
    assert not np.allclose(out, closed_form(dict(c, diag=0), c["A"], c["B"], c["C"]))
    # herk without the conjugate transpose
    c = random_case("Cherk", rng, n=5, k=3, uplo=1, trans=0)
    out, _, _ = oracle_result(c, c["A"], None, c["C"])
    assert not np.allclose(out, closed_form(dict(c, routine="Csyrk"), c["A"], None, c["C"]))


def test_oracle_never_reads_poisoned_parts():
    rng = np.random.default_rng(11)
    for routine, v in [("Ssymm", dict(side=1, uplo=0)), ("Chemm", dict(side=0, uplo=1)),
                       ("Strmm", dict(side=0, uplo=1, trans=2, diag=1)), ("Ctrmm", dict(side=1, uplo=0, trans=2, diag=1)),
                       ("Ssyrk", dict(uplo=1, trans=1)), ("Cher2k", dict(uplo=0, trans=2))]:
        c = random_case(routine, rng, m=9, n=6, k=5, beta=0.0, **v)
        out, _, _ = oracle_result(c, c["A"], c["B"], c["C"])
        rc = out.shape[0] - 3
        if routine in RANKK or routine in RANK2K:
            tri = np.tril(np.ones((rc, rc), bool)) if v["uplo"] == 1 else np.triu(np.ones((rc, rc), bool))
            assert np.all(np.isfinite(out[:rc][tri]))
        else:
            assert np.all(np.isfinite(out[:rc]))


# ---------------------------------------------------------------- deviations
def _strmm_as_built(c):
    """What Strmm_naive.cl:16-232 computes (for lda = ldb = ldc = m = n, as in
    all of the reference's known answers): every output element in the uplo
    triangle of C *positions* gets alpha * sum_k op_full(A)(row, k) * B(k, col)
    (left) or alpha * sum_k B(row, k) * op_full(A)(k, col) (right) over the
    whole square A, with the unit diagonal replaced by 1 on the kernel's own
    index test; elements outside that triangle keep their input value (the
    tests pass B as C)."""
    m, n, alpha = c["m"], c["n"], c["alpha"]
    A, B = c["A"].astype(np.float32), c["B"].astype(np.float32)
    out = c["C"].copy()
    K = m if c["side"] == 0 else n
    for row in range(m):
        for col in range(n):
            inside = col >= row if c["uplo"] == 0 else row >= col
            if not inside:
                continue
            acc = np.float32(0)
            for k in range(K):
                if c["side"] == 0:
                    if c["trans"] == 0:
                        unit = k == row
                        a = A[row, k]
                    else:
                        unit = (k == col) if c["uplo"] == 0 else (k == row and col == row)
                        a = A[k, row]
                    b = B[k, col]
                    acc += np.float32(alpha) * (np.float32(1) if (unit and c["diag"] == 1) else a) * b
                else:
                    unit = k == col
                    a = A[k, col] if c["trans"] == 0 else A[col, k]
                    acc += np.float32(alpha) * B[row, k] * (np.float32(1) if (unit and c["diag"] == 1) else a)
            out[row, col] = acc
    return out


@pytest.mark.parametrize("case", [c for c in DEVIATION if c["routine"] == "Strmm"], ids=lambda c: c["name"])
def test_strmm_reference_kats_follow_the_kernel_not_the_header(case):
    c = kat_case(case)
    ref = expected_matrix(case, "reference_expected")
    built = _strmm_as_built(c)
    assert all(float_eq(e, g) for e, g in zip(ref.T.reshape(-1), built.T.reshape(-1))), (ref, built)
    contract = expected_matrix(case, "contract_expected")
    assert np.abs(ref - contract).max() > 1.0  # the header's C = alpha op(A) B differs


def _cherk_as_built(c):
    """Cherk_naive.cl:16-55 (and the reference test's own loop): the value of
    every (i, j) is sum_l |A(l, i)|^2, computed from column i of A only."""
    A, C = c["A"].astype(np.complex128), c["C"].astype(np.complex128)
    n, k = c["n"], c["k"]
    out = C.copy()
    for i in range(n):
        v = sum(abs(A[l, i]) ** 2 for l in range(k))
        for j in range(n):
            if i == j:
                out[i, i] = out[i, i].real
            if (c["uplo"] == 0 and j >= i) or (c["uplo"] == 1 and j <= i):
                out[i, j] = c["alpha"] * v + c["beta"] * out[i, j]
    return out


@pytest.mark.parametrize("case", [c for c in DEVIATION if c["routine"] == "Cherk"], ids=lambda c: c["name"])
def test_cherk_reference_expectation_is_not_the_header(case):
    c = kat_case(case)
    built = _cherk_as_built(c)
    contract = expected_matrix(case, "contract_expected")
    assert np.abs(built - contract).max() > 1.0


@pytest.mark.parametrize("case", [c for c in DEVIATION if c["routine"] == "Ctrmm"], ids=lambda c: c["name"])
def test_ctrmm_reference_kats_differ_from_the_header(case):
    ref = expected_matrix(case, "reference_expected")
    contract = expected_matrix(case, "contract_expected")
    assert np.abs(ref - contract).max() > 1.0
