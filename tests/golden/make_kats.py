"""Writes tests/golden/kats.json: the known-answer tests of the reference's
API-level suites, transcribed as data.

Sources (reference tree, read as text, nothing executed):
  tests/iclBLAS/test_Sgemm.cpp:18-287   6 SGEMM cases
  tests/iclBLAS/test_Cgemm.cpp:19-431   8 CGEMM cases
  tests/iclBLAS/test_Strsm.cpp:23-334   8 STRSM cases
  tests/iclBLAS/test_Ctrsm.cpp:36-637  12 CTRSM cases

Inputs and literal expected arrays are copied value for value. Expected
values the reference writes as float expressions (STRSM, e.g. `1.1f*2.f/3.f`)
are evaluated here with float32 arithmetic in C's left-to-right order, so
they are bit-identical to the reference's constants. Expected values the
reference computes with an inline loop (SGEMM 64^3 closed form, the three
hermitian CGEMM cases) are computed exactly in float64; the test then applies
the same comparison the reference uses where that is exact, and the FP64
error bound otherwise (the `compare` field).

Run: python tests/golden/make_kats.py
"""
import json
from pathlib import Path

import numpy as np

f = np.float32


def fl(*vals):
    return [float(f(v)) for v in vals]


def cplx(pairs):
    return [[float(a), float(b)] for a, b in pairs]


def sgemm_cases():
    rep = [1, 2, 3] * 3
    cases = [
        dict(name="ntransAB_3x3", transa=0, transb=0, expected=[7, 14, 21] * 3),
        dict(name="transAB_3x3", transa=1, transb=1, expected=[7, 8, 9, 13, 14, 15, 19, 20, 21]),
        dict(name="ntransA_transB_3x3", transa=0, transb=1, expected=[4, 8, 12, 7, 14, 21, 10, 20, 30]),
        dict(name="transA_ntransB_3x3", transa=1, transb=0, expected=[15, 16, 17] * 3),
    ]
    out = []
    for c in cases:
        out.append(dict(c, m=3, n=3, k=3, alpha=1.0, beta=1.0, lda=3, ldb=3, ldc=3, A=fl(*rep), B=fl(*rep),
                        C=fl(*rep), expected=fl(*c["expected"]), compare="ulp4"))
    out.append(dict(name="ntransAB_ldx_3x3", transa=0, transb=0, m=3, n=3, k=3, alpha=1.0, beta=1.0, lda=4, ldb=5,
                    ldc=4, A=fl(*([1, 2, 3, -1] * 3)), B=fl(*([1, 2, 3, -1, -1] * 3)),
                    C=fl(*([1, 2, 3, -1] * 3)), expected=fl(*([7, 14, 21, -1] * 3)), compare="ulp4"))
    # ntransAB_optim (test_Sgemm.cpp:233-287): A[i] = 1.25 i, B[i] = 1.5 i, C[i] = 1.75 i
    m = n = k = 64
    A = (np.arange(64 * 64) * f(1.25)).astype(f)
    B = (np.arange(64 * 64) * f(1.5)).astype(f)
    C = (np.arange(64 * 64) * f(1.75)).astype(f)
    Af = A.astype(np.float64).reshape(k, m).T
    Bf = B.astype(np.float64).reshape(n, k).T
    Cf = C.astype(np.float64).reshape(n, m).T
    exp = (Af @ Bf + Cf).T.reshape(-1)
    out.append(dict(name="ntransAB_optim", transa=0, transb=0, m=m, n=n, k=k, alpha=1.0, beta=1.0, lda=64, ldb=64,
                    ldc=64, A=A.tolist(), B=B.tolist(), C=C.tolist(), expected=exp.tolist(), compare="bound"))
    return out


def cgemm_cases():
    rep = [(1, 0), (2, 0), (3, 0)] * 3
    base = [
        dict(name="ntransAB_3x3", transa=0, transb=0, expected=[(7, 0), (14, 0), (21, 0)] * 3),
        dict(name="transAB_3x3", transa=1, transb=1,
             expected=[(7, 0), (8, 0), (9, 0), (13, 0), (14, 0), (15, 0), (19, 0), (20, 0), (21, 0)]),
        dict(name="ntransA_transB_3x3", transa=0, transb=1,
             expected=[(4, 0), (8, 0), (12, 0), (7, 0), (14, 0), (21, 0), (10, 0), (20, 0), (30, 0)]),
        dict(name="transA_ntransB_3x3", transa=1, transb=0, expected=[(15, 0), (16, 0), (17, 0)] * 3),
    ]
    out = []
    for c in base:
        out.append(dict(c, m=3, n=3, k=3, alpha=[1.0, 0.0], beta=[1.0, 0.0], lda=3, ldb=3, ldc=3, A=cplx(rep),
                        B=cplx(rep), C=cplx(rep), expected=cplx(c["expected"]), compare="complex_ulp4"))
    pad = [(1, 0), (2, 0), (3, 0), (-1, 0)] * 3
    out.append(dict(name="ntransAB_ldx_3x3", transa=0, transb=0, m=3, n=3, k=3, alpha=[1.0, 0.0], beta=[1.0, 0.0],
                    lda=4, ldb=5, ldc=4, A=cplx(pad), B=cplx([(1, 0), (2, 0), (3, 0), (-1, 0), (-1, 0)] * 3),
                    C=cplx(pad), expected=cplx([(7, 0), (14, 0), (21, 0), (-1, 0)] * 3), compare="complex_ulp4"))
    A = [(1, 0), (2, -1), (3, 5), (2, 1), (2, 0), (3, 1), (3, -5), (3, 1), (6, 0)]
    B = [(1, 0), (1, 1), (3, 5), (1, -1), (2, 0), (3, -2), (3, -5), (3, 2), (3, 0)]
    Creal = [(1, 0), (2, 0), (3, 0)] * 3
    Cnc = [(1, -1), (2, 5), (3, -8), (1, 2), (2, -18), (3, 1), (1, 15), (2, 7), (3, -3)]
    Az = np.array([complex(*p) for p in A])
    Bz = np.array([complex(*p) for p in B])

    def expected(transa, transb, Cl):
        Cz = np.array([complex(*p) for p in Cl])
        res = np.zeros(9, complex)
        for i in range(3):
            for j in range(3):
                v = 0
                for l in range(3):
                    a = np.conj(Az[i * 3 + l]) if transa == 2 else Az[l * 3 + i]
                    b = np.conj(Bz[l * 3 + j]) if transb == 2 else Bz[j * 3 + l]
                    v += a * b
                res[j * 3 + i] = v + Cz[j * 3 + i]
        return [[r.real, r.imag] for r in res]

    # The reference's inline loops (test_Cgemm.cpp:266-281, 329-344, 392-407):
    #   CC: conj(A[i*lda+l]) * conj(B[l*ldb+j]);  CN: conj(A[i*lda+l]) * B[j*ldb+l];
    #   NC: A[l*lda+i] * conj(B[l*ldb+j]).
    for name, ta, tb, Cl in (("hermitAB_3x3", 2, 2, Creal), ("hermitA_ntransB_3x3", 2, 0, Creal),
                             ("ntransA_hermitB_3x3", 0, 2, Cnc)):
        out.append(dict(name=name, transa=ta, transb=tb, m=3, n=3, k=3, alpha=[1.0, 0.0], beta=[1.0, 0.0], lda=3,
                        ldb=3, ldc=3, A=cplx(A), B=cplx(B), C=cplx(Cl), expected=expected(ta, tb, Cl),
                        compare="complex_ulp4"))
    return out


def strsm_cases():
    # float32 evaluation of the reference's constant expressions
    a11, m11, a13, a01 = f(1.1), f(-1.1), f(1.3), f(0.1)
    m13, m01 = f(-1.3), f(-0.1)

    def e(*parts):  # left-to-right float32 product/quotient chain: e(x, '*', y, '/', z)
        v = f(parts[0])
        for op, x in zip(parts[1::2], parts[2::2]):
            v = f(v * f(x)) if op == "*" else f(v / f(x))
        return float(v)

    A1 = fl(1, -1, -2, 2, 3, -3, 4, 5, 6)
    B1 = fl(7, 8, 9, 10, 11, 12, 13, 14, 15)
    cases = [
        dict(name="3x3_left_up_ntrans_ndiag", side=0, uplo=0, trans=0, diag=0, m=3, n=3, lda=3, ldb=3,
             alpha=float(a11), A=A1, B=B1,
             expected=[e(a11, "*", 2, "/", 3), e(a11, "*", 1, "/", 6), e(a11, "*", 3, "/", 2),
                       e(a11, "*", 4, "/", 3), e(a11, "*", 1, "/", 3), e(a11, "*", 2),
                       e(a11, "*", 2), e(a11, "*", 1, "/", 2), e(a11, "*", 5, "/", 2)]),
        dict(name="3x3_right_up_ntrans_ndiag", side=1, uplo=0, trans=0, diag=0, m=3, n=3, lda=3, ldb=3,
             alpha=float(a11), A=A1, B=B1,
             expected=[e(a11, "*", 7), e(a11, "*", 8), e(a11, "*", 9),
                       e(m11, "*", 4, "/", 3), e(m11, "*", 5, "/", 3), e(m11, "*", 6, "/", 3),
                       e(m11, "*", 25, "/", 18), e(m11, "*", 29, "/", 18), e(m11, "*", 33, "/", 18)]),
        dict(name="3x3_left_low_ntrans_ndiag", side=0, uplo=1, trans=0, diag=0, m=3, n=3, lda=4, ldb=3,
             alpha=float(m11), A=fl(1, 2, 3, -1, -2, 4, 5, -3, -4, -5, 6, -7), B=B1,
             expected=[e(m11, "*", 7), e(a11, "*", 3, "/", 2), e(a11, "*", 9, "/", 12),
                       e(m11, "*", 10), e(a11, "*", 9, "/", 4), e(a11, "*", 27, "/", 24),
                       e(m11, "*", 13), e(a11, "*", 3), e(a11, "*", 9, "/", 6)]),
        # A[6] is written `-  -2.f` in the reference (= +2); it lies in the unused upper triangle.
        dict(name="3x3_right_low_ntrans_ndiag", side=1, uplo=1, trans=0, diag=0, m=3, n=3, lda=3, ldb=4,
             alpha=float(a01), A=fl(1, 2, 3, -1, 4, 5, 2, -3, 6),
             B=fl(7, 8, 9, -1, 10, 11, 12, -2, 13, 14, 15, -3),
             expected=[e(a01, "*", 11, "/", 12), e(a01, "*", 4, "/", 3), e(a01, "*", 7, "/", 4), -1.0,
                       e(m01, "*", 5, "/", 24), e(m01, "*", 1, "/", 6), e(m01, "*", 1, "/", 8), -2.0,
                       e(a01, "*", 13, "/", 6), e(a01, "*", 14, "/", 6), e(a01, "*", 15, "/", 6), -3.0]),
        dict(name="4x3_left_up_trans_ndiag", side=0, uplo=0, trans=1, diag=0, m=4, n=3, lda=4, ldb=4,
             alpha=float(a13), A=fl(1, -1, -2, -3, 2, 3, -4, -5, 4, 5, 6, -6, 7, 8, 9, 10),
             B=fl(11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22),
             expected=[e(a13, "*", 11), e(m13, "*", 10, "/", 3), e(m13, "*", 43, "/", 18), e(m13, "*", 89, "/", 60),
                       e(a13, "*", 15), e(m13, "*", 14, "/", 3), e(m13, "*", 59, "/", 18), e(m13, "*", 121, "/", 60),
                       e(a13, "*", 19), e(m13, "*", 18, "/", 3), e(m13, "*", 75, "/", 18),
                       e(m13, "*", 153, "/", 60)]),
        dict(name="3x4_right_up_trans_ndiag", side=1, uplo=0, trans=2, diag=0, m=3, n=4, lda=4, ldb=3,
             alpha=1.0, A=fl(1, -1, -2, -3, 2, 3, -4, -5, 4, 5, 6, -6, 7, 8, 9, 10),
             B=fl(11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22),
             expected=[e(-14, "/", 9), e(-14, "/", 10), e(-56, "/", 45), e(-7, "/", 18), e(-21, "/", 60),
                       e(-14, "/", 45), e(-1, "/", 6), e(-9, "/", 60), e(-8, "/", 60), e(20, "/", 10),
                       e(21, "/", 10), e(22, "/", 10)]),
        dict(name="3x3_left_low_trans_diag", side=0, uplo=1, trans=1, diag=1, m=3, n=3, lda=3, ldb=3,
             alpha=float(a01), A=fl(1, 2, 3, -2, 4, 5, -4, -5, 6), B=B1,
             expected=[e(a01, "*", 54), e(m01, "*", 37), e(a01, "*", 9), e(a01, "*", 72), e(m01, "*", 49),
                       e(a01, "*", 12), e(a01, "*", 90), e(m01, "*", 61), e(a01, "*", 15)]),
        dict(name="5x2_right_low_trans_diag", side=1, uplo=1, trans=1, diag=1, m=5, n=2, lda=3, ldb=6,
             alpha=float(a01), A=fl(1, 2, -1, -2, 3, -3), B=fl(4, 5, 6, 7, 8, -1, 10, 11, 12, 13, 14, -2),
             expected=[e(a01, "*", 4), e(a01, "*", 5), e(a01, "*", 6), e(a01, "*", 7), e(a01, "*", 8), -1.0,
                       e(a01, "*", 2), e(a01, "*", 1), 0.0, e(m01, "*", 1), e(m01, "*", 2), -2.0]),
    ]
    for c in cases:
        c["compare"] = "adaptive"
    return cases


def _pairs(lo, hi):
    """(lo, lo+1), (lo+2, lo+3), ... up to hi: the reference's running b arrays."""
    return [(v, v + 1) for v in range(lo, hi, 2)]


def ctrsm_cases():
    """tests/iclBLAS/test_Ctrsm.cpp:36-637. The reference computes each expected
    array with an inline std::complex<float> substitution and compares with
    EXPECT_COMPLEX_NEAR (relative 4e-4 per component, absolute for zero
    components, :21-34). Here the expected values are the exact solution of the
    same system, op(A) X = alpha B (left) or X op(A) = alpha B (right), solved in
    complex128 by numpy from the mathematical definition (triangle of A, unit
    diagonal, conjugation), independent of the oracle's substitution order; the
    padding rows of B (ldb > m) keep their input values."""
    A1 = [(1, 2), (-1, -2), (-3, -4), (3, 4), (5, 6), (-5, -6), (7, 8), (9, 10), (11, 12)]
    A16 = [(1, 2), (-1, -2), (-3, -4), (-5, -6), (3, 4), (5, 6), (-7, -8), (-9, -10), (7, 8), (9, 10), (11, 12),
           (-11, -12), (13, 14), (15, 16), (17, 18), (19, 20)]
    A4 = [(1, 2), (3, 4), (5, 6), (-11, -12), (-1, -2), (7, 8), (9, 10), (-11, -12), (-3, -4), (-5, -6), (11, 12),
          (-11, -12)]
    AL = [(1, 2), (3, 4), (5, 6), (-1, -2), (7, 8), (9, 10), (-3, -4), (-5, -6), (11, 12)]
    A20 = [(1, 2), (3, 4), (5, 6), (7, 8), (-1, -2), (-3, -4), (9, 10), (11, 12), (13, 14), (-5, -6), (-7, -8),
           (-9, -10), (15, 16), (17, 18), (-11, -12), (-13, -14), (-15, -16), (-17, -18), (19, 20), (-19, -20)]
    B5 = [(13, 14), (15, 16), (17, 18), (-1, -2), (19, 20), (21, 22), (23, 24), (-3, -4), (25, 26), (27, 28),
          (29, 30), (-5, -6), (31, 32), (33, 34), (35, 36), (-7, -8)]
    B6 = [(21, 22), (23, 24), (25, 26), (27, 28), (-1, -2), (29, 30), (31, 32), (33, 34), (35, 36), (-3, -4),
          (37, 38), (39, 40), (41, 42), (43, 44), (-5, -6)]
    B11 = [(21, 22), (23, 24), (25, 26), (-1, -2), (-3, -4), (27, 28), (29, 30), (31, 32), (-5, -6), (-7, -8),
           (33, 34), (35, 36), (37, 38), (-9, -10), (-11, -12), (39, 40), (41, 42), (43, 44), (-13, -14), (-15, -16)]
    raw = [
        ("3x3_left_up_ntrans_ndiag", 0, 0, 0, 0, 3, 3, 3, 3, (1.1, 2.3), A1, _pairs(13, 31)),
        ("4x3_left_up_trans_diag", 0, 0, 1, 1, 4, 3, 4, 4, (-1.1, 2.3), A16, _pairs(21, 45)),
        ("3x4_left_up_conj_ndiag", 0, 0, 2, 0, 3, 4, 3, 3, (-1.1, -2.3), A1, _pairs(13, 37)),
        ("3x3_left_low_ntrans_diag_lda4", 0, 1, 0, 1, 3, 3, 4, 3, (1.1, -2.3), A4, _pairs(13, 31)),
        ("3x4_left_low_trans_ndiag_ldb4", 0, 1, 1, 0, 3, 4, 3, 4, (2.1, 1.3), AL, B5),
        ("4x3_left_low_conj_diag_lds5", 0, 1, 2, 1, 4, 3, 5, 5, (-2.1, 1.3), A20, B6),
        ("3x3_right_up_ntrans_ndiag", 1, 0, 0, 0, 3, 3, 3, 3, (-2.1, -1.3), A1, _pairs(13, 31)),
        ("4x3_right_up_trans_ndiag_ldb5", 1, 0, 1, 0, 4, 3, 3, 4, (2.1, -1.3), A1, _pairs(13, 37)),
        ("3x4_right_up_conj_diag_lda5", 1, 0, 2, 1, 3, 4, 4, 3, (1.3, 2.3), A16, _pairs(21, 45)),
        ("3x3_right_low_ntrans_diag", 1, 1, 0, 1, 3, 3, 3, 3, (-1.3, -2.3), AL, _pairs(13, 31)),
        ("3x4_right_low_trans_ndiag_lds5", 1, 1, 1, 0, 3, 4, 5, 5, (1.3, -2.3), A20, B11),
        ("4x3_right_low_conj_ndiag", 1, 1, 2, 0, 4, 3, 3, 4, (1.3, 2.3), AL, _pairs(13, 37)),
    ]
    out = []
    for name, side, uplo, trans, diag, m, n, lda, ldb, alpha, A, B in raw:
        order = m if side == 0 else n
        assert len(A) == lda * order and len(B) == ldb * n, name
        al = complex(f(alpha[0]), f(alpha[1]))
        Am = np.array([complex(f(a), f(b)) for a, b in A]).reshape(order, lda).T  # column-major
        Bm = np.array([complex(f(a), f(b)) for a, b in B]).reshape(n, ldb).T
        T = Am[:order, :order].copy()
        T = np.tril(T) if uplo == 1 else np.triu(T)
        if diag == 1:
            np.fill_diagonal(T, 1.0)
        opT = T if trans == 0 else (T.T if trans == 1 else T.conj().T)
        rhs = al * Bm[:m, :n]
        X = np.linalg.solve(opT, rhs) if side == 0 else np.linalg.solve(opT.T, rhs.T).T
        E = Bm.copy()
        E[:m, :n] = X
        out.append(dict(name=name, side=side, uplo=uplo, trans=trans, diag=diag, m=m, n=n, lda=lda, ldb=ldb,
                        alpha=[float(al.real), float(al.imag)], A=cplx(A), B=cplx(B),
                        expected=[[float(z.real), float(z.imag)] for z in E.T.reshape(-1)],
                        compare="complex_near_4e-4"))
    return out


def main():
    data = {"source": "intel/clGPU tests/iclBLAS/test_{Sgemm,Cgemm,Strsm,Ctrsm}.cpp",
            "sgemm": sgemm_cases(), "cgemm": cgemm_cases(), "strsm": strsm_cases(), "ctrsm": ctrsm_cases()}
    out = Path(__file__).with_name("kats.json")
    out.write_text(json.dumps(data, indent=None, separators=(",", ":")))
    print(f"wrote {out} ({out.stat().st_size} bytes): "
          f"{len(data['sgemm'])} sgemm, {len(data['cgemm'])} cgemm, {len(data['strsm'])} strsm, "
          f"{len(data['ctrsm'])} ctrsm")


if __name__ == "__main__":
    main()
