"""Writes tests/golden/kats_gemv.json: the reference's GEMV known-answer tests.

Sources (reference tree, read as text, nothing executed or compiled):
  tests/iclBLAS/test_Sgemv.cpp:20-495  8 cases. Data are built by loops,
      A[i] = x[i] = y[i] = 1.f * i / n / n over each vector's full length
      (A: lda * n; x: n * incx for op N, m * incx otherwise; y: m * incy for op N,
      n * incy otherwise), rebuilt by l3_data.sgemv_arrays; stored as parameters.
      The reference's expectation is its inline FP32 loop (l_result +=
      alpha * A * x, then + beta * y), restated here as `reference_expected`;
      it compares with EXPECT_NEAR 1e-5 (N_256x256) or EXPECT_FLOAT_EQ.
  tests/iclBLAS/test_Cgemv.cpp:19-290   7 cases with literal inputs and
      expectations (EXPECT_ARRAYS_EQ, ComplexEqual: 4 ULP per part).
Every case also gets `contract_expected` (float64 / complex128 from the oracle,
include/iclBLAS/iclBLAS.h:1020-1049, 1135-1164).

Run: python tests/golden/make_kats_gemv.py   (needs the reference tree at
/root/reference and the oracle built; only for regenerating the fixture)
"""
import json
import re
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
sys.path.insert(0, str(HERE))
import oracle  # noqa: E402
from l3_data import sgemv_arrays  # noqa: E402
from make_kats_l3 import ENUMS, PAIR, f32_eval, split_top, strip_comments  # noqa: E402

REF = Path("/root/reference/tests/iclBLAS")
f32 = np.float32

# test_Sgemv.cpp: (name, trans, m, n, lda, incx, incy, alpha, beta, compare)
SGEMV = [
    ("N_256x256", 0, 256, 256, 256, 1, 1, 1.0, 1.0, "near_1e-5"),   # :20-77
    ("N_15x25_inc", 0, 25, 20, 25, 2, 4, 1.0, 1.0, "ulp4"),         # :79-136
    ("N_35x25_beta0", 0, 25, 35, 25, 1, 1, 1.0, 0.0, "ulp4"),       # :138-195
    ("N_25x25_lda", 0, 25, 25, 35, 1, 1, 1.0, 1.0, "ulp4"),         # :197-254
    ("T_15x25", 1, 25, 15, 25, 1, 1, 1.0, 1.0, "ulp4"),             # :256-314
    ("T_15x25_inc", 1, 25, 15, 25, 2, 4, 1.0, 1.0, "ulp4"),         # :316-374
    ("T_15x25_beta0", 1, 25, 15, 25, 1, 1, 1.0, 0.0, "ulp4"),       # :376-434
    ("T_15x25_lda", 1, 25, 15, 35, 1, 1, 1.0, 1.0, "ulp4"),         # :436-495
]


def sgemv_reference_loop(c, A, x, y):
    """The reference test's own expectation: an FP32 loop per element of y,
    l_result += alpha * A(.) * x(.), then l_result + beta * y (test_Sgemv.cpp:49-60)."""
    m, n, t = c["m"], c["n"], c["trans"]
    alpha, beta = f32(c["alpha"]), f32(c["beta"])
    out = y.copy()
    leny, lenr = (m, n) if t == 0 else (n, m)
    for r in range(leny):
        acc = f32(0)
        for k in range(lenr):
            a = A[r, k] if t == 0 else A[k, r]
            acc = f32(acc + f32(f32(alpha * a) * x[k * c["incx"]]))
        out[r * c["incy"]] = f32(acc + f32(beta * y[r * c["incy"]]))
    return out


def cgemv_cases():
    text = strip_comments((REF / "test_Cgemv.cpp").read_text())
    out = []
    for m in re.finditer(r"TEST\(\s*Cgemv\s*,\s*(\w+)\s*\)", text):
        start = m.end()
        nxt = text.find("TEST(", start)
        body = text[start: nxt if nxt >= 0 else len(text)]
        call = re.search(r"iclblasCgemv\(\s*handle\s*,(.*?)\)\s*;", body, re.S)
        env, arrays = {}, {}
        for stmt in body[: call.start()].split(";"):
            st = " ".join(stmt.split()).lstrip("{} ")
            a = re.match(r"^oclComplex_t (\w+)\s*\[[^\]]*\]\s*=\s*\{(.*)\}$", st)
            if a:
                arrays[a.group(1)] = [[float(f32_eval(p, env)), float(f32_eval(q, env))]
                                      for p, q in PAIR.findall(a.group(2))]
                continue
            a = re.match(r"^oclComplex_t (\w+)\s*=\s*\{(.*)\}$", st)
            if a:
                re_, im_ = split_top(a.group(2))
                env[a.group(1)] = [float(f32_eval(re_, env)), float(f32_eval(im_, env))]
                continue
            a = re.match(r"^(?:const )?(?:int|iclblasOperation_t) (\w+)\s*=\s*(.+)$", st)
            if a:
                v = a.group(2).strip()
                env[a.group(1)] = ENUMS[v] if v in ENUMS else int(f32_eval(v, env))
        args = [x.strip().lstrip("&") for x in split_top(call.group(1))]
        names = "trans m n alpha A lda x incx beta y incy".split()
        c = {"name": m.group(1), "routine": "Cgemv"}
        for pn, a in zip(names, args):
            c[pn] = arrays[a] if pn in ("A", "x", "y") else env[a]
        c["reference_expected"] = arrays["ex_result"]
        c["compare"] = "ulp4"
        out.append(c)
    return out


def contract(c, A, x, y):
    al = complex(*c["alpha"]) if isinstance(c["alpha"], list) else c["alpha"]
    be = complex(*c["beta"]) if isinstance(c["beta"], list) else c["beta"]
    res, _ = oracle.zgemv(c["trans"], c["m"], c["n"], al, A, x, c["incx"], be, y, c["incy"])
    out = np.array(y, np.complex128 if np.iscomplexobj(A) else np.float64)
    out[:: c["incy"]][: len(res)] = res
    return out


def main():
    cases = []
    for name, t, m, n, lda, incx, incy, alpha, beta, cmp in SGEMV:
        c = dict(name=name, routine="Sgemv", trans=t, m=m, n=n, lda=lda, incx=incx, incy=incy, alpha=alpha,
                 beta=beta, compare=cmp, generator="ratio")
        A, x, y = sgemv_arrays(c)
        c["reference_expected"] = [float(v) for v in sgemv_reference_loop(c, A, x, y)]
        c["contract_expected"] = [float(v) for v in contract(c, A, x, y)]
        cases.append(c)
    for c in cgemv_cases():
        lda = c["lda"]
        v = np.asarray(c["A"], np.float64)
        A = np.asfortranarray((v[:, 0] + 1j * v[:, 1]).astype(np.complex64).reshape(-1, lda).T)
        x = np.asarray([complex(*p) for p in c["x"]], np.complex64)
        y = np.asarray([complex(*p) for p in c["y"]], np.complex64)
        out = contract(c, A, x, y)
        c["contract_expected"] = [[float(z.real), float(z.imag)] for z in out]
        cases.append(c)
    print(len(cases), [c["name"] for c in cases])
    (HERE / "kats_gemv.json").write_text(json.dumps({"cases": cases}, separators=(",", ":")) + "\n")


if __name__ == "__main__":
    main()
