"""Writes tests/golden/kats_l3.json: the known-answer tests of the reference's
API-level suites for the GEMM-based level-3 routines, transcribed as data.

Sources (reference tree, read as text, nothing executed or compiled):
  tests/iclBLAS/test_Ssymm.cpp   5 cases     tests/iclBLAS/test_Csymm.cpp   4 cases
  tests/iclBLAS/test_Ssyrk.cpp   4 cases     tests/iclBLAS/test_Chemm.cpp   4 cases
  tests/iclBLAS/test_Ssyr2k.cpp  4 cases     tests/iclBLAS/test_Csyrk.cpp   6 cases
  tests/iclBLAS/test_Strmm.cpp  11 cases     tests/iclBLAS/test_Csyr2k.cpp  6 cases
                                             tests/iclBLAS/test_Cherk.cpp   7 cases
                                             tests/iclBLAS/test_Cher2k.cpp  7 cases
                                             tests/iclBLAS/test_Ctrmm.cpp   8 cases

The parser reads each TEST block's scalar declarations, array initialisers and
the iclblas call (which names the arrays passed as A, B, C). Inputs are copied
value for value. A literal `expected` array is kept as `reference_expected`,
with float expressions (e.g. `1.3f*10.f + 1.1f*66.f`) evaluated in float32 in
C's order. Tests that compute their expectation with an inline loop have
`reference_expected = null`. Every case also gets `contract_expected`: the
header-contract result (include/iclBLAS/iclBLAS.h) in float64 from the oracle.

`deviation` marks the routines whose reference expectation is not the header
contract (the reference kernels Strmm_naive.cl, Ctrmm_naive.cl mask the output
instead of A; Cherk_naive.cl, Cher2k_naive.cl index only column i of A, and
the tests restate those kernels): the library follows the header for them.

test_Ssymm.cpp's 256x256_left_up builds its data with loops; it is stored as
`generator: ssymm_256` and rebuilt by tests/golden/l3_data.py.

Run: python tests/golden/make_kats_l3.py   (needs the reference tree at
/root/reference and the oracle built; only for regenerating the fixture)
"""
import ast
import json
import re
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
sys.path.insert(0, str(HERE))
import oracle  # noqa: E402
from l3_data import ssymm_256  # noqa: E402

REF = Path("/root/reference/tests/iclBLAS")
f32 = np.float32

ENUMS = {
    "ICLBLAS_SIDE_LEFT": 0, "ICLBLAS_SIDE_RIGHT": 1,
    "ICLBLAS_FILL_MODE_UPPER": 0, "ICLBLAS_FILL_MODE_LOWER": 1,
    "ICLBLAS_OP_N": 0, "ICLBLAS_OP_T": 1, "ICLBLAS_OP_C": 2,
    "ICLBLAS_DIAG_NON_UNIT": 0, "ICLBLAS_DIAG_UNIT": 1,
}

# Parameter names of each routine after the handle (reference signatures,
# include/iclBLAS/iclBLAS.h).
SIGS = {
    "Ssymm": "side uplo m n alpha A lda B ldb beta C ldc",
    "Csymm": "side uplo m n alpha A lda B ldb beta C ldc",
    "Chemm": "side uplo m n alpha A lda B ldb beta C ldc",
    "Ssyrk": "uplo trans n k alpha A lda beta C ldc",
    "Csyrk": "uplo trans n k alpha A lda beta C ldc",
    "Cherk": "uplo trans n k alpha A lda beta C ldc",
    "Ssyr2k": "uplo trans n k alpha A lda B ldb beta C ldc",
    "Csyr2k": "uplo trans n k alpha A lda B ldb beta C ldc",
    "Cher2k": "uplo trans n k alpha A lda B ldb beta C ldc",
    "Strmm": "side uplo trans diag m n alpha A lda B ldb C ldc",
    "Ctrmm": "side uplo trans diag m n alpha A lda B ldb C ldc",
}
DEVIATION = {"Strmm", "Ctrmm", "Cherk", "Cher2k"}
COMPLEX = {"Csymm", "Chemm", "Csyrk", "Cherk", "Csyr2k", "Cher2k", "Ctrmm"}


def strip_comments(s):
    s = re.sub(r"/\*.*?\*/", "", s, flags=re.S)
    return re.sub(r"//[^\n]*", "", s)


def f32_eval(expr, env):
    """Evaluates a C float expression in float32, left to right as C does."""
    e = re.sub(r"(\d+\.\d*|\.\d+|\d+)([eE][+-]?\d+)?[fF]\b", r"\1\2", expr.strip())

    def ev(node):
        if isinstance(node, ast.Expression):
            return ev(node.body)
        if isinstance(node, ast.Constant):
            return f32(node.value) if isinstance(node.value, float) else node.value
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub):
            return -ev(node.operand)
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.UAdd):
            return ev(node.operand)
        if isinstance(node, ast.Name):
            return env[node.id]
        if isinstance(node, ast.BinOp):
            a, b = ev(node.left), ev(node.right)
            if isinstance(a, int) and isinstance(b, int):
                if isinstance(node.op, ast.Div):
                    return int(a / b)
                return {ast.Add: a + b, ast.Sub: a - b, ast.Mult: a * b}[type(node.op)]
            a, b = f32(a), f32(b)
            op = type(node.op)
            return {ast.Add: a + b, ast.Sub: a - b, ast.Mult: a * b, ast.Div: a / b}[op]
        raise ValueError(f"unsupported expression {expr!r}")

    return ev(ast.parse(e, mode="eval"))


def split_top(s, sep=","):
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[<":
            depth += 1
        elif ch in ")}]>":
            depth -= 1
        if ch == sep and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur)
    return [x.strip() for x in out]


PAIR = re.compile(r"\{\s*([^{},]+?)\s*,\s*([^{},]+?)\s*\}")


def parse_block(routine, name, body):
    env, arrays = {}, {}
    body = strip_comments(body)
    call = re.search(r"iclblas" + routine + r"\(\s*handle\s*,(.*?)\)\s*;", body, re.S)
    if not call:
        return None
    for stmt in body[: call.start()].split(";"):
        st = " ".join(stmt.split()).lstrip("{} ")
        m = re.match(r"^(?:const )?(?:float|std::complex<float>|oclComplex_t) (\w+)\s*\[[^\]]*\]\s*=\s*\{(.*)\}$", st)
        if m:
            nm, init = m.group(1), m.group(2)
            pairs = PAIR.findall(init)
            if pairs:
                arrays[nm] = [[float(f32_eval(a, env)), float(f32_eval(b, env))] for a, b in pairs]
            else:
                arrays[nm] = [float(f32_eval(x, env)) for x in split_top(init) if x]
            continue
        m = re.match(r"^(?:const )?(?:std::complex<float>|oclComplex_t) (\w+)\s*(?:\((.*)\)|=\s*\{(.*)\})$", st)
        if m:
            nm, init = m.group(1), m.group(2) or m.group(3)
            re_, im_ = split_top(init)
            env[nm] = [float(f32_eval(re_, env)), float(f32_eval(im_, env))]
            continue
        m = re.match(r"^(?:const )?(?:auto|int|float|iclblas\w+) (\w+)\s*=\s*(.+)$", st)
        if m and not st.startswith("iclblasStatus_t") and not st.startswith("iclblasHandle_t"):
            nm, init = m.group(1), m.group(2).strip()
            if init in ENUMS:
                env[nm] = ENUMS[init]
            else:
                try:
                    v = f32_eval(init, env)
                except (ValueError, SyntaxError, KeyError):
                    continue  # computed by a helper call (the loop-built case)
                env[nm] = int(v) if isinstance(v, int) else float(v)
    args = []
    for a in split_top(call.group(1)):
        a = re.sub(r"reinterpret_cast<[^>]*>\((.*)\)$", r"\1", a.strip())
        a = a.lstrip("&").replace(".data()", "")
        args.append(a)
    case = {"name": name, "routine": routine}
    for pname, a in zip(SIGS[routine].split(), args):
        if pname in ("A", "B", "C"):
            case[pname] = a  # array name, resolved below
        elif a in env:
            case[pname] = env[a]
        elif a in ENUMS:
            case[pname] = ENUMS[a]
        else:
            case[pname] = f32_eval(a, env)
            case[pname] = float(case[pname]) if not isinstance(case[pname], int) else case[pname]
    for pname in ("A", "B", "C"):
        if pname in case:
            nm = case[pname]
            if nm not in arrays:
                return {"name": name, "routine": routine, "unparsed": nm}
            case[pname] = arrays[nm]
    exp = arrays.get("expected")
    case["reference_expected"] = exp
    return case


def to_mat(vals, ld, cols, cplx):
    v = np.asarray(vals, np.float64)
    if cplx:
        v = (v[:, 0] + 1j * v[:, 1]).astype(np.complex64)
    else:
        v = v.astype(np.float32)
    rows = len(v) // cols
    assert rows == ld and rows * cols == len(v), (len(v), ld, cols)
    return np.asfortranarray(v.reshape(cols, rows).T)


def scalar(v):
    return complex(v[0], v[1]) if isinstance(v, list) else float(v)


def contract(c):
    r, cplx = c["routine"], c["routine"] in COMPLEX
    al = scalar(c["alpha"])
    if r in ("Ssymm", "Csymm", "Chemm"):
        order = c["m"] if c["side"] == 0 else c["n"]
        A = to_mat(c["A"], c["lda"], order, cplx)
        B = to_mat(c["B"], c["ldb"], c["n"], cplx)
        C = to_mat(c["C"], c["ldc"], c["n"], cplx)
        X, _ = oracle.zsymm(c["side"], c["uplo"], r == "Chemm", c["m"], c["n"], al, A, B, scalar(c["beta"]), C)
        out = np.array(C, np.complex128 if cplx else np.float64)
        out[: c["m"], :] = X
    elif r in ("Ssyrk", "Csyrk", "Cherk", "Ssyr2k", "Csyr2k", "Cher2k"):
        rows, cols = (c["n"], c["k"]) if c["trans"] == 0 else (c["k"], c["n"])
        A = to_mat(c["A"], c["lda"], cols, cplx)
        C = to_mat(c["C"], c["ldc"], c["n"], cplx)
        herm = r in ("Cherk", "Cher2k")
        if "2k" in r:
            B = to_mat(c["B"], c["ldb"], cols, cplx)
            X, _ = oracle.zsyr2k(c["uplo"], c["trans"], herm, c["n"], c["k"], al, A, B, scalar(c["beta"]), C)
        else:
            X, _ = oracle.zsyrk(c["uplo"], c["trans"], herm, c["n"], c["k"], al, A, scalar(c["beta"]), C)
        out = np.array(C, np.complex128 if cplx else np.float64)
        out[: c["n"], :] = X
    else:  # trmm
        order = c["m"] if c["side"] == 0 else c["n"]
        A = to_mat(c["A"], c["lda"], order, cplx)
        B = to_mat(c["B"], c["ldb"], c["n"], cplx)
        X, _ = oracle.ztrmm(c["side"], c["uplo"], c["trans"], c["diag"], c["m"], c["n"], al, A, B)
        C = to_mat(c["C"], c["ldc"], c["n"], cplx) if isinstance(c["C"], list) else None
        out = np.array(C, np.complex128 if cplx else np.float64)
        out[: c["m"], :] = X
    flat = out.T.reshape(-1)
    if cplx:
        return [[float(z.real), float(z.imag)] for z in flat]
    return [float(x) for x in flat]


def cases_of(routine):
    text = (REF / f"test_{routine}.cpp").read_text()
    out = []
    for m in re.finditer(r"TEST\(\s*" + routine + r"\s*,\s*(\w+)\s*\)", text):
        start = m.end()
        nxt = text.find("TEST(", start)
        body = text[start: nxt if nxt >= 0 else len(text)]
        c = parse_block(routine, m.group(1), body)
        if c is None:
            continue
        out.append(c)
    return out


def main():
    all_cases = []
    for routine in SIGS:
        for c in cases_of(routine):
            if c.get("unparsed") == "A" and c["name"] == "256x256_left_up":
                c = dict(ssymm_256(), name=c["name"], routine=routine, generator="ssymm_256")
                c["reference_expected"] = None
                c["contract_expected"] = None  # computed by the test from the rebuilt data
            else:
                assert "unparsed" not in c, c
                # Strmm / Ctrmm tests pass B as C (or a separate C); C's initial
                # contents are what the reference leaves outside its writes.
                c["contract_expected"] = contract(c)
            c["deviation"] = routine in DEVIATION
            c["complex"] = routine in COMPLEX
            all_cases.append(c)
    counts = {}
    for c in all_cases:
        counts[c["routine"]] = counts.get(c["routine"], 0) + 1
    print(counts, len(all_cases))
    (HERE / "kats_l3.json").write_text(json.dumps({"cases": all_cases}, indent=None, separators=(",", ":")) + "\n")


if __name__ == "__main__":
    main()
