"""GEMM-based level-3 routines on the GPU through the C ABI (Ssymm, Csymm,
Chemm, Ssyrk, Csyrk, Cherk, Ssyr2k, Csyr2k, Cher2k, Strmm, Ctrmm):
  - the reference's known answers (tests/golden/kats_l3.json): the literal
    float expectations with the reference's 4-ULP EXPECT_FLOAT_EQ, the loop-built
    SSYMM 256 case with its 1e-4, the rest against the FP64 header contract;
  - every side / uplo / trans / diag variant against the FP64 oracle at shapes
    that are and are not multiples of the GEMM tiles, within the elementwise
    bound |err| <= 4 (k + 2) 2^-24 (|alpha| |X||Y| + |beta| |C|);
  - never-read parts poisoned with NaN (the other triangle of A, a unit
    diagonal, hermitian diagonal imaginary parts, C when beta == 0), the other
    triangle of C and ld padding compared bit for bit with their input;
  - in-place TRMM (C is B) in both memory modes, and large problems in
    device memory against float64 numpy products;
  - the split-BF16 math mode at depths of at least 256, every variant, within
    the same bound."""
import numpy as np
import pytest

from blas_helpers import complex_eq, float_eq
from clgpu_amd import iclblas as ib
from l3_helpers import (COMPLEX, KATS, RANK2K, RANKK, SYMM, TRMM, call_library, expected_matrix, kat_case,
                        oracle_result, random_case, shapes)

pytestmark = pytest.mark.gpu
U = 2.0 ** -24


@pytest.fixture
def h(handle):
    handle.set_math_mode(ib.MATH_FP32)
    handle.set_implementation(None)
    yield handle


def _run(h, c):
    A, B, C = c["A"].copy(order="F"), (c["B"].copy(order="F") if c.get("B") is not None else None), \
        c["C"].copy(order="F")
    ib.check(call_library(h.value, c, A, B, C), c["routine"])
    return C


def _bound_ok(c, got, exp, mag, k):
    (rc, cc) = shapes(c)[2]
    g = np.asarray(got[:rc, :cc], np.complex128)
    e = exp[:rc, :cc]
    C0 = np.asarray(c["C"][:rc, :cc], np.complex128)
    beta = c.get("beta", 0.0)
    bc = np.abs(beta) * np.where(np.isfinite(C0), np.abs(C0), 0.0) if beta != 0 else 0.0
    bound = 4.0 * (k + 2) * U * (abs(c["alpha"]) * mag + bc)
    fin = np.isfinite(e)
    err = np.abs(g - e)
    ok = (err <= bound + 1e-30) | ~fin
    return bool(np.all(ok)), float(np.max(np.where(fin, err / np.maximum(bound, 1e-30), 0)))


# ---------------------------------------------------------------- known answers
@pytest.mark.parametrize("case", KATS, ids=lambda c: f"{c['routine']}-{c['name']}")
def test_level3_kats(h, case):
    c = kat_case(case)
    got = _run(h, c)
    cplx = c["routine"] in COMPLEX
    if case.get("generator") == "ssymm_256":
        exp, _, _ = oracle_result(c, c["A"], c["B"], c["C"])
        assert np.abs(got - exp).max() <= 1e-4
        return
    if case["reference_expected"] is not None and not case["deviation"]:
        exp = expected_matrix(case, "reference_expected")
        ok = all((complex_eq if cplx else float_eq)(e, g) for e, g in zip(exp.T.reshape(-1), got.T.reshape(-1)))
        assert ok, (exp, got)
        return
    # inline-loop expectations and the routines that follow the header contract
    exp, mag, k = oracle_result(c, c["A"], c.get("B"), c["C"])
    ok, ratio = _bound_ok(c, got, exp, mag, k)
    assert ok, ratio
    (rc, cc) = shapes(c)[2]
    # everything outside the computed region is untouched
    rest = np.ones(got.shape, bool)
    rest[:rc, :cc] = False
    assert np.array_equal(got[rest], c["C"][rest])


# ---------------------------------------------------------------- variants
def _variants(routine):
    if routine in SYMM:
        return [dict(side=s, uplo=u) for s in (0, 1) for u in (0, 1)]
    if routine in RANKK or routine in RANK2K:
        return [dict(uplo=u, trans=t) for u in (0, 1) for t in (0, 1, 2)]
    return [dict(side=s, uplo=u, trans=t, diag=d) for s in (0, 1) for u in (0, 1) for t in (0, 1, 2) for d in (0, 1)]


ALL = [(r, v) for r in SYMM + RANKK + RANK2K + TRMM for v in _variants(r)]
SHAPES = [(5, 3, 4), (64, 64, 64), (130, 70, 33), (257, 129, 300)]


def _vid(x):
    if isinstance(x, str):
        return x
    if isinstance(x, tuple):
        return "x".join(map(str, x))
    return "".join(f"{k[0]}{v}" for k, v in x.items())


@pytest.mark.parametrize("shape", SHAPES, ids=_vid)
@pytest.mark.parametrize("routine,v", ALL, ids=_vid)
def test_level3_variants_vs_oracle(h, routine, v, shape):
    m, n, k = shape
    rng = np.random.default_rng(m * 1000 + n)
    c = random_case(routine, rng, m=m, n=n, k=k, **v)
    got = _run(h, c)
    exp, mag, kk = oracle_result(c, c["A"], c["B"], c["C"])
    (rc, cc) = shapes(c)[2]
    ok, ratio = _bound_ok(c, got, exp, mag, kk)
    assert ok, ratio
    assert np.all(np.isfinite(got[:rc, :cc][np.isfinite(exp[:rc, :cc])]))
    if routine in RANKK or routine in RANK2K:
        other = np.triu(np.ones((rc, cc), bool), 1) if v["uplo"] == 1 else np.tril(np.ones((rc, cc), bool), -1)
        assert np.array_equal(got[:rc, :cc][other], c["C"][:rc, :cc][other])  # other triangle untouched
    assert np.array_equal(got[rc:], c["C"][rc:])  # ld padding untouched
    assert h.last_implementation().startswith(("sgemm_f32", "cgemm_f32"))


DMA_L3 = [("Ssyrk", dict(uplo=u, trans=t), (0, 512, 272)) for u in (0, 1) for t in (0, 1)] + [
    ("Ssyr2k", dict(uplo=1, trans=1), (0, 512, 256)),
    ("Strmm", dict(side=0, uplo=0, trans=1, diag=0), (512, 256, 0)),
    ("Strmm", dict(side=1, uplo=1, trans=0, diag=1), (256, 512, 0)),
    ("Cherk", dict(uplo=0, trans=2), (0, 256, 136)),
    ("Cherk", dict(uplo=1, trans=0), (0, 256, 136)),
    ("Csyrk", dict(uplo=1, trans=1), (0, 256, 128)),
    ("Cher2k", dict(uplo=0, trans=0), (0, 256, 128)),
    ("Ctrmm", dict(side=1, uplo=1, trans=2, diag=1), (256, 384, 0)),
    ("Ctrmm", dict(side=0, uplo=0, trans=0, diag=0), (384, 256, 0)),
    ("Ssyrk", dict(uplo=0, trans=1), (0, 384, 144), "sgemm_f32_dma_128x128"),
    ("Ssyr2k", dict(uplo=1, trans=0), (0, 384, 128), "sgemm_f32_dma_128x128"),
    ("Strmm", dict(side=1, uplo=0, trans=1, diag=0), (128, 384, 0), "sgemm_f32_dma_128x128")]


@pytest.mark.parametrize("routine,v,shape,impl", [
    c if len(c) == 4 else c + ("cgemm_f32_dma_128x128" if c[0] in COMPLEX else "sgemm_f32_dma_256x256w8",)
    for c in DMA_L3], ids=_vid)
def test_level3_dma_kernels_forced(h, routine, v, shape, impl):
    """The LDS-DMA FP32 GEMM kernels forced on level-3 problems without edges:
    the written triangle's tiles only (SYRK / HERK types), the k windows of a
    triangular operand (TRMM), the conjugated accumulator and real diagonal of
    the complex epilogue."""
    m, n, k = shape
    rng = np.random.default_rng(m + 7 * n + 13 * k)
    c = random_case(routine, rng, m=m or None, n=n, k=k or None, pad=0, **v)
    h.set_implementation(impl)
    got = _run(h, c)
    exp, mag, kk = oracle_result(c, c["A"], c["B"], c["C"])
    (rc, cc) = shapes(c)[2]
    ok, ratio = _bound_ok(c, got, exp, mag, kk)
    assert ok, ratio
    if routine in RANKK or routine in RANK2K:
        other = np.triu(np.ones((rc, cc), bool), 1) if v["uplo"] == 1 else np.tril(np.ones((rc, cc), bool), -1)
        assert np.array_equal(got[:rc, :cc][other], c["C"][:rc, :cc][other])
    assert h.last_implementation() == impl


@pytest.mark.parametrize("routine,v", ALL, ids=_vid)
def test_level3_split_bf16_vs_oracle(h, routine, v):
    """Split-BF16 math mode with every product at least 256 deep: the split
    kernels' level-3 instantiation (triangle tiles and stores, k windows,
    conjugated accumulator, real diagonal) within the same elementwise bound,
    the other triangle and the padding untouched."""
    h.set_math_mode(ib.MATH_SPLIT_BF16X3)
    m, n, k = 300, 270, 280
    rng = np.random.default_rng(77)
    c = random_case(routine, rng, m=m, n=n, k=k, **v)
    got = _run(h, c)
    exp, mag, kk = oracle_result(c, c["A"], c["B"], c["C"])
    (rc, cc) = shapes(c)[2]
    ok, ratio = _bound_ok(c, got, exp, mag, kk)
    assert ok, ratio
    if routine in RANKK or routine in RANK2K:
        other = np.triu(np.ones((rc, cc), bool), 1) if v["uplo"] == 1 else np.tril(np.ones((rc, cc), bool), -1)
        assert np.array_equal(got[:rc, :cc][other], c["C"][:rc, :cc][other])
        if routine in ("Cherk", "Cher2k"):
            assert np.all(np.diag(got[:rc, :cc]).imag == 0)
    assert np.array_equal(got[rc:], c["C"][rc:])
    assert h.last_implementation().startswith(("sgemm_bf16x3", "cgemm_bf16x3"))


@pytest.mark.parametrize("routine,v", [(r, v) for r, v in ALL if r not in TRMM][::2], ids=_vid)
def test_level3_beta_zero_does_not_read_c(h, routine, v):
    rng = np.random.default_rng(5)
    c = random_case(routine, rng, m=70, n=45, k=20, beta=0.0, **v)
    got = _run(h, c)
    exp, mag, kk = oracle_result(c, c["A"], c["B"], c["C"])
    ok, ratio = _bound_ok(c, got, exp, mag, kk)
    assert ok, ratio


@pytest.mark.parametrize("routine", ["Ssyrk", "Cherk", "Csyr2k", "Ssymm", "Chemm"])
def test_level3_alpha_zero_scales_c_only(h, routine):
    rng = np.random.default_rng(9)
    v = dict(uplo=1, trans=0) if routine not in SYMM else dict(side=0, uplo=1)
    c = random_case(routine, rng, m=40, n=33, k=12, alpha=0.0, beta=(0.5 if routine == "Cherk" else 0.5 + 0.25j)
                    if routine in COMPLEX else 0.5, **v)
    c["A"][:] = np.nan
    if c.get("B") is not None:
        c["B"][:] = np.nan
    got = _run(h, c)
    exp, mag, kk = oracle_result(dict(c, alpha=0.0), np.zeros_like(c["A"]), np.zeros_like(c["B"]), c["C"])
    (rc, cc) = shapes(c)[2]
    assert np.allclose(got[:rc, :cc], exp[:rc, :cc], rtol=1e-6, atol=1e-7, equal_nan=True)


@pytest.mark.parametrize("routine", ["Strmm", "Ctrmm"])
def test_trmm_alpha_zero_writes_zeros_without_reading(h, routine):
    rng = np.random.default_rng(2)
    c = random_case(routine, rng, m=30, n=20, side=1, uplo=0, trans=2, diag=0, alpha=0.0)
    c["A"][:] = np.nan
    c["B"][:] = np.nan
    got = _run(h, c)
    assert np.all(got[:30] == 0) and np.array_equal(got[30:], c["C"][30:])


@pytest.mark.parametrize("routine", ["Strmm", "Ctrmm"])
def test_trmm_in_place_host(h, routine):
    """C is the same host matrix as B, as in the reference's tests."""
    rng = np.random.default_rng(4)
    c = random_case(routine, rng, m=200, n=150, side=0, uplo=1, trans=0, diag=0, pad=0, poison=False)
    B = c["B"].copy(order="F")
    ib.check(call_library(h.value, dict(c, ldc=c["ldb"]), c["A"], B, B), routine)
    exp, mag, kk = oracle_result(c, c["A"], c["B"], c["C"])
    ok, ratio = _bound_ok(dict(c, C=c["B"]), B, exp, mag, kk)
    assert ok, ratio


# ---------------------------------------------------------------- device memory, large
def _torch():
    import torch
    return torch


@pytest.mark.parametrize("routine,v", [("Strmm", dict(side=0, uplo=1, trans=0, diag=0)),
                                       ("Strmm", dict(side=1, uplo=0, trans=1, diag=1)),
                                       ("Ctrmm", dict(side=0, uplo=0, trans=2, diag=0))], ids=_vid)
def test_trmm_in_place_device(dev_handle, routine, v):
    torch = _torch()
    rng = np.random.default_rng(8)
    m, n = 520, 300
    c = random_case(routine, rng, m=m, n=n, pad=0, poison=False, **v)
    dt = torch.complex64 if routine in COMPLEX else torch.float32
    dA = torch.from_numpy(np.ascontiguousarray(c["A"].T)).to("cuda")
    dB = torch.from_numpy(np.ascontiguousarray(c["B"].T)).to("cuda")
    assert dA.dtype == dt
    ib.check(call_library(dev_handle.value, dict(c, ldc=c["ldb"]), dA.data_ptr(), dB.data_ptr(), dB.data_ptr()),
             routine)
    torch.cuda.synchronize()
    got = dB.cpu().numpy().T
    exp, mag, kk = oracle_result(c, c["A"], c["B"], c["C"])
    ok, ratio = _bound_ok(dict(c, C=c["B"]), got, exp, mag, kk)
    assert ok, ratio


LARGE = [("Ssyrk", dict(uplo=1, trans=0), (0, 2048, 1024)), ("Cherk", dict(uplo=0, trans=2), (0, 1536, 512)),
         ("Ssyr2k", dict(uplo=0, trans=1), (0, 1024, 768)), ("Ssymm", dict(side=1, uplo=1), (1536, 2048, 0)),
         ("Chemm", dict(side=0, uplo=0), (1024, 768, 0)), ("Strmm", dict(side=0, uplo=1, trans=0, diag=0), (2048, 1024, 0)),
         ("Strmm", dict(side=1, uplo=0, trans=1, diag=1), (1024, 2048, 0)),
         ("Ctrmm", dict(side=1, uplo=1, trans=2, diag=0), (768, 1024, 0))]


@pytest.mark.parametrize("routine,v,shape", LARGE, ids=lambda x: _vid(x))
def test_level3_large_device(dev_handle, routine, v, shape):
    """Device memory mode at sizes of thousands, against float64 numpy products
    (the closed forms of the header contract)."""
    from l3_helpers import closed_form
    torch = _torch()
    m, n, k = shape
    rng = np.random.default_rng(12)
    c = random_case(routine, rng, m=m or None, n=n, k=k or None, pad=0, poison=False, **v)
    dev = {key: torch.from_numpy(np.ascontiguousarray(c[key].T)).to("cuda") for key in ("A", "B", "C")}
    ib.check(call_library(dev_handle.value, c, dev["A"].data_ptr(), dev["B"].data_ptr(), dev["C"].data_ptr()), routine)
    torch.cuda.synchronize()
    got = dev["C"].cpu().numpy().T
    exp = closed_form(c, c["A"], c["B"], c["C"])
    (rc, cc) = shapes(c)[2]
    kk = {"Ssyr2k": 2 * (k or 0)}.get(routine, k or (m if v.get("side", 0) == 0 else n))
    scale = np.abs(exp[:rc, :cc]).max()
    err = np.abs(got[:rc, :cc] - exp[:rc, :cc]).max()
    assert err <= 4 * kk * U * max(scale, 1.0) * 4, (err, scale)


@pytest.mark.parametrize("routine", ["Ssyrk", "Ssyr2k"])
@pytest.mark.parametrize("uplo,trans", [(ib.FILL_MODE_LOWER, ib.OP_N), (ib.FILL_MODE_UPPER, ib.OP_T)],
                         ids=["LN", "UT"])
def test_syrk_band_split_8192(dev_handle, routine, uplo, trans):
    """SSYRK / SSYR2K at n = 8192, k = 2048 on the FP32 path, where the selection
    runs each product as two launches (the tiles off the diagonal 256 x 256
    blocks, then the diagonal band): two timed launches per product, the written
    triangle against a float64 product on the GPU within the elementwise bound,
    the other triangle bit for bit untouched."""
    torch = _torch()
    n, k, alpha, beta = 8192, 2048, 0.75, -0.5
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    rows, cols = (n, k) if trans == ib.OP_N else (k, n)
    A = torch.rand((cols, rows), generator=g, device="cuda") * 2 - 1  # column-major rows x cols, ld = rows
    B = torch.rand((cols, rows), generator=g, device="cuda") * 2 - 1
    C0 = torch.rand((n, n), generator=g, device="cuda") * 2 - 1
    C = C0.clone()
    h = dev_handle
    h.set_math_mode(ib.MATH_FP32)
    h.set_implementation(None)
    h.set_kernel_timing(True)
    if routine == "Ssyrk":
        st = ib.iclblasSsyrk(h.value, uplo, trans, n, k, alpha, A.data_ptr(), rows, beta, C.data_ptr(), n)
    else:
        st = ib.iclblasSsyr2k(h.value, uplo, trans, n, k, alpha, A.data_ptr(), rows, B.data_ptr(), rows, beta,
                              C.data_ptr(), n)
    ib.check(st, routine)
    torch.cuda.synchronize()
    _, launches = h.kernel_timing()
    h.set_kernel_timing(False)
    assert launches == (2 if routine == "Ssyrk" else 4)
    opA = A.double().t() if trans == ib.OP_N else A.double()  # n x k
    opB = B.double().t() if trans == ib.OP_N else B.double()
    if routine == "Ssyrk":
        P, mag, kk = opA @ opA.t(), opA.abs() @ opA.abs().t(), k
    else:
        P = opA @ opB.t() + opB @ opA.t()
        mag, kk = opA.abs() @ opB.abs().t() + opB.abs() @ opA.abs().t(), 2 * k
    ref = alpha * P + beta * C0.double().t()
    del P
    keep = torch.ones((n, n), dtype=torch.bool, device="cuda")
    keep = keep.tril() if uplo == ib.FILL_MODE_LOWER else keep.triu()
    got = C.t()
    bound = 4.0 * (kk + 2) * U * (abs(alpha) * mag + abs(beta) * C0.double().t().abs())
    err = (got.double() - ref).abs()
    assert bool((err[keep] <= bound[keep]).all()), float((err[keep] / bound[keep]).max())
    assert torch.equal(got[~keep], C0.t()[~keep])
