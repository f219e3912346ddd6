"""CGEMM parity on the GPU through the C ABI: the reference's known answers
(tests/iclBLAS/test_Cgemm.cpp, 4 ULP per component), the implementation-level
grid the reference left disabled (tests/blas_funcs/test_gemm.cpp:102-122:
N/T/C x N/T/C, padded leading dimensions, beta zero / non-zero) on
quarter-integer data with exact equality against cpu_gemm<complex<float>>,
and the FP64 bound on random data, for the FP32 and split-BF16 paths."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from blas_helpers import U32, bound_ratio, cmat, complex_eq, flat
from clgpu_amd import datagen
from clgpu_amd import iclblas as ib

pytestmark = pytest.mark.gpu
KATS = json.loads((Path(__file__).parent / "golden" / "kats.json").read_text())
MODES = [ib.MATH_FP32, ib.MATH_SPLIT_BF16X3]
IMPLS = [n for n in ib.implementation_names() if n.startswith("cgemm")]


@pytest.fixture
def h(handle):
    handle.set_math_mode(ib.MATH_FP32)
    handle.set_implementation(None)
    yield handle
    handle.set_math_mode(ib.MATH_FP32)
    handle.set_implementation(None)


def run(h, ta, tb, m, n, k, alpha, A, B, beta, C):
    ib.check(ib.iclblasCgemm(h.value, ta, tb, m, n, k, alpha, A, A.shape[0], B, B.shape[0], beta, C, C.shape[0]),
             "iclblasCgemm")
    return C


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("case", KATS["cgemm"], ids=lambda c: c["name"])
def test_cgemm_kat(h, case, mode):
    c = case
    h.set_math_mode(mode)
    A = cmat(c["A"], c["lda"], c["k"] if c["transa"] == 0 else c["m"])
    B = cmat(c["B"], c["ldb"], c["n"] if c["transb"] == 0 else c["k"])
    C = cmat(c["C"], c["ldc"], c["n"])
    run(h, c["transa"], c["transb"], c["m"], c["n"], c["k"], complex(*c["alpha"]), A, B, complex(*c["beta"]), C)
    exp = np.asarray(c["expected"], np.float64)
    assert complex_eq(flat(C), exp[:, 0] + 1j * exp[:, 1])


def _qcmat(rows, cols, ld, stream):
    v = datagen.quarter_integers(2 * ld * cols, seed=13, stream=stream)
    z = (v[0::2] + 1j * v[1::2]).astype(np.complex64)
    return np.asfortranarray(z.reshape(cols, ld).T)


@pytest.mark.parametrize("impl", IMPLS)
@pytest.mark.parametrize("ta", [0, 1, 2])
@pytest.mark.parametrize("tb", [0, 1, 2])
@pytest.mark.parametrize("beta_zero", [False, True])
def test_cgemm_quarter_integer_grid_exact(h, impl, ta, tb, beta_zero):
    m, n, k = 96, 80, 112
    pad = 29 if (ta + tb) % 2 else 0
    ra, ca = (m, k) if ta == 0 else (k, m)
    rb, cb = (k, n) if tb == 0 else (n, k)
    A = _qcmat(ra, ca, ra + pad, 0)
    B = _qcmat(rb, cb, rb + pad, 1)
    C = _qcmat(m, n, m + pad, 2)
    alpha = 0.75 - 0.5j
    beta = 0.0 if beta_zero else -0.25 + 1.0j
    if beta_zero:
        C[:m] = np.nan
    ref = C.copy(order="F")
    oracle.cgemm(ta, tb, m, n, k, alpha, A, B, beta, ref)
    h.set_implementation(impl)
    run(h, ta, tb, m, n, k, alpha, A, B, beta, C)
    assert h.last_implementation() == impl
    assert np.array_equal(C, ref, equal_nan=True)


@pytest.mark.parametrize("ta", [0, 1, 2])
@pytest.mark.parametrize("tb", [0, 1, 2])
def test_cgemm_f32_dma_all_ops_exact(h, ta, tb):
    """The LDS-DMA FP32 kernel on a problem without edges (M, N multiples of
    128, K of 8, even leading dimensions), every op pair: each operand is copied
    in the rows-contiguous or the k-contiguous layout by its op, op C negates
    the imaginary fragment, and the k order inside a slab is permuted for both
    operands alike. Quarter-integer data make every sum exact, so the result
    equals cpu_gemm<complex<float>> bit for bit."""
    m, n, k = 256, 128, 136
    ra, ca = (m, k) if ta == 0 else (k, m)
    rb, cb = (k, n) if tb == 0 else (n, k)
    A = _qcmat(ra, ca, ra + 2, 0)
    B = _qcmat(rb, cb, rb + 4, 1)
    C = _qcmat(m, n, m + 1, 2)
    alpha, beta = 0.75 - 0.5j, -0.25 + 1.0j
    ref = C.copy(order="F")
    oracle.cgemm(ta, tb, m, n, k, alpha, A, B, beta, ref)
    h.set_implementation("cgemm_f32_dma_128x128")
    run(h, ta, tb, m, n, k, alpha, A, B, beta, C)
    assert h.last_implementation() == "cgemm_f32_dma_128x128"
    assert np.array_equal(C, ref)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("ta,tb", [(0, 0), (2, 0), (1, 2), (2, 2), (0, 1)])
@pytest.mark.parametrize("shape", [(1, 1, 1), (37, 21, 53), (130, 70, 200)], ids=lambda s: "x".join(map(str, s)))
def test_cgemm_error_bound(h, mode, ta, tb, shape):
    m, n, k = shape
    h.set_math_mode(mode)
    ra, ca = (m, k) if ta == 0 else (k, m)
    rb, cb = (k, n) if tb == 0 else (n, k)
    A = datagen.complex_matrix(ra, ca, ra + 1, "A", seed=31)
    B = datagen.complex_matrix(rb, cb, rb + 2, "B", seed=31)
    C = datagen.complex_matrix(m, n, m + 1, "C", seed=31)
    alpha, beta = 1 + 0.5j, 0.25
    C64, ab = oracle.zgemm(ta, tb, m, n, k, alpha, A, B, beta, C)
    C0 = C.copy(order="F")
    run(h, ta, tb, m, n, k, alpha, A, B, beta, C)
    err = np.abs(C[:m].astype(np.complex128) - C64)
    extra = 4 * U32 * (abs(beta) * np.abs(C0[:m]).astype(np.float64) + abs(alpha) * ab)
    assert bound_ratio(err, k, abs(alpha) * ab, extra) <= 4.0
    assert np.array_equal(C[m:], C0[m:])
    split = mode == ib.MATH_SPLIT_BF16X3 and k >= 256
    assert h.last_implementation().startswith("cgemm_bf16x3" if split else "cgemm_f32")


@pytest.mark.parametrize("impl", [n for n in IMPLS if "bf16x3" in n])
def test_cgemm_split_forced_small_k_bound(h, impl):
    m, n, k = 40, 24, 20
    A = datagen.complex_matrix(m, k, name="A", seed=33)
    B = datagen.complex_matrix(k, n, name="B", seed=33)
    C = np.zeros((m, n), np.complex64, order="F")
    h.set_implementation(impl)
    run(h, 0, 0, m, n, k, 1.0, A, B, 0.0, C)
    assert h.last_implementation() == impl
    C64, ab = oracle.zgemm(0, 0, m, n, k, 1.0, A, B)
    # each real product of the complex MAC carries at most 3*2^-16 |x||y|
    assert np.all(np.abs(C - C64) <= (3 * 2.0 ** -16 + 4 * k * 2.0 ** -24) * 2 * ab)


def _to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a.T)).cuda()


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("ta", [0, 2])
def test_cgemm_config3_4096(dev_handle, mode, ta):
    """CGEMM 4096^3, NN and CN, alpha = (1, 0.5), beta = (0.25, 0)."""
    import torch
    n = 4096
    A = datagen.complex_matrix(n, n, name="A", seed=0)
    B = datagen.complex_matrix(n, n, name="B", seed=0)
    C = datagen.complex_matrix(n, n, name="C", seed=0)
    h = dev_handle
    h.set_math_mode(mode)
    dA, dB, dC = _to_dev(A), _to_dev(B), _to_dev(C)
    alpha, beta = 1 + 0.5j, 0.25
    ib.check(ib.iclblasCgemm(h.value, ta, 0, n, n, n, alpha, dA.data_ptr(), n, dB.data_ptr(), n, beta, dC.data_ptr(),
                             n))
    torch.cuda.synchronize()
    rng = np.random.default_rng(ta)
    rows = rng.integers(0, n, 256)
    cols = rng.integers(0, n, 256)
    got = dC.t()[torch.from_numpy(rows).cuda(), torch.from_numpy(cols).cuda()].cpu().numpy().astype(np.complex128)
    c64, ab = oracle.zgemm_sample(ta, 0, n, alpha, A, B, rows, cols, beta, C)
    extra = 4 * U32 * (abs(beta) * np.abs(C[rows, cols]).astype(np.float64) + abs(alpha) * ab)
    assert bound_ratio(np.abs(got - c64), n, abs(alpha) * ab, extra) <= 4.0
    h.set_math_mode(ib.MATH_FP32)


@pytest.mark.parametrize("tb", [0, 2], ids=["BN", "BC"])
@pytest.mark.parametrize("beta", [0.0, 0.5 - 0.25j], ids=["beta0", "beta"])
def test_cgemm_host_panel_pipeline(h, tb, beta):
    """Host pointers at a size that takes the panel pipeline (B and C in column
    panels, transfers overlapped with the compute, a partial last panel), padded
    leading dimensions, B as is and conjugate-transposed, C not read when
    beta == 0; against a complex128 product within the elementwise bound."""
    m, n, k = 1024, 2304, 2048
    rb, cb = (k, n) if tb == 0 else (n, k)
    A = datagen.complex_matrix(m, k, m + 1, "A", seed=32)
    B = datagen.complex_matrix(rb, cb, rb + 3, "B", seed=32)
    C = datagen.complex_matrix(m, n, m + 2, "C", seed=32)
    C[m:] = -7.0 + 1.0j
    if beta == 0.0:
        C[:m] = np.nan
    C0 = C.copy(order="F")
    alpha = 0.5 + 0.75j
    run(h, 0, tb, m, n, k, alpha, A, B, beta, C)
    A64 = A[:m].astype(np.complex128)
    opB = B[:rb].astype(np.complex128) if tb == 0 else B[:rb].astype(np.complex128).conj().T
    ref = alpha * (A64 @ opB)
    mag = abs(alpha) * (np.abs(A64) @ np.abs(opB))
    if beta != 0.0:
        ref += beta * C0[:m].astype(np.complex128)
        mag += abs(beta) * np.abs(C0[:m].astype(np.complex128))
    err = np.abs(C[:m].astype(np.complex128) - ref)
    assert np.all(err <= 8 * (k + 2) * U32 * mag), float((err / np.maximum(mag, 1e-30)).max())
    assert np.all(C[m:] == np.complex64(-7.0 + 1.0j))
