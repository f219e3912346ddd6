"""SGEMM parity on the GPU through the C ABI (libiclBLAS.so), against the
oracle: the reference's known answers, its implementation-level grid
(tests/blas_funcs/test_gemm.cpp:56-72, quarter-integer data, exact equality),
and the FP64 elementwise bound |C - C64|_ij <= 4*k*2^-24*(|A||B|)_ij for
random data, on the plain FP32-MFMA and the split-BF16 paths."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from blas_helpers import U32, bound_ratio, flat, float_eq, fmat
from clgpu_amd import datagen
from clgpu_amd import iclblas as ib

pytestmark = pytest.mark.gpu
KATS = json.loads((Path(__file__).parent / "golden" / "kats.json").read_text())
MODES = [ib.MATH_FP32, ib.MATH_SPLIT_BF16X3]

def _registered(prefix):
    from clgpu_amd import iclblas
    return [n for n in iclblas.implementation_names() if n.startswith(prefix)]


F32_IMPLS = _registered("sgemm_f32")
BF16_IMPLS = _registered("sgemm_bf16x3")


@pytest.fixture
def h(handle):
    handle.set_math_mode(ib.MATH_FP32)
    handle.set_implementation(None)
    yield handle
    handle.set_math_mode(ib.MATH_FP32)
    handle.set_implementation(None)


def run(h, ta, tb, m, n, k, alpha, A, B, beta, C):
    ib.check(ib.iclblasSgemm(h.value, ta, tb, m, n, k, alpha, A, A.shape[0], B, B.shape[0], beta, C, C.shape[0]),
             "iclblasSgemm")
    return C


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("case", KATS["sgemm"], ids=lambda c: c["name"])
def test_sgemm_kat(h, case, mode):
    c = case
    h.set_math_mode(mode)
    A = fmat(c["A"], c["lda"], c["k"] if c["transa"] == 0 else c["m"])
    B = fmat(c["B"], c["ldb"], c["n"] if c["transb"] == 0 else c["k"])
    C = fmat(c["C"], c["ldc"], c["n"])
    C0 = C.copy(order="F")
    run(h, c["transa"], c["transb"], c["m"], c["n"], c["k"], c["alpha"], A, B, c["beta"], C)
    exp = np.asarray(c["expected"], np.float64)
    if c["compare"] == "ulp4":
        assert float_eq(flat(C), exp)
    else:
        _, ab = oracle.dgemm(c["transa"], c["transb"], c["m"], c["n"], c["k"], 1.0, A, B)
        err = np.abs(C[: c["m"]].astype(np.float64) - exp.reshape(c["n"], c["ldc"]).T[: c["m"]])
        assert bound_ratio(err, c["k"], ab, extra=U32 * np.abs(C0[: c["m"]]).astype(np.float64) * 2) <= 4.0


# ---- the reference's implementation-level grid, every implementation forced by name
_grid_cache = {}


def _grid_inputs(lda_add, ldb_add, ldc_add, beta_zero):
    key = (lda_add, ldb_add, ldc_add, beta_zero)
    if key not in _grid_cache:
        m = n = k = 256
        q = lambda cnt, s: datagen.quarter_integers(cnt, seed=11, stream=s).astype(np.float32)  # noqa: E731
        A = np.asfortranarray(q((m + lda_add) * k, 0).reshape(k, m + lda_add).T)
        B = np.asfortranarray(q((k + ldb_add) * n, 1).reshape(n, k + ldb_add).T)
        C = np.asfortranarray(q((m + ldc_add) * n, 2).reshape(n, m + ldc_add).T)
        alpha = float(q(1, 3)[0]) or 0.5
        beta = 0.0 if beta_zero else (float(q(1, 4)[0]) or -0.25)
        if beta_zero:
            C[:m] = np.nan  # beta == 0: C need not be initialised
        ref = C.copy(order="F")
        oracle.sgemm(0, 0, m, n, k, alpha, A, B, beta, ref)
        _grid_cache[key] = (A, B, C, alpha, beta, ref)
    return _grid_cache[key]


@pytest.mark.parametrize("impl", F32_IMPLS + BF16_IMPLS)
@pytest.mark.parametrize("lda_add", [0, 29])
@pytest.mark.parametrize("ldb_add", [0, 29])
@pytest.mark.parametrize("beta_zero", [False, True])
@pytest.mark.parametrize("ldc_add", [0, 29])
def test_sgemm_quarter_integer_grid_exact(h, impl, lda_add, ldb_add, beta_zero, ldc_add):
    A, B, C, alpha, beta, ref = _grid_inputs(lda_add, ldb_add, ldc_add, beta_zero)
    h.set_implementation(impl)
    C = C.copy(order="F")
    run(h, 0, 0, 256, 256, 256, alpha, A, B, beta, C)
    assert h.last_implementation() == impl
    assert np.array_equal(C, ref, equal_nan=True)  # padding rows untouched, values exact


SHAPES = [(1, 1, 1), (3, 5, 7), (77, 45, 131), (128, 128, 16), (300, 257, 65), (513, 130, 1000)]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1), (2, 0), (0, 2)])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_sgemm_error_bound(h, mode, ta, tb, shape):
    m, n, k = shape
    h.set_math_mode(mode)
    ra, ca = (m, k) if ta == 0 else (k, m)
    rb, cb = (k, n) if tb == 0 else (n, k)
    A = datagen.matrix(ra, ca, ra + 3, "A", seed=21)
    B = datagen.matrix(rb, cb, rb + 1, "B", seed=21)
    C = datagen.matrix(m, n, m + 2, "C", seed=21)
    C[m:] = -7.0  # sentinel padding
    C0 = C.copy(order="F")
    alpha, beta = -1.5, 0.5
    C64, ab = oracle.dgemm(ta, tb, m, n, k, alpha, A, B, beta, C)
    run(h, ta, tb, m, n, k, alpha, A, B, beta, C)
    err = np.abs(C[:m].astype(np.float64) - C64)
    extra = 2 * U32 * (abs(beta) * np.abs(C0[:m].astype(np.float64)) + abs(alpha) * ab)
    assert bound_ratio(err, k, abs(alpha) * ab, extra) <= 4.0
    assert np.all(C[m:] == -7.0)
    # split-BF16 requests with k < 256 run on FP32 MFMA (the split cannot meet the bound there)
    split = mode == ib.MATH_SPLIT_BF16X3 and k >= 256
    assert h.last_implementation().startswith("sgemm_bf16x3" if split else "sgemm_f32")


@pytest.mark.parametrize("impl", ["sgemm_f32_dma_256x256w8", "sgemm_f32_dma_128x128"])
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_sgemm_f32_dma_all_ops(h, ta, tb, impl):
    """The LDS-DMA FP32 kernel on a problem without edges (M, N multiples of
    256, K of 16, leading dimensions multiples of 4), in every transpose case:
    each operand is copied in the rows-contiguous or the k-contiguous layout by
    its transpose, and the k order inside a slab is permuted for both alike."""
    m, n, k = 512, 256, 272
    ra, ca = (m, k) if ta == 0 else (k, m)
    rb, cb = (k, n) if tb == 0 else (n, k)
    A = datagen.matrix(ra, ca, ra + 4, "A", seed=24)
    B = datagen.matrix(rb, cb, rb + 8, "B", seed=24)
    C = datagen.matrix(m, n, m + 2, "C", seed=24)
    C[m:] = -7.0
    C0 = C.copy(order="F")
    alpha, beta = 0.75, -0.5
    C64, ab = oracle.dgemm(ta, tb, m, n, k, alpha, A, B, beta, C)
    h.set_implementation(impl)
    run(h, ta, tb, m, n, k, alpha, A, B, beta, C)
    assert h.last_implementation() == impl
    err = np.abs(C[:m].astype(np.float64) - C64)
    extra = 2 * U32 * (abs(beta) * np.abs(C0[:m].astype(np.float64)) + abs(alpha) * ab)
    assert bound_ratio(err, k, abs(alpha) * ab, extra) <= 4.0
    assert np.all(C[m:] == -7.0)


@pytest.mark.parametrize("impl", BF16_IMPLS)
@pytest.mark.parametrize("shape", [(1, 1, 1), (33, 17, 16), (130, 70, 100)], ids=lambda s: "x".join(map(str, s)))
def test_sgemm_split_forced_small_k_bound(h, impl, shape):
    """The split kernel itself at small k: per product it leaves at most
    3*2^-16 |a||b| (residual of the two-term split plus the dropped lo*lo term)."""
    m, n, k = shape
    A = datagen.matrix(m, k, name="A", seed=23)
    B = datagen.matrix(k, n, name="B", seed=23)
    C = np.zeros((m, n), np.float32, order="F")
    h.set_implementation(impl)
    run(h, 0, 0, m, n, k, 1.0, A, B, 0.0, C)
    assert h.last_implementation() == impl
    C64, ab = oracle.dgemm(0, 0, m, n, k, 1.0, A, B)
    assert np.all(np.abs(C - C64) <= (3 * 2.0 ** -16 + 4 * k * U32) * ab)


@pytest.mark.parametrize("mode", MODES)
def test_sgemm_beta_zero_does_not_read_c_and_alpha_zero_does_not_read_ab(h, mode):
    h.set_math_mode(mode)
    m, n, k = 70, 33, 40
    A = datagen.matrix(m, k, name="A")
    B = datagen.matrix(k, n, name="B")
    C = np.full((m, n), np.nan, np.float32, order="F")
    run(h, 0, 0, m, n, k, 1.0, A, B, 0.0, C)
    assert np.all(np.isfinite(C))
    An = np.full_like(A, np.nan)
    C2 = datagen.matrix(m, n, name="C")
    exp = (np.float32(0.25) * C2).astype(np.float32)
    run(h, 0, 0, m, n, k, 0.0, An, An, 0.25, C2)
    assert np.array_equal(C2, exp)
    C3 = datagen.matrix(m, n, name="C")
    exp3 = (np.float32(-2.0) * C3).astype(np.float32)
    run(h, 0, 0, m, n, 0, 1.0, A, B, -2.0, C3)  # k == 0: C = beta*C
    assert np.array_equal(C3, exp3)


@pytest.mark.parametrize("mode", MODES)
def test_sgemm_config1_256_beta_half(h, mode):
    """SGEMM NN 256^3, ld 256, alpha 1, beta 0.5, uniform[-1,1] from seed 0."""
    h.set_math_mode(mode)
    m = n = k = 256
    A = datagen.matrix(m, k, name="A", seed=0)
    B = datagen.matrix(k, n, name="B", seed=0)
    C = datagen.matrix(m, n, name="C", seed=0)
    C64, ab = oracle.dgemm(0, 0, m, n, k, 1.0, A, B, 0.5, C)
    Cc = C.copy(order="F")
    oracle.sgemm(0, 0, m, n, k, 1.0, A, B, 0.5, Cc)
    run(h, 0, 0, m, n, k, 1.0, A, B, 0.5, C)
    ratio = bound_ratio(np.abs(C - C64), k, ab, U32 * 0.5 * np.abs(Cc) * 2)
    assert ratio <= 4.0
    # the reference CPU path (double accumulation) is within the same bound
    assert bound_ratio(np.abs(Cc - C64), k, ab, U32 * np.abs(Cc)) <= 4.0


# ---- full-size shapes in device memory (sampled FP64 check + size-independent identities)
def _to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a.T)).cuda()


def _from_dev(t):
    return np.asfortranarray(t.cpu().numpy().T)


@pytest.fixture(scope="module")
def big():
    n = 8192
    A = datagen.matrix(n, n, name="A", seed=0)
    B = datagen.matrix(n, n, name="B", seed=0)
    return A, B


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_sgemm_config2_8192(dev_handle, big, mode, ta, tb):
    import torch
    A, B = big
    n = 8192
    h = dev_handle
    h.set_math_mode(mode)
    dA, dB = _to_dev(A), _to_dev(B)
    dC = torch.empty((n, n), device="cuda", dtype=torch.float32)
    ib.check(ib.iclblasSgemm(h.value, ta, tb, n, n, n, 1.0, dA.data_ptr(), n, dB.data_ptr(), n, 0.0, dC.data_ptr(), n))
    torch.cuda.synchronize()
    rng = np.random.default_rng(ta * 2 + tb)
    rows = rng.integers(0, n, 384)
    cols = rng.integers(0, n, 384)
    got = dC.t()[torch.from_numpy(rows).cuda(), torch.from_numpy(cols).cuda()].cpu().numpy().astype(np.float64)
    c64, ab = oracle.dgemm_sample(ta, tb, n, 1.0, A, B, rows, cols)
    assert bound_ratio(np.abs(got - c64), n, ab) <= 4.0
    h.set_math_mode(ib.MATH_FP32)


@pytest.mark.parametrize("mode", MODES)
def test_sgemm_config2_8192_full_matrix(dev_handle, big, mode):
    """C2 NN at 8192^3, every one of the 8192^2 elements against a float64
    product formed on the device (the sampled oracle check above pins the float64
    product's convention on the same data): an error confined to a few tiles (a
    tile raster, XCD remap or partial-product launch mistake) cannot pass."""
    import torch
    A, B = big
    n = 8192
    h = dev_handle
    h.set_math_mode(mode)
    dA, dB = _to_dev(A), _to_dev(B)  # storage (cols, ld): the transposes of the column-major matrices
    dC = torch.empty((n, n), device="cuda", dtype=torch.float32)
    ib.check(ib.iclblasSgemm(h.value, 0, 0, n, n, n, 1.0, dA.data_ptr(), n, dB.data_ptr(), n, 0.0, dC.data_ptr(), n))
    torch.cuda.synchronize()
    h.set_math_mode(ib.MATH_FP32)
    # storage(C) = C^T = B^T A^T
    At, Bt = dA.double(), dB.double()
    ref = Bt @ At
    err = (dC.double() - ref).abs()
    del ref
    absab = Bt.abs() @ At.abs()
    ratio = float((err / (n * 2.0 ** -24 * absab).clamp_min(1e-300)).max())
    assert ratio <= 4.0, ratio
    # the sampled oracle elements agree with the device product used here
    rows = np.array([0, 1, 4095, 4096, 8191])
    cols = np.array([8191, 4096, 1, 4095, 0])
    c64, _ = oracle.dgemm_sample(0, 0, n, 1.0, A, B, rows, cols)
    dev64 = (Bt @ At)[torch.from_numpy(cols).cuda(), torch.from_numpy(rows).cuda()].cpu().numpy()
    assert np.allclose(dev64, c64, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("mode", MODES)
def test_sgemm_config2_nonsquare_padded(dev_handle, mode):
    import torch
    m, n, k, ld = 4096, 8192, 2048, 4160
    A = datagen.matrix(m, k, ld, "A", seed=0)
    B = datagen.matrix(k, n, ld, "B", seed=0)
    h = dev_handle
    h.set_math_mode(mode)
    dA, dB = _to_dev(A), _to_dev(B)
    dC = torch.full((n, ld), -3.0, device="cuda", dtype=torch.float32)
    ib.check(ib.iclblasSgemm(h.value, 0, 0, m, n, k, 1.0, dA.data_ptr(), ld, dB.data_ptr(), ld, 0.0, dC.data_ptr(), ld))
    torch.cuda.synchronize()
    Ct = dC.t()
    assert bool((Ct[m:] == -3.0).all())  # padding rows of C untouched
    rng = np.random.default_rng(7)
    rows = rng.integers(0, m, 384)
    cols = rng.integers(0, n, 384)
    got = Ct[torch.from_numpy(rows).cuda(), torch.from_numpy(cols).cuda()].cpu().numpy().astype(np.float64)
    c64, ab = oracle.dgemm_sample(0, 0, k, 1.0, A, B, rows, cols)
    assert bound_ratio(np.abs(got - c64), k, ab) <= 4.0
    # linearity in B at full size: op(A)(B1 + B2) == op(A)B1 + op(A)B2 to rounding
    dC2 = torch.empty((n, ld), device="cuda", dtype=torch.float32)
    dB2 = dB * 0.5
    ib.check(ib.iclblasSgemm(h.value, 0, 0, m, n, k, 2.0, dA.data_ptr(), ld, dB2.data_ptr(), ld, 0.0, dC2.data_ptr(),
                             ld))
    torch.cuda.synchronize()
    assert bool((dC2.t()[:m] == Ct[:m]).all())  # 2 * (A (B/2)) is bitwise A B (power-of-two scaling)
    h.set_math_mode(ib.MATH_FP32)


@pytest.mark.parametrize("ta,tb,m", [(0, 0, 8192), (1, 1, 8192), (0, 1, 8448), (1, 0, 8448)])
def test_sgemm_split_pack_overlap_bit_identical(dev_handle, ta, tb, m):
    """A large split-BF16 product (>= 1024 tiles) packs half of its operands on
    the handle's side stream and runs as four quadrant launches; the
    allocation-free entry point runs the same kernel in one launch after a
    single prepass. Both give the same bits (same per-element summation order),
    including an M whose halves are not equal (8448 = 4352 + 4096) and a K that
    is not a multiple of the 32-deep slab."""
    import torch
    n, k, ld = 8192, 300, 8512
    rows_a, cols_a = (m, k) if ta == 0 else (k, m)
    rows_b, cols_b = (k, n) if tb == 0 else (n, k)
    A = datagen.matrix(ld, cols_a, name="A", seed=21)
    B = datagen.matrix(ld, cols_b, name="B", seed=21)
    C = datagen.matrix(ld, n, name="C", seed=21)
    dA, dB = _to_dev(A), _to_dev(B)
    outs = []
    for no_alloc in (False, True):
        h = dev_handle
        h.set_math_mode(ib.MATH_SPLIT_BF16X3)
        dC = _to_dev(C)
        fn = ib.iclblasSgemmNoAlloc if no_alloc else ib.iclblasSgemm
        ib.check(fn(h.value, ta, tb, m, n, k, 0.5, dA.data_ptr(), ld, dB.data_ptr(), ld, -1.0, dC.data_ptr(), ld))
        torch.cuda.synchronize()
        assert h.last_implementation().startswith("sgemm_bf16x3d")
        outs.append(dC.cpu().numpy())
        h.set_math_mode(ib.MATH_FP32)
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[0][:, m:], np.ascontiguousarray(C.T)[:, m:])  # rows m.. of C untouched
    rng = np.random.default_rng(5)
    rows, cols = rng.integers(0, m, 256), rng.integers(0, n, 256)
    got = outs[0].T[rows, cols].astype(np.float64)
    c64, ab = oracle.dgemm_sample(ta, tb, k, 0.5, A, B, rows, cols, beta=-1.0, C=C)
    assert bound_ratio(np.abs(got - c64), k, 0.5 * ab, U32 * np.abs(C[rows, cols].astype(np.float64))) <= 4.0


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("tb", [0, 1], ids=["BN", "BT"])
@pytest.mark.parametrize("beta", [0.0, -0.5], ids=["beta0", "beta"])
def test_sgemm_host_panel_pipeline(h, mode, tb, beta):
    """Host pointers at a size where the FP32 math mode takes the panel
    pipeline (B and C in column panels, transfers overlapped with the compute, a
    partial last panel) and the split mode one transfer per matrix; padded
    leading dimensions, both layouts of B, C not read when beta == 0; against a
    float64 product within the elementwise bound."""
    m, n, k = 2048, 2304, 1024
    h.set_math_mode(mode)
    rb, cb = (k, n) if tb == 0 else (n, k)
    A = datagen.matrix(m, k, m + 3, "A", seed=31)
    B = datagen.matrix(rb, cb, rb + 5, "B", seed=31)
    C = datagen.matrix(m, n, m + 7, "C", seed=31)
    C[m:] = -7.0
    if beta == 0.0:
        C[:m] = np.nan
    C0 = C.copy(order="F")
    alpha = 0.75
    run(h, 0, tb, m, n, k, alpha, A, B, beta, C)
    A64 = A[:m].astype(np.float64)
    opB = B[:rb].astype(np.float64) if tb == 0 else B[:rb].astype(np.float64).T
    ref = alpha * (A64 @ opB)
    mag = abs(alpha) * (np.abs(A64) @ np.abs(opB))
    if beta != 0.0:
        ref += beta * C0[:m].astype(np.float64)
        mag += abs(beta) * np.abs(C0[:m].astype(np.float64))
    err = np.abs(C[:m].astype(np.float64) - ref)
    assert np.all(err <= 4 * (k + 2) * U32 * mag), float((err / np.maximum(mag, 1e-30)).max())
    assert np.all(C[m:] == -7.0)
