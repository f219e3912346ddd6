"""SGEMV / CGEMV on the GPU through the C ABI:
  - the reference's 15 known answers (tests/golden/kats_gemv.json): the literal
    CGEMV expectations with the reference's ComplexEqual (4 ULP per part), and
    every case against the FP64 oracle within the elementwise bound;
  - every op (N, T, C) and stride against the FP64 oracle at shapes that are and
    are not multiples of the kernels' row / column blocks, including shapes that
    split the reduction over the grid, within |err| <= 4 (k + 2) 2^-24
    (|alpha| |op(A)||x| + |beta||y|);
  - NaN in A's padding rows, in the gaps of x and (beta == 0) in y: never read;
    the gaps of y keep their values; alpha == 0 scales y without reading A or x;
  - a 16384 x 16384 problem in device memory against a float64 product."""
import numpy as np
import pytest

import oracle
from blas_helpers import complex_eq
from clgpu_amd import iclblas as ib
from gemv_helpers import KATS, expected, kat_arrays, random_problem, scalars

pytestmark = pytest.mark.gpu
U = 2.0 ** -24


@pytest.fixture
def h(handle):
    handle.set_math_mode(ib.MATH_FP32)
    handle.set_implementation(None)
    yield handle


def _call(h, cplx, trans, m, n, alpha, A, x, incx, beta, y, incy):
    fn = ib.iclblasCgemv if cplx else ib.iclblasSgemv
    ib.check(fn(h.value, trans, m, n, alpha, A, A.shape[0], x, incx, beta, y, incy), "gemv")


def _check(trans, m, n, alpha, A, x, incx, beta, y0, incy, got):
    res, mag = oracle.zgemv(trans, m, n, alpha, A, x, incx, beta, y0, incy)
    leny = len(res)
    g = np.asarray(got[::incy][:leny], np.complex128)
    k = n if trans == 0 else m
    yb = np.abs(np.asarray(y0[::incy][:leny], np.complex128)) if beta != 0 else 0.0
    bound = 4.0 * (k + 2) * U * (abs(alpha) * mag + abs(beta) * yb) + 1e-30
    err = np.abs(g - res)
    assert np.all(err <= bound), float((err / bound).max())
    # the gaps between the elements of y are untouched
    gaps = np.ones(len(got), bool)
    gaps[::incy][:leny] = False
    assert np.array_equal(got[gaps], y0[gaps])


@pytest.mark.parametrize("case", KATS, ids=lambda c: f"{c['routine']}-{c['name']}")
def test_gemv_kats(h, case):
    A, x, y = kat_arrays(case)
    al, be = scalars(case)
    cplx = case["routine"] == "Cgemv"
    got = y.copy()
    _call(h, cplx, case["trans"], case["m"], case["n"], al, A, x, case["incx"], be, got, case["incy"])
    if cplx:
        ref = expected(case, "reference_expected").astype(np.complex64)
        assert all(complex_eq(e, g) for e, g in zip(ref, got)), (ref, got)
    _check(case["trans"], case["m"], case["n"], al, A, x, case["incx"], be, y, case["incy"], got)
    assert h.last_implementation() in ("sgemv_hbm", "cgemv_hbm", "sgemv_hbm_split", "cgemv_hbm_split")


SHAPES = [(1, 1), (7, 3), (255, 257), (256, 1000), (1000, 33), (129, 4097), (20000, 40), (40, 20000), (3000, 2500)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s[0]}x{s[1]}")
@pytest.mark.parametrize("trans", [0, 1, 2], ids=["N", "T", "C"])
@pytest.mark.parametrize("cplx", [False, True], ids=["S", "C"])
@pytest.mark.parametrize("incs", [(1, 1), (2, 3)], ids=["inc1", "inc23"])
def test_gemv_vs_oracle(h, cplx, trans, shape, incs):
    m, n = shape
    rng = np.random.default_rng(m * 7 + n)
    A, x, y, al, be = random_problem(cplx, trans, m, n, *incs, rng)
    got = y.copy()
    _call(h, cplx, trans, m, n, al, A, x, incs[0], be, got, incs[1])
    _check(trans, m, n, al, A, x, incs[0], be, y, incs[1], got)


@pytest.mark.parametrize("trans", [0, 1, 2], ids=["N", "T", "C"])
@pytest.mark.parametrize("cplx", [False, True], ids=["S", "C"])
def test_gemv_beta_zero_does_not_read_y(h, cplx, trans):
    rng = np.random.default_rng(3)
    m, n = 600, 900
    A, x, y, al, _ = random_problem(cplx, trans, m, n, 1, 2, rng, beta=0.0)
    got = y.copy()
    _call(h, cplx, trans, m, n, al, A, x, 1, 0.0, got, 2)
    _check(trans, m, n, al, A, x, 1, 0.0, y, 2, got)


@pytest.mark.parametrize("cplx", [False, True], ids=["S", "C"])
def test_gemv_alpha_zero_scales_y_only(h, cplx):
    rng = np.random.default_rng(4)
    m, n = 50, 70
    A, x, y, _, _ = random_problem(cplx, 0, m, n, 1, 1, rng)
    A[:] = np.nan
    x[:] = np.nan
    be = (0.5 - 1j) if cplx else 0.5
    got = y.copy()
    _call(h, cplx, 0, m, n, 0.0, A, x, 1, be, got, 1)
    exp = be * y.astype(np.complex128)
    np.testing.assert_allclose(got, (exp if cplx else exp.real).astype(got.dtype), rtol=1e-6)
    assert h.last_implementation() == "scale_y"


@pytest.mark.parametrize("trans", [0, 1], ids=["N", "T"])
def test_sgemv_16384_device(dev_handle, trans):
    import torch
    n = 16384
    g = torch.Generator(device="cuda")
    g.manual_seed(1)
    A = torch.rand((n, n), generator=g, device="cuda") * 2 - 1  # column-major view: A^T of the tensor
    x = torch.rand(n, generator=g, device="cuda") * 2 - 1
    y = torch.rand(n, generator=g, device="cuda") * 2 - 1
    y0 = y.clone()
    ib.check(ib.iclblasSgemv(dev_handle.value, trans, n, n, 0.5, A.data_ptr(), n, x.data_ptr(), 1, -1.0,
                             y.data_ptr(), 1))
    torch.cuda.synchronize()
    M = A.t().double()  # the column-major matrix
    op = M if trans == 0 else M.t()
    ref = 0.5 * (op @ x.double()) - y0.double()
    mag = 0.5 * (op.abs() @ x.double().abs()) + y0.double().abs()
    err = (y.double() - ref).abs()
    assert bool((err <= 4 * (n + 2) * U * mag).all()), float((err / mag).max())
