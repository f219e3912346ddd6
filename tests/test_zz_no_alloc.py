"""iclblasSgemmNoAlloc (the SGEMM entry point that allocates nothing) against
iclblasSgemm in the device-pointer memory mode, through the C ABI: bit-identical
results on the split-BF16 path, whose packed operands go to the library's static
device scratch, and on the FP32 path, for two calls in a row, the second one
smaller (it runs on what the first left in the scratch). A call whose packed
operands exceed the scratch is refused before anything is queued."""
import numpy as np
import pytest

from clgpu_amd import datagen
from clgpu_amd import iclblas as ib

pytestmark = pytest.mark.gpu


def _dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x.T)).cuda()  # column-major ld x cols


@pytest.fixture(scope="module")
def handles():
    import torch
    torch.cuda.init()
    stream = torch.cuda.current_stream().cuda_stream
    ref = ib.Handle(memory_mode=ib.MEMORY_DEVICE, stream=stream)
    # the no-alloc entry point takes device pointers in either memory mode
    na = ib.Handle(memory_mode=ib.MEMORY_HOST, stream=stream)
    yield ref, na
    na.close()
    ref.close()


def _pair(ref, na, mode, ta, tb, m, n, k, lda, ldb, ldc, alpha, beta, seed):
    import torch
    rows_a, cols_a = (m, k) if ta == 0 else (k, m)
    rows_b, cols_b = (k, n) if tb == 0 else (n, k)
    A = datagen.matrix(lda, cols_a, name="A", seed=seed)
    B = datagen.matrix(ldb, cols_b, name="B", seed=seed)
    C = datagen.matrix(ldc, n, name="C", seed=seed)
    tA, tB = _dev(A), _dev(B)
    c_ref, c_na = _dev(C), _dev(C)
    for h in (ref, na):
        h.set_math_mode(mode)
        h.set_implementation(None)
    ib.check(ib.iclblasSgemm(ref.value, ta, tb, m, n, k, alpha, tA.data_ptr(), lda, tB.data_ptr(), ldb, beta,
                             c_ref.data_ptr(), ldc), "iclblasSgemm")
    ref.synchronize()
    ib.check(ib.iclblasSgemmNoAlloc(na.value, ta, tb, m, n, k, alpha, tA.data_ptr(), lda, tB.data_ptr(), ldb, beta,
                                    c_na.data_ptr(), ldc), "iclblasSgemmNoAlloc")
    na.synchronize()
    torch.cuda.synchronize()
    assert na.last_implementation() == ref.last_implementation()
    return c_ref.cpu().numpy(), c_na.cpu().numpy(), na.last_implementation()


@pytest.mark.parametrize("mode", [ib.MATH_SPLIT_BF16X3, ib.MATH_FP32], ids=["split", "fp32"])
def test_no_alloc_matches_sgemm_twice(handles, mode):
    ref, na = handles
    # first call: 1024 x 1536 x 1024 NN (packed operands 10 MiB)
    r1, n1, impl1 = _pair(ref, na, mode, 0, 0, 1024, 1536, 1024, 1024, 1024, 1024, 1.0, 0.0, seed=11)
    if mode == ib.MATH_SPLIT_BF16X3:
        assert impl1.startswith("sgemm_bf16x3d"), impl1  # the packed kernel, i.e. the scratch was used
    assert np.array_equal(r1, n1)
    # second call, smaller, with edges and padding: 300 x 200 x 520 TN, beta != 0
    r2, n2, impl2 = _pair(ref, na, mode, 1, 0, 300, 200, 520, 530, 530, 310, 0.75, -0.5, seed=12)
    if mode == ib.MATH_SPLIT_BF16X3:
        assert impl2.startswith("sgemm_bf16x3"), impl2
    assert np.array_equal(r2, n2)
    assert np.isfinite(n2).all()


def test_no_alloc_refuses_oversized_scratch(handles):
    import torch
    _, na = handles
    na.set_math_mode(ib.MATH_SPLIT_BF16X3)
    na.set_implementation(None)
    m = n = k = 4096  # packed operands (4096 + 4096) * 4096 * 4 bytes = 128 MiB > 64 MiB
    assert (m + n) * k * 4 > ib.NOALLOC_SCRATCH_BYTES
    A = torch.zeros((k, m), device="cuda")
    B = torch.zeros((n, k), device="cuda")
    C = torch.full((n, m), 7.0, device="cuda")
    st = ib.iclblasSgemmNoAlloc(na.value, 0, 0, m, n, k, 1.0, A.data_ptr(), m, B.data_ptr(), k, 0.0, C.data_ptr(), m)
    assert st == ib.STATUS_INVALID_VALUE
    na.synchronize()
    torch.cuda.synchronize()
    assert bool((C == 7.0).all())  # nothing was queued
    # the FP32 path has no scratch and no such limit
    na.set_math_mode(ib.MATH_FP32)
    ib.check(ib.iclblasSgemmNoAlloc(na.value, 0, 0, m, n, k, 1.0, A.data_ptr(), m, B.data_ptr(), k, 0.0,
                                    C.data_ptr(), m), "iclblasSgemmNoAlloc")
    na.synchronize()
    torch.cuda.synchronize()
    assert bool((C == 0.0).all())


def test_no_alloc_argument_checks(handles):
    _, na = handles
    # the reference's GEMM checks, before any HIP call
    assert ib.iclblasSgemmNoAlloc(na.value, 0, 0, -1, 4, 4, 1.0, 16, 4, 16, 4, 0.0, 16, 4) == ib.STATUS_INVALID_VALUE
    assert ib.iclblasSgemmNoAlloc(na.value, 0, 0, 0, 4, 4, 1.0, None, 4, None, 4, 0.0, None, 4) == ib.STATUS_SUCCESS
    assert ib.iclblasSgemmNoAlloc(na.value, 0, 0, 4, 4, 4, 1.0, None, 4, 16, 4, 0.0, 16, 4) == ib.STATUS_INVALID_VALUE
    assert ib.iclblasSgemmNoAlloc(None, 0, 0, 4, 4, 4, 1.0, 16, 4, 16, 4, 0.0, 16, 4) == ib.STATUS_INVALID_VALUE
