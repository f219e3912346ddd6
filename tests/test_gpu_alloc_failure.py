"""ICLBLAS_STATUS_ALLOC_FAILED through the C ABI. The reference maps an
allocation failure to ALLOC_FAILED (libs/blas/shlib/iclBLASImpl.hpp:45-73);
here the handle's scratch cap (iclblasSetScratchLimit) makes a call that would
grow the handle's device buffers beyond it fail the same way, so the path runs
without exhausting the device. The failing call queues no kernel (its output is
untouched), and the handle works normally once the cap is lifted."""
import numpy as np
import pytest

import oracle
from blas_helpers import bound_ratio
from clgpu_amd import datagen
from clgpu_amd import iclblas as ib

pytestmark = pytest.mark.gpu


def _dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x.T)).cuda()


def test_device_split_sgemm_over_the_cap():
    import torch
    n = 1024  # packed split-BF16 operands: (1024 + 1024) * 1024 * 4 B = 8 MiB
    A = datagen.matrix(n, n, name="A", seed=3)
    B = datagen.matrix(n, n, name="B", seed=3)
    tA, tB = _dev(A), _dev(B)
    stream = torch.cuda.current_stream().cuda_stream
    with ib.Handle(memory_mode=ib.MEMORY_DEVICE, math_mode=ib.MATH_SPLIT_BF16X3, stream=stream) as h:
        h.set_scratch_limit(1 << 20)
        C = torch.full((n, n), 5.0, device="cuda")
        st = ib.iclblasSgemm(h.value, 0, 0, n, n, n, 1.0, tA.data_ptr(), n, tB.data_ptr(), n, 0.0, C.data_ptr(), n)
        assert st == ib.STATUS_ALLOC_FAILED
        h.synchronize()
        assert bool((C == 5.0).all())
        held, limit = h.scratch_usage()
        assert limit == 1 << 20 and held <= limit
        h.set_scratch_limit(0)
        ib.check(ib.iclblasSgemm(h.value, 0, 0, n, n, n, 1.0, tA.data_ptr(), n, tB.data_ptr(), n, 0.0,
                                 C.data_ptr(), n), "iclblasSgemm")
        h.synchronize()
        assert h.last_implementation().startswith("sgemm_bf16x3d")
        assert h.scratch_usage()[0] >= 8 << 20
        got = C.cpu().numpy().T
    with ib.Handle(memory_mode=ib.MEMORY_DEVICE, math_mode=ib.MATH_SPLIT_BF16X3, stream=stream) as h2:
        C2 = torch.empty((n, n), device="cuda")
        ib.check(ib.iclblasSgemm(h2.value, 0, 0, n, n, n, 1.0, tA.data_ptr(), n, tB.data_ptr(), n, 0.0,
                                 C2.data_ptr(), n), "iclblasSgemm")
        h2.synchronize()
        assert np.array_equal(got, C2.cpu().numpy().T)


def test_host_sgemm_staging_over_the_cap():
    m = n = k = 512  # staging: 1 MiB per matrix
    A = datagen.matrix(m, k, name="A", seed=4)
    B = datagen.matrix(k, n, name="B", seed=4)
    C = datagen.matrix(m, n, name="C", seed=4)
    with ib.Handle() as h:
        h.set_scratch_limit(1536 << 10)
        Cg = C.copy(order="F")
        st = ib.iclblasSgemm(h.value, 0, 0, m, n, k, 1.0, A, m, B, k, 0.5, Cg, m)
        assert st == ib.STATUS_ALLOC_FAILED
        assert np.array_equal(Cg, C)  # the host result is written only by a call that completes
        h.set_scratch_limit(0)
        ib.check(ib.iclblasSgemm(h.value, 0, 0, m, n, k, 1.0, A, m, B, k, 0.5, Cg, m), "iclblasSgemm")
    C64, ab = oracle.dgemm(0, 0, m, n, k, 1.0, A, B, 0.5, C)
    err = np.abs(Cg.astype(np.float64) - C64)
    assert bound_ratio(err, k, ab, extra=0.5 * 2.0 ** -24 * np.abs(C.astype(np.float64))) <= 4.0


def test_device_strsm_workspace_over_the_cap():
    import torch
    m, nr = 1024, 256
    T = datagen.triangular(m, lower=True, seed=5)
    R = datagen.matrix(m, nr, name="B", seed=5)
    tT, tR = _dev(T), _dev(R)
    stream = torch.cuda.current_stream().cuda_stream
    with ib.Handle(memory_mode=ib.MEMORY_DEVICE, stream=stream) as h:
        h.set_scratch_limit(64 << 10)
        X = tR.clone()
        st = ib.iclblasStrsm(h.value, 0, 1, 0, 0, m, nr, 1.0, tT.data_ptr(), m, X.data_ptr(), m)
        assert st == ib.STATUS_ALLOC_FAILED
        h.synchronize()
        assert bool(torch.equal(X, tR))
        h.set_scratch_limit(0)
        ib.check(ib.iclblasStrsm(h.value, 0, 1, 0, 0, m, nr, 1.0, tT.data_ptr(), m, X.data_ptr(), m), "iclblasStrsm")
        h.synchronize()
        Xh = X.cpu().numpy().T
    X64 = oracle.dtrsm(0, 1, 0, 0, m, nr, 1.0, T, R)
    assert float(np.abs(Xh - X64).max() / np.abs(X64).max()) <= 1e-5
