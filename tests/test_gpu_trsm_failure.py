"""Failure reporting of the dataflow STRSM (strsm_dataflow64) through the C ABI.

Its waits on other workgroups' tiles are bounded; a wait that runs out sets an
abort word so the launch drains, and the call must then report
ICLBLAS_STATUS_EXECUTION_FAILED (the reference maps device failures to that
status: libs/blas/shlib/iclBLASImpl.hpp:45-73). The condition is forced from
the host: ICLBLAS_DATAFLOW_WAIT_POLLS < 0 makes the kernel abort its first wait
without polling, and every kernel still exits normally. Host-pointer calls
report the failure themselves; in the device-pointer mode iclblasSynchronize
reports it. The next call on the same handle must succeed and be correct."""
import os

import numpy as np
import pytest

import oracle
from clgpu_amd import datagen
from clgpu_amd import iclblas as ib

pytestmark = pytest.mark.gpu
ENV = "ICLBLAS_DATAFLOW_WAIT_POLLS"
M, N = 256, 64  # four 64-row blocks: every tile below the first waits on another


def _problem():
    A = datagen.triangular(M, lower=True, seed=7, poison=True)
    B = datagen.matrix(M, N, name="B", seed=7)
    X64 = oracle.dtrsm(0, 1, 0, 0, M, N, 1.0, A, B)
    return A, B, X64


def _forced_abort(fn):
    os.environ[ENV] = "-1"
    try:
        return fn()
    finally:
        del os.environ[ENV]


def _ok(X, X64):
    return np.abs(X - X64).max() / np.abs(X64).max() <= 4 * M * 2.0 ** -24


def test_dataflow_abort_reported_in_host_mode(handle):
    h = handle
    h.set_math_mode(ib.MATH_FP32)
    h.set_implementation("strsm_dataflow64")
    try:
        A, B, X64 = _problem()
        X = B.copy(order="F")
        st = _forced_abort(lambda: ib.iclblasStrsm(h.value, 0, 1, 0, 0, M, N, 1.0, A, M, X, M))
        assert st == ib.STATUS_EXECUTION_FAILED
        assert h.last_implementation() == "strsm_dataflow64"
        # the next calls succeed, with the dataflow solve and with the default one
        X = B.copy(order="F")
        ib.check(ib.iclblasStrsm(h.value, 0, 1, 0, 0, M, N, 1.0, A, M, X, M))
        assert _ok(X, X64)
        h.set_implementation(None)
        X = B.copy(order="F")
        ib.check(ib.iclblasStrsm(h.value, 0, 1, 0, 0, M, N, 1.0, A, M, X, M))
        assert _ok(X, X64)
    finally:
        h.set_implementation(None)


def test_dataflow_abort_reported_by_synchronize_in_device_mode(dev_handle):
    import torch
    h = dev_handle
    h.set_math_mode(ib.MATH_FP32)
    h.set_implementation("strsm_dataflow64")
    try:
        A, B, X64 = _problem()
        dev = torch.device("cuda:0")
        tA = torch.from_numpy(np.ascontiguousarray(A.T)).to(dev)  # column-major storage as (cols, ld)
        tB = torch.from_numpy(np.ascontiguousarray(B.T)).to(dev)
        X = tB.clone()
        st = _forced_abort(lambda: ib.iclblasStrsm(h.value, 0, 1, 0, 0, M, N, 1.0, tA.data_ptr(), M, X.data_ptr(), M))
        assert st == ib.STATUS_SUCCESS  # queued; the failure shows at the synchronising call
        assert ib.load().iclblasSynchronize(h.value) == ib.STATUS_EXECUTION_FAILED
        assert ib.load().iclblasSynchronize(h.value) == ib.STATUS_SUCCESS  # reported once
        X = tB.clone()
        ib.check(ib.iclblasStrsm(h.value, 0, 1, 0, 0, M, N, 1.0, tA.data_ptr(), M, X.data_ptr(), M))
        ib.check(ib.load().iclblasSynchronize(h.value))
        assert _ok(X.cpu().numpy().T, X64)
    finally:
        h.set_implementation(None)
