"""Host-side checks of the C-ABI library that need no GPU: it loads, exports
every symbol include/*.h declares, and its argument checks return the
reference's status codes in the reference's order before any HIP call
(reference libs/blas/shlib/iclBLAS_S.cpp:487-510, iclBLAS_C.cpp:458-469,
iclBLASImpl.hpp:108-245, iclBLAS.cpp:35-59)."""
import ctypes as ct
import subprocess

import numpy as np
import pytest

from clgpu_amd import datagen
from clgpu_amd import iclblas as ib


def test_library_loads_and_exports_every_declared_symbol():
    declared = ib.declared_symbols()
    assert len(declared) == 78 + 19  # reference API + extensions
    assert ib.exported_symbols() == declared
    out = subprocess.run(["nm", "-D", "--defined-only", str(ib.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(declared) <= exported
    # nothing but the API leaks out of the library
    assert all(name.startswith("iclblas") for name in exported), sorted(n for n in exported
                                                                         if not n.startswith("iclblas"))


def test_reference_enum_and_status_values():
    assert (ib.OP_N, ib.OP_T, ib.OP_C) == (0, 1, 2)
    assert (ib.FILL_MODE_UPPER, ib.FILL_MODE_LOWER) == (0, 1)
    assert (ib.DIAG_NON_UNIT, ib.DIAG_UNIT) == (0, 1)
    assert (ib.SIDE_LEFT, ib.SIDE_RIGHT) == (0, 1)
    assert (ib.STATUS_SUCCESS, ib.STATUS_INVALID_VALUE, ib.STATUS_NOT_SUPPORTED, ib.STATUS_ERROR) == (0, 7, 15, 17)


class _Forged(ct.Structure):
    _fields_ = [("tag", ct.c_int), ("pad", ct.c_char * 252)]


@pytest.fixture
def forged():
    f = _Forged()
    f.tag = 0x1234
    return ct.cast(ct.pointer(f), ct.c_void_p), f


def _sgemm(h, m=3, n=3, k=3, A="a", lda=3, B="b", ldb=3, C="c", ldc=3, alpha=1.0, beta=0.0, ta=0, tb=0):
    bufs = {"a": np.zeros((max(lda, 1), 8), np.float32, order="F"), "b": np.zeros((max(ldb, 1), 8), np.float32, order="F"),
            "c": np.zeros((max(ldc, 1), 8), np.float32, order="F")}
    return ib.iclblasSgemm(h, ta, tb, m, n, k, alpha, bufs.get(A) if A else None, lda, bufs.get(B) if B else None,
                           ldb, beta, bufs.get(C) if C else None, ldc)


def test_sgemm_status_order(forged):
    bad_handle, _keep = forged
    INV, OK = ib.STATUS_INVALID_VALUE, ib.STATUS_SUCCESS
    # negative sizes first, whatever the handle
    assert _sgemm(None, m=-1) == INV
    assert _sgemm(None, k=-1) == INV
    # m == 0 / n == 0 succeed before the handle is looked at (iclBLASImpl.hpp:110-120, 237-238)
    assert _sgemm(None, m=0) == OK
    assert _sgemm(bad_handle, n=0, A=None) == OK
    # k == 0 is work (C = beta*C), so the handle is checked
    assert _sgemm(None, k=0) == INV
    # null operands and bad leading dimensions
    assert _sgemm(None, A=None) == INV
    assert _sgemm(None, lda=0) == INV
    assert _sgemm(None, B=None) == INV
    assert _sgemm(None, ldb=-2) == INV
    assert _sgemm(None, C=None) == INV
    assert _sgemm(None, ldc=0) == INV
    # checks the reference does not make: ld smaller than the rows, bad enum, null scalars
    assert _sgemm(None, m=4, lda=3, ldc=4) == INV
    assert _sgemm(None, ta=7) == INV
    assert _sgemm(None, alpha=None) == INV
    # all arguments fine, handle invalid
    assert _sgemm(None) == INV
    assert _sgemm(bad_handle) == INV


def test_cgemm_and_trsm_status_order(forged):
    bad_handle, _keep = forged
    z = np.zeros((4, 4), np.complex64, order="F")
    f = np.zeros((4, 4), np.float32, order="F")
    INV, OK = ib.STATUS_INVALID_VALUE, ib.STATUS_SUCCESS
    assert ib.iclblasCgemm(None, 0, 0, -1, 2, 2, 1, z, 4, z, 4, 0, z, 4) == INV
    assert ib.iclblasCgemm(None, 0, 0, 2, 0, 2, 1, z, 4, z, 4, 0, z, 4) == OK
    assert ib.iclblasCgemm(None, 2, 2, 2, 2, 2, 1, z, 4, z, 4, 0, None, 4) == INV
    assert ib.iclblasCgemm(bad_handle, 2, 2, 2, 2, 2, 1, z, 4, z, 4, 0, z, 4) == INV
    # STRSM: m == 0 or n == 0 -> SUCCESS before anything else (iclBLAS_S.cpp:503-504)
    assert ib.iclblasStrsm(None, 0, 0, 0, 0, 0, -5, None, None, 0, None, 0) == OK
    assert ib.iclblasStrsm(None, 0, 0, 0, 0, -1, 2, 1.0, f, 4, f, 4) == INV
    assert ib.iclblasStrsm(None, 0, 0, 0, 0, 2, 2, 1.0, None, 4, f, 4) == INV
    assert ib.iclblasStrsm(None, 0, 0, 0, 0, 2, 2, 1.0, f, 0, f, 4) == INV
    assert ib.iclblasStrsm(None, 0, 0, 0, 0, 2, 2, 1.0, f, 4, f, 1) == INV  # ldb < m
    assert ib.iclblasStrsm(None, 3, 0, 0, 0, 2, 2, 1.0, f, 4, f, 4) == INV  # side out of range
    assert ib.iclblasStrsm(bad_handle, 0, 0, 0, 0, 2, 2, 1.0, f, 4, f, 4) == INV
    assert ib.iclblasStrsmRefined(None, 0, 0, 0, 0, 2, 2, 1.0, f, 4, f, 4, 1) == INV
    # CTRSM: the same chain (iclBLAS_C.cpp:525-535 returns SUCCESS for m == 0 or n == 0 first)
    assert ib.iclblasCtrsm(None, 0, 0, 0, 0, 2, 0, None, None, 0, None, 0) == OK
    assert ib.iclblasCtrsm(None, 0, 0, 0, 0, -1, 2, 1, z, 4, z, 4) == INV
    assert ib.iclblasCtrsm(None, 0, 0, 2, 0, 2, 2, 1, z, 4, None, 4) == INV
    assert ib.iclblasCtrsm(None, 1, 0, 2, 0, 2, 5, 1, z, 4, z, 4) == INV  # lda < n for a right-side solve
    assert ib.iclblasCtrsm(None, 0, 0, 3, 0, 2, 2, 1, z, 4, z, 4) == INV  # trans out of range
    assert ib.iclblasCtrsm(bad_handle, 0, 0, 2, 1, 2, 2, 1, z, 4, z, 4) == INV


def test_level3_status_order(forged):
    """The level-3 entry points' own early exits (reference iclBLAS_S.cpp:387-420,
    512-520; iclBLAS_C.cpp:472-589), then the generic chain and the handle."""
    bad_handle, _keep = forged
    f = np.zeros((4, 4), np.float32, order="F")
    z = np.zeros((4, 4), np.complex64, order="F")
    INV, OK = ib.STATUS_INVALID_VALUE, ib.STATUS_SUCCESS
    # SYMM / HEMM: m == 0 or n == 0 or (alpha == 0 and beta == 1) -> SUCCESS first
    assert ib.iclblasSsymm(None, 0, 0, 0, -3, None, None, 0, None, 0, None, None, 0) == OK
    assert ib.iclblasSsymm(bad_handle, 0, 0, 2, 2, 0.0, None, 0, None, 0, 1.0, None, 0) == OK
    assert ib.iclblasChemm(bad_handle, 1, 1, 2, 2, 0j, None, 0, None, 0, 1 + 0j, None, 0) == OK
    assert ib.iclblasCsymm(None, 0, 0, 2, 2, 0j, None, 0, None, 0, 1 + 1j, z, 4) == INV  # beta != 1: work
    assert ib.iclblasSsymm(None, 0, 0, -1, 2, 1.0, f, 4, f, 4, 0.0, f, 4) == INV
    assert ib.iclblasSsymm(None, 0, 0, 2, 2, None, f, 4, f, 4, 0.0, f, 4) == INV
    assert ib.iclblasSsymm(None, 1, 0, 2, 5, 1.0, f, 4, f, 4, 0.0, f, 4) == INV  # lda < n (right side)
    assert ib.iclblasSsymm(None, 0, 2, 2, 2, 1.0, f, 4, f, 4, 0.0, f, 4) == INV  # uplo out of range
    assert ib.iclblasSsymm(bad_handle, 0, 0, 2, 2, 1.0, f, 4, f, 4, 0.0, f, 4) == INV
    # SYRK / SYR2K (real and complex symmetric): n == 0 or k == 0 -> SUCCESS first
    assert ib.iclblasSsyrk(None, 0, 0, 0, -1, None, None, 0, None, None, 0) == OK
    assert ib.iclblasSsyrk(bad_handle, 0, 0, 3, 0, None, None, 0, None, None, 0) == OK
    assert ib.iclblasCsyr2k(bad_handle, 0, 0, 2, 2, 0j, None, 0, None, 0, 1 + 0j, None, 0) == OK
    assert ib.iclblasSsyrk(None, 0, 0, -2, 2, 1.0, f, 4, 0.0, f, 4) == INV
    assert ib.iclblasSsyr2k(None, 0, 0, 2, 2, 1.0, f, 4, None, 4, 0.0, f, 4) == INV
    assert ib.iclblasSsyrk(None, 0, 1, 2, 5, 1.0, f, 4, 0.0, f, 4) == INV  # op T: lda < k
    assert ib.iclblasCsyrk(None, 0, 3, 2, 2, 1, z, 4, 0, z, 4) == INV  # trans out of range
    assert ib.iclblasSsyrk(bad_handle, 1, 0, 2, 2, 1.0, f, 4, 0.0, f, 4) == INV
    # CHERK / CHER2K: n < 0 or k < 0 -> INVALID_VALUE first; k == 0 is work
    assert ib.iclblasCherk(None, 0, 0, 2, -1, None, None, 0, None, None, 0) == INV
    assert ib.iclblasCher2k(None, 0, 0, -1, 2, None, None, 0, None, 0, None, None, 0) == INV
    assert ib.iclblasCherk(bad_handle, 0, 0, 0, 2, None, None, 0, None, None, 0) == OK
    assert ib.iclblasCherk(bad_handle, 0, 0, 2, 0, 1.0, z, 4, 1.0, z, 4) == INV
    assert ib.iclblasCher2k(None, 0, 2, 2, 2, 1, z, 4, z, 4, None, z, 4) == INV
    # STRMM: the generic chain only (n checked before m)
    assert ib.iclblasStrmm(None, 0, 0, 0, 0, -1, 0, None, None, 0, None, 0, None, 0) == OK
    assert ib.iclblasStrmm(None, 0, 0, 0, 0, 0, -1, None, None, 0, None, 0, None, 0) == INV
    assert ib.iclblasStrmm(None, 0, 0, 0, 0, 2, 2, 1.0, f, 4, f, 4, None, 4) == INV
    assert ib.iclblasStrmm(None, 0, 0, 0, 2, 2, 2, 1.0, f, 4, f, 4, f, 4) == INV  # diag out of range
    assert ib.iclblasStrmm(bad_handle, 0, 0, 0, 0, 2, 2, 1.0, f, 4, f, 4, f, 4) == INV
    # CTRMM: m <= 0, n <= 0 or uplo outside {0, 1} -> SUCCESS first (iclBLAS_C.cpp:568-569)
    assert ib.iclblasCtrmm(None, 0, 0, 0, 0, -2, 2, None, None, 0, None, 0, None, 0) == OK
    assert ib.iclblasCtrmm(bad_handle, 0, 5, 0, 0, 2, 2, 1, z, 4, z, 4, z, 4) == OK
    assert ib.iclblasCtrmm(None, 0, 1, 0, 0, 2, 2, 1, z, 1, z, 4, z, 4) == INV  # lda < m
    assert ib.iclblasCtrmm(bad_handle, 1, 1, 2, 1, 2, 2, 1, z, 4, z, 4, z, 4) == INV


def test_gemv_status_order(forged):
    """iclBLAS_S.cpp:477-479 / iclBLAS_C.cpp:238-242, then iclBLASImpl.hpp:129-166
    (x, incx, y, incy, A, lda), then the handle."""
    bad_handle, _keep = forged
    f = np.zeros((4, 4), np.float32, order="F")
    v = np.zeros(8, np.float32)
    z = np.zeros((4, 4), np.complex64, order="F")
    w = np.zeros(8, np.complex64)
    INV, OK, NS = ib.STATUS_INVALID_VALUE, ib.STATUS_SUCCESS, ib.STATUS_NOT_SUPPORTED
    assert ib.iclblasSgemv(None, 0, 0, -1, None, None, 0, None, 0, None, None, 0) == OK
    assert ib.iclblasSgemv(bad_handle, 0, 2, 2, 0.0, None, 0, None, 0, 1.0, None, 0) == OK
    assert ib.iclblasCgemv(bad_handle, 2, 2, 2, 0j, None, 0, None, 0, 1 + 0j, None, 0) == OK
    assert ib.iclblasSgemv(None, 0, -1, 2, 1.0, f, 4, v, 1, 0.0, v, 1) == INV
    assert ib.iclblasSgemv(None, 0, 2, 2, 1.0, f, 4, v, 0, 0.0, v, 1) == INV
    assert ib.iclblasSgemv(None, 0, 2, 2, 1.0, f, 4, None, 1, 0.0, v, 1) == INV
    assert ib.iclblasSgemv(None, 0, 2, 2, 1.0, f, 4, v, -1, 0.0, v, 1) == NS
    assert ib.iclblasSgemv(None, 0, 2, 2, 1.0, f, 4, v, 1, 0.0, None, 1) == INV
    assert ib.iclblasCgemv(None, 0, 2, 2, 1, z, 4, w, 1, 0, w, -2) == NS
    assert ib.iclblasCgemv(None, 0, 2, 2, 1, None, 4, w, 1, 0, w, 1) == INV
    assert ib.iclblasSgemv(None, 0, 5, 2, 1.0, f, 4, v, 1, 0.0, v, 1) == INV  # lda < m
    assert ib.iclblasSgemv(None, 3, 2, 2, 1.0, f, 4, v, 1, 0.0, v, 1) == INV  # trans out of range
    assert ib.iclblasSgemv(None, 0, 2, 2, None, f, 4, v, 1, 0.0, v, 1) == INV
    assert ib.iclblasCgemv(bad_handle, 2, 2, 2, 1, z, 4, w, 1, 0, w, 1) == INV


def test_handle_entry_points_reject_invalid_handles(forged):
    bad_handle, _keep = forged
    L = ib.load()
    assert L.iclblasDestroy(None) == ib.STATUS_INVALID_VALUE
    assert L.iclblasDestroy(bad_handle) == ib.STATUS_INVALID_VALUE
    assert L.iclblasSetMemoryMode(bad_handle, 0) == ib.STATUS_INVALID_VALUE
    assert L.iclblasSetMathMode(None, 1) == ib.STATUS_INVALID_VALUE
    assert L.iclblasSynchronize(None) == ib.STATUS_INVALID_VALUE
    assert L.iclblasSetImplementation(None, b"x") == ib.STATUS_INVALID_VALUE
    assert L.iclblasCreate(None) == ib.STATUS_INVALID_VALUE


def test_unimplemented_entry_points_report_not_supported():
    L = ib.load()
    f = L.iclblasSaxpy
    f.restype = ct.c_int
    assert f(None, 0, None, None, 1, None, 1) == ib.STATUS_NOT_SUPPORTED
    g = L.iclblasSrotg
    g.restype = ct.c_int
    assert g(None, None, None, None, None) == ib.STATUS_NOT_SUPPORTED


def test_create_without_a_gpu_reports_not_initialized():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    st, h = ib.iclblasCreate()
    assert st == ib.STATUS_NOT_INITIALIZED and not h.value


def test_generator_is_deterministic_and_exact_in_fp32():
    a = datagen.uniform_pm1(10000, seed=0, stream=0)
    b = datagen.uniform_pm1(10000, seed=0, stream=0)
    assert np.array_equal(a, b)
    assert np.all(np.abs(a) < 1.0)
    assert np.array_equal(a.astype(np.float32).astype(np.float64), a)
    assert not np.array_equal(a, datagen.uniform_pm1(10000, seed=0, stream=1))
    assert abs(a.mean()) < 0.03 and abs(a.var() - 1 / 3) < 0.02
    w = datagen.normal(20000, seed=0, stream=5)
    assert abs(w.mean()) < 0.03 and abs(w.std() - 1.0) < 0.03
    T = datagen.triangular(64, lower=True)
    assert np.all(np.isnan(T[np.triu_indices(64, 1)]))
    d = np.abs(np.diag(T))
    assert np.all((d >= 1.0) & (d <= 2.0))
    q = datagen.quarter_integers(1000)
    assert set(np.unique(q)) <= {x / 4 for x in range(-4, 5)}
