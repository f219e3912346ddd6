"""ctypes binding of libiclBLAS.so, mirroring the reference C API.

Names, enum values and argument order are those of the reference's
include/iclBLAS/iclBLAS.h (iclblasCreate :172, iclblasDestroy :179,
iclblasSgemm :1685, iclblasStrsm :1719, iclblasCgemm :1799, iclblasCtrsm
:1925; the level-3 routines Ssymm :1579, Ssyrk :1608, Ssyr2k :1645, Strmm
:1753, Csymm :1825, Csyr2k :1862, Csyrk :1891, Cherk :1954, Cher2k :1991,
Ctrmm :2025, Chemm :2051; Sgemv :1049, Cgemv :1164) plus the extensions of
include/iclBLAS_ext.h. Every function returns the C status;
`check()` turns a non-success status into IclblasError.

Host-pointer mode (the default, as in the reference) takes numpy arrays in
Fortran order whose shape[0] is the leading dimension. Device-pointer mode
takes integer device addresses (e.g. torch tensor .data_ptr()).

The shared library must have been built (python -m clgpu_amd.build); loading
fails loudly otherwise -- there is no fallback path.
"""
from __future__ import annotations

import ctypes as ct
import os
from pathlib import Path

import numpy as np

# ICLBLAS_LIBRARY selects another build of the same library (the sanitizer build
# of the host layer, clgpu_amd/_build/san/libiclBLAS.so, runs the CPU tests).
LIB_PATH = Path(os.environ.get("ICLBLAS_LIBRARY") or Path(__file__).resolve().parent / "libiclBLAS.so")

# iclblasStatus_t (iclBLAS.h:104-117)
STATUS_SUCCESS = 0
STATUS_NOT_INITIALIZED = 1
STATUS_ALLOC_FAILED = 3
STATUS_INVALID_VALUE = 7
STATUS_ARCH_MISMATCH = 8
STATUS_MAPPING_ERROR = 11
STATUS_EXECUTION_FAILED = 13
STATUS_INTERNAL_ERROR = 14
STATUS_NOT_SUPPORTED = 15
STATUS_LICENSE_ERROR = 16
STATUS_ERROR = 17
STATUS_NAMES = {v: k for k, v in dict(globals()).items() if k.startswith("STATUS_") and isinstance(v, int)}

OP_N, OP_T, OP_C = 0, 1, 2
FILL_MODE_UPPER, FILL_MODE_LOWER = 0, 1
DIAG_NON_UNIT, DIAG_UNIT = 0, 1
SIDE_LEFT, SIDE_RIGHT = 0, 1
MEMORY_HOST, MEMORY_DEVICE = 0, 1
MATH_FP32, MATH_SPLIT_BF16X3 = 0, 1


class IclblasError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        super().__init__(f"{what}: {STATUS_NAMES.get(status, status)} ({status})")
        self.status = status


def check(status: int, what: str = "iclblas") -> None:
    if status != STATUS_SUCCESS:
        raise IclblasError(status, what)


class _Complex(ct.Structure):
    _fields_ = [("re", ct.c_float), ("im", ct.c_float)]


_lib = None

_I, _F, _P = ct.c_int, ct.c_float, ct.c_void_p
_FP = ct.POINTER(ct.c_float)
_CP = ct.POINTER(_Complex)
_SIGS = {
    "iclblasCreate": [ct.POINTER(_P)],
    "iclblasDestroy": [_P],
    "iclblasSgemm": [_P, _I, _I, _I, _I, _I, _FP, _P, _I, _P, _I, _FP, _P, _I],
    "iclblasCgemm": [_P, _I, _I, _I, _I, _I, _CP, _P, _I, _P, _I, _CP, _P, _I],
    "iclblasStrsm": [_P, _I, _I, _I, _I, _I, _I, _FP, _P, _I, _P, _I],
    "iclblasCtrsm": [_P, _I, _I, _I, _I, _I, _I, _CP, _P, _I, _P, _I],
    "iclblasStrsmRefined": [_P, _I, _I, _I, _I, _I, _I, _FP, _P, _I, _P, _I, _I],
    "iclblasSgemv": [_P, _I, _I, _I, _FP, _P, _I, _P, _I, _FP, _P, _I],
    "iclblasCgemv": [_P, _I, _I, _I, _CP, _P, _I, _P, _I, _CP, _P, _I],
    "iclblasSsymm": [_P, _I, _I, _I, _I, _FP, _P, _I, _P, _I, _FP, _P, _I],
    "iclblasCsymm": [_P, _I, _I, _I, _I, _CP, _P, _I, _P, _I, _CP, _P, _I],
    "iclblasChemm": [_P, _I, _I, _I, _I, _CP, _P, _I, _P, _I, _CP, _P, _I],
    "iclblasSsyrk": [_P, _I, _I, _I, _I, _FP, _P, _I, _FP, _P, _I],
    "iclblasCsyrk": [_P, _I, _I, _I, _I, _CP, _P, _I, _CP, _P, _I],
    "iclblasCherk": [_P, _I, _I, _I, _I, _FP, _P, _I, _FP, _P, _I],
    "iclblasSsyr2k": [_P, _I, _I, _I, _I, _FP, _P, _I, _P, _I, _FP, _P, _I],
    "iclblasCsyr2k": [_P, _I, _I, _I, _I, _CP, _P, _I, _P, _I, _CP, _P, _I],
    "iclblasCher2k": [_P, _I, _I, _I, _I, _CP, _P, _I, _P, _I, _FP, _P, _I],
    "iclblasStrmm": [_P, _I, _I, _I, _I, _I, _I, _FP, _P, _I, _P, _I, _P, _I],
    "iclblasCtrmm": [_P, _I, _I, _I, _I, _I, _I, _CP, _P, _I, _P, _I, _P, _I],
    "iclblasSgemmBackward": [_P, _I, _I, _I, _I, _I, _FP, _P, _I, _P, _I, _P, _I, _P, _I, _P, _I],
    "iclblasStrsmBackward": [_P, _I, _I, _I, _I, _I, _I, _FP, _P, _I, _P, _I, _P, _I, _P, _I, _P, _I],
    "iclblasSetMemoryMode": [_P, _I],
    "iclblasGetMemoryMode": [_P, ct.POINTER(_I)],
    "iclblasSetMathMode": [_P, _I],
    "iclblasGetMathMode": [_P, ct.POINTER(_I)],
    "iclblasSetStream": [_P, _P],
    "iclblasGetStream": [_P, ct.POINTER(_P)],
    "iclblasSynchronize": [_P],
    "iclblasSetImplementation": [_P, ct.c_char_p],
    "iclblasGetLastImplementation": [_P, ct.c_char_p, _I],
    "iclblasGetLastRefinementSweeps": [_P, ct.POINTER(_I)],
    "iclblasGetImplementations": [ct.c_char_p, _I],
    "iclblasSetKernelTiming": [_P, _I],
    "iclblasGetKernelTiming": [_P, _FP, ct.POINTER(_I)],
    "iclblasSgemmNoAlloc": [_P, _I, _I, _I, _I, _I, _FP, _P, _I, _P, _I, _FP, _P, _I],
    "iclblasSetScratchLimit": [_P, ct.c_size_t],
    "iclblasGetScratchUsage": [_P, ct.POINTER(ct.c_size_t), ct.POINTER(ct.c_size_t)],
}

# Bytes of the library's static device scratch (ICLBLAS_NOALLOC_SCRATCH_BYTES, iclBLAS_ext.h)
NOALLOC_SCRATCH_BYTES = 64 << 20


def implementation_names() -> list[str]:
    """Names of every registered GEMM implementation (iclblasGetImplementations)."""
    buf = ct.create_string_buffer(1 << 14)
    check(load().iclblasGetImplementations(buf, len(buf)), "iclblasGetImplementations")
    return buf.value.decode().split("\n") if buf.value else []


def load() -> ct.CDLL:
    """Loads libiclBLAS.so from the package directory (raises if it is missing)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is not built; run `python -m clgpu_amd.build`")
        # PyTorch-ROCm ships its own libamdhip64.so.7 with the same SONAME as the
        # system ROCm runtime the library links. Whichever is loaded first serves the
        # whole process; loading torch first keeps one HIP runtime that both torch and
        # the library use (torch cannot initialise the GPU on a runtime it was not
        # built with).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ct.CDLL(str(LIB_PATH))
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ct.c_int
        _lib = L
    return _lib


def _ptr(x):
    """numpy array -> its data pointer; int -> device address; None -> NULL."""
    if x is None:
        return None
    if isinstance(x, np.ndarray):
        return x.ctypes.data_as(ct.c_void_p)
    if isinstance(x, int):
        return ct.c_void_p(x)
    if hasattr(x, "data_ptr"):
        return ct.c_void_p(x.data_ptr())
    raise TypeError(type(x))


def _fscalar(v):
    return None if v is None else ct.byref(ct.c_float(v))


def _cscalar(v):
    if v is None:
        return None
    v = complex(v)
    return ct.byref(_Complex(v.real, v.imag))


# ----------------------------------------------------------------------------
# C API, argument for argument
# ----------------------------------------------------------------------------
def iclblasCreate():
    """Returns (status, handle)."""
    h = _P()
    st = load().iclblasCreate(ct.byref(h))
    return st, h


def iclblasDestroy(handle):
    return load().iclblasDestroy(handle)


def iclblasSgemm(handle, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    return load().iclblasSgemm(handle, transa, transb, m, n, k, _fscalar(alpha), _ptr(A), lda, _ptr(B), ldb,
                               _fscalar(beta), _ptr(C), ldc)


def iclblasSgemmNoAlloc(handle, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    """SGEMM on device pointers that allocates nothing (iclBLAS_ext.h)."""
    return load().iclblasSgemmNoAlloc(handle, transa, transb, m, n, k, _fscalar(alpha), _ptr(A), lda, _ptr(B), ldb,
                                      _fscalar(beta), _ptr(C), ldc)


def iclblasCgemm(handle, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    return load().iclblasCgemm(handle, transa, transb, m, n, k, _cscalar(alpha), _ptr(A), lda, _ptr(B), ldb,
                               _cscalar(beta), _ptr(C), ldc)


def iclblasStrsm(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb):
    return load().iclblasStrsm(handle, side, uplo, trans, diag, m, n, _fscalar(alpha), _ptr(A), lda, _ptr(B), ldb)


def iclblasCtrsm(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb):
    return load().iclblasCtrsm(handle, side, uplo, trans, diag, m, n, _cscalar(alpha), _ptr(A), lda, _ptr(B), ldb)


def iclblasStrsmRefined(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, max_sweeps):
    return load().iclblasStrsmRefined(handle, side, uplo, trans, diag, m, n, _fscalar(alpha), _ptr(A), lda, _ptr(B),
                                      ldb, max_sweeps)


def iclblasSgemv(handle, trans, m, n, alpha, A, lda, x, incx, beta, y, incy):
    return load().iclblasSgemv(handle, trans, m, n, _fscalar(alpha), _ptr(A), lda, _ptr(x), incx, _fscalar(beta),
                               _ptr(y), incy)


def iclblasCgemv(handle, trans, m, n, alpha, A, lda, x, incx, beta, y, incy):
    return load().iclblasCgemv(handle, trans, m, n, _cscalar(alpha), _ptr(A), lda, _ptr(x), incx, _cscalar(beta),
                               _ptr(y), incy)


def iclblasSsymm(handle, side, uplo, m, n, alpha, A, lda, B, ldb, beta, C, ldc):
    return load().iclblasSsymm(handle, side, uplo, m, n, _fscalar(alpha), _ptr(A), lda, _ptr(B), ldb, _fscalar(beta),
                               _ptr(C), ldc)


def iclblasCsymm(handle, side, uplo, m, n, alpha, A, lda, B, ldb, beta, C, ldc):
    return load().iclblasCsymm(handle, side, uplo, m, n, _cscalar(alpha), _ptr(A), lda, _ptr(B), ldb, _cscalar(beta),
                               _ptr(C), ldc)


def iclblasChemm(handle, side, uplo, m, n, alpha, A, lda, B, ldb, beta, C, ldc):
    return load().iclblasChemm(handle, side, uplo, m, n, _cscalar(alpha), _ptr(A), lda, _ptr(B), ldb, _cscalar(beta),
                               _ptr(C), ldc)


def iclblasSsyrk(handle, uplo, trans, n, k, alpha, A, lda, beta, C, ldc):
    return load().iclblasSsyrk(handle, uplo, trans, n, k, _fscalar(alpha), _ptr(A), lda, _fscalar(beta), _ptr(C), ldc)


def iclblasCsyrk(handle, uplo, trans, n, k, alpha, A, lda, beta, C, ldc):
    return load().iclblasCsyrk(handle, uplo, trans, n, k, _cscalar(alpha), _ptr(A), lda, _cscalar(beta), _ptr(C), ldc)


def iclblasCherk(handle, uplo, trans, n, k, alpha, A, lda, beta, C, ldc):
    """alpha and beta are real."""
    return load().iclblasCherk(handle, uplo, trans, n, k, _fscalar(alpha), _ptr(A), lda, _fscalar(beta), _ptr(C), ldc)


def iclblasSsyr2k(handle, uplo, trans, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    return load().iclblasSsyr2k(handle, uplo, trans, n, k, _fscalar(alpha), _ptr(A), lda, _ptr(B), ldb, _fscalar(beta),
                                _ptr(C), ldc)


def iclblasCsyr2k(handle, uplo, trans, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    return load().iclblasCsyr2k(handle, uplo, trans, n, k, _cscalar(alpha), _ptr(A), lda, _ptr(B), ldb, _cscalar(beta),
                                _ptr(C), ldc)


def iclblasCher2k(handle, uplo, trans, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    """alpha is complex, beta real."""
    return load().iclblasCher2k(handle, uplo, trans, n, k, _cscalar(alpha), _ptr(A), lda, _ptr(B), ldb, _fscalar(beta),
                                _ptr(C), ldc)


def iclblasStrmm(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, C, ldc):
    return load().iclblasStrmm(handle, side, uplo, trans, diag, m, n, _fscalar(alpha), _ptr(A), lda, _ptr(B), ldb,
                               _ptr(C), ldc)


def iclblasCtrmm(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, C, ldc):
    return load().iclblasCtrmm(handle, side, uplo, trans, diag, m, n, _cscalar(alpha), _ptr(A), lda, _ptr(B), ldb,
                               _ptr(C), ldc)


def iclblasSgemmBackward(handle, transa, transb, m, n, k, alpha, A, lda, B, ldb, dC, lddc, dA, ldda, dB, lddb):
    return load().iclblasSgemmBackward(handle, transa, transb, m, n, k, _fscalar(alpha), _ptr(A), lda, _ptr(B), ldb,
                                       _ptr(dC), lddc, _ptr(dA), ldda, _ptr(dB), lddb)


def iclblasStrsmBackward(handle, side, uplo, trans, diag, m, n, alpha, A, lda, X, ldx, dX, lddx, dA, ldda, dB, lddb):
    return load().iclblasStrsmBackward(handle, side, uplo, trans, diag, m, n, _fscalar(alpha), _ptr(A), lda, _ptr(X),
                                       ldx, _ptr(dX), lddx, _ptr(dA), ldda, _ptr(dB), lddb)


# ----------------------------------------------------------------------------
# Handle convenience wrapper
# ----------------------------------------------------------------------------
class Handle:
    """RAII wrapper: `with Handle() as h: ...`."""

    def __init__(self, memory_mode: int = MEMORY_HOST, math_mode: int = MATH_FP32, stream=None):
        st, self.h = iclblasCreate()
        check(st, "iclblasCreate")
        self.set_memory_mode(memory_mode)
        self.set_math_mode(math_mode)
        if stream is not None:
            self.set_stream(stream)

    def close(self):
        if self.h:
            check(iclblasDestroy(self.h), "iclblasDestroy")
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def value(self):
        return self.h

    def set_memory_mode(self, mode: int):
        check(load().iclblasSetMemoryMode(self.h, mode), "iclblasSetMemoryMode")

    def set_math_mode(self, mode: int):
        check(load().iclblasSetMathMode(self.h, mode), "iclblasSetMathMode")

    def set_stream(self, stream):
        """stream: a hipStream_t as int (torch: torch.cuda.current_stream().cuda_stream)."""
        check(load().iclblasSetStream(self.h, _P(int(stream))), "iclblasSetStream")

    def synchronize(self):
        check(load().iclblasSynchronize(self.h), "iclblasSynchronize")

    def set_implementation(self, name: str | None):
        check(load().iclblasSetImplementation(self.h, (name or "").encode()), "iclblasSetImplementation")

    def last_implementation(self) -> str:
        buf = ct.create_string_buffer(128)
        check(load().iclblasGetLastImplementation(self.h, buf, 128), "iclblasGetLastImplementation")
        return buf.value.decode()

    def last_refinement_sweeps(self) -> int:
        v = ct.c_int()
        check(load().iclblasGetLastRefinementSweeps(self.h, ct.byref(v)), "iclblasGetLastRefinementSweeps")
        return v.value

    def set_kernel_timing(self, on: bool):
        check(load().iclblasSetKernelTiming(self.h, 1 if on else 0), "iclblasSetKernelTiming")

    def set_scratch_limit(self, nbytes: int):
        """Cap of the handle's device buffers in bytes (0 = none); beyond it calls fail with ALLOC_FAILED."""
        check(load().iclblasSetScratchLimit(self.h, nbytes), "iclblasSetScratchLimit")

    def scratch_usage(self) -> tuple[int, int]:
        """(bytes held by the handle's device buffers, cap)."""
        held, limit = ct.c_size_t(), ct.c_size_t()
        check(load().iclblasGetScratchUsage(self.h, ct.byref(held), ct.byref(limit)), "iclblasGetScratchUsage")
        return held.value, limit.value

    def kernel_timing(self) -> tuple[float, int]:
        """(summed ms, launches) of the main GEMM kernels timed since the last read."""
        ms, n = ct.c_float(), ct.c_int()
        check(load().iclblasGetKernelTiming(self.h, ct.byref(ms), ct.byref(n)), "iclblasGetKernelTiming")
        return ms.value, n.value


def exported_symbols() -> list[str]:
    """Names of the iclblas* functions the library exports (read with nm-like
    dlsym probing of the declarations in include/*.h)."""
    L = load()
    out = []
    for name in declared_symbols():
        if hasattr(L, name):
            out.append(name)
    return out


def declared_symbols() -> list[str]:
    """Every function declared in include/iclBLAS.h and include/iclBLAS_ext.h."""
    import re
    root = Path(__file__).resolve().parent.parent / "include"
    names = []
    for hdr in ("iclBLAS.h", "iclBLAS_ext.h"):
        text = (root / hdr).read_text()
        names += re.findall(r"ICLBLAS_API\s+iclblasStatus_t\s+(\w+)\s*\(", text)
    return names
