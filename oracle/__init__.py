"""CPU oracle for the level-3 hot path -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, by __graft_entry__.smoke() (as the checker) and by
bench.py's CPU-baseline leg. The library (clgpu_amd) never imports it.

liboracle.so (oracle.c) restates the reference's CPU semantics:
  sgemm/cgemm  <- cpu_gemm, reference tests/blas_funcs/test_gemm.hpp:83-201
  strsm        <- Strsm_naive, reference libs/blas/funcs/Strsm/Strsm_naive.cl:18-126
  ctrsm        <- Ctrsm_naive, reference libs/blas/funcs/Ctrsm/Ctrsm_naive.cl:16-183
and provides FP64 references (dgemm, zgemm, dtrsm, ztrsm) for error reporting.
level3.c adds FP64 oracles of the GEMM-based level-3 routines (symm / hemm,
syrk / herk, syr2k / her2k, trmm) and level2.c one of gemv (the sources they
restate are listed there).
_ref/libref_gemm.so, when present, is the reference's own cpu_gemm compiled from
its headers (oracle/Makefile, target `ref`).

All matrices are numpy arrays in Fortran (column-major) order whose shape[0]
is the leading dimension.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
# ICLBLAS_ORACLE_LIBRARY selects another build of the oracle (oracle/_san, the
# sanitizer build: `make -C oracle san`).
LIB = Path(os.environ.get("ICLBLAS_ORACLE_LIBRARY") or HERE / "liboracle.so")
REF_LIB = HERE / "_ref" / "libref_gemm.so"

_f32p = np.ctypeslib.ndpointer(np.float32, flags="F_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_lib = None
_ref = None


def build(reference: str | None = None) -> None:
    """Compile liboracle.so (and oracle/_ref when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "all"], check=True)
    ref = reference or "/root/reference"
    if os.path.isdir(os.path.join(ref, "tests", "blas_funcs")):
        subprocess.run(["make", "-s", "-C", str(HERE), "ref", f"REFERENCE={ref}"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ct.CDLL(str(LIB))
        I, F, D, P = ct.c_int, ct.c_float, ct.c_double, ct.c_void_p
        L.orc_sgemm.argtypes = [I, I, I, I, I, F, P, I, P, I, F, P, I]
        L.orc_sgemm_mt.argtypes = [I, I, I, I, I, I, F, P, I, P, I, F, P, I]
        L.orc_sgemm_mt.restype = I
        L.orc_cgemm.argtypes = [I, I, I, I, I, P, P, I, P, I, P, P, I]
        L.orc_strsm.argtypes = [I, I, I, I, I, I, F, P, I, P, I]
        L.orc_dtrsm.argtypes = [I, I, I, I, I, I, D, P, I, P, I]
        L.orc_ctrsm.argtypes = [I, I, I, I, I, I, P, P, I, P, I]
        L.orc_ztrsm.argtypes = [I, I, I, I, I, I, P, P, I, P, I]
        L.orc_dgemm.argtypes = [I, I, I, I, I, D, P, I, P, I, D, P, I, P, P]
        L.orc_dgemm_sample.argtypes = [I, I, I, D, P, I, P, I, D, P, I, I, P, P, P, P]
        L.orc_zgemm.argtypes = [I, I, I, I, I, P, P, I, P, I, P, P, I, P, P]
        L.orc_zgemm_sample.argtypes = [I, I, I, P, P, I, P, I, P, P, I, I, P, P, P, P]
        L.orc_zsymm.argtypes = [I, I, I, I, I, P, P, I, P, I, P, P, I, P]
        L.orc_zsyrk.argtypes = [I, I, I, I, I, P, P, I, P, P, I, P]
        L.orc_zsyr2k.argtypes = [I, I, I, I, I, P, P, I, P, I, P, P, I, P]
        L.orc_ztrmm.argtypes = [I, I, I, I, I, I, P, P, I, P, I, P, I, P]
        L.orc_zgemv.argtypes = [I, I, I, P, P, I, P, I, P, P, I, P, P]
        _lib = L
    return _lib


def ref_available() -> bool:
    return REF_LIB.exists()


def ref_lib():
    global _ref
    if _ref is None:
        R = ct.CDLL(str(REF_LIB))
        I, F, P = ct.c_int, ct.c_float, ct.c_void_p
        R.ref_cpu_sgemm.argtypes = [I, I, I, I, I, F, P, I, P, I, F, P, I]
        R.ref_cpu_cgemm.argtypes = [I, I, I, I, I, P, P, I, P, I, P, P, I]
        _ref = R
    return _ref


def _p(a: np.ndarray):
    return a.ctypes.data_as(ct.c_void_p)


def _chk(a: np.ndarray, dtype) -> np.ndarray:
    assert a.dtype == dtype and a.flags.f_contiguous, (a.dtype, a.flags)
    return a


def _ld(a: np.ndarray) -> int:
    return max(1, a.shape[0])


def op_shape(trans: int, rows: int, cols: int):
    return (rows, cols) if trans == 0 else (cols, rows)


# ---------------------------------------------------------------- layer 1
def sgemm(transa, transb, m, n, k, alpha, A, B, beta, C, reference=False):
    """C <- cpu_gemm(...) in place (C float32 Fortran array)."""
    _chk(A, np.float32), _chk(B, np.float32), _chk(C, np.float32)
    fn = ref_lib().ref_cpu_sgemm if reference else lib().orc_sgemm
    fn(transa, transb, m, n, k, alpha, _p(A), _ld(A), _p(B), _ld(B), beta, _p(C), _ld(C))
    return C


def sgemm_mt(threads, transa, transb, m, n, k, alpha, A, B, beta, C):
    _chk(A, np.float32), _chk(B, np.float32), _chk(C, np.float32)
    return lib().orc_sgemm_mt(threads, transa, transb, m, n, k, alpha, _p(A), _ld(A), _p(B), _ld(B), beta, _p(C),
                              _ld(C))


def cgemm(transa, transb, m, n, k, alpha, A, B, beta, C, reference=False):
    _chk(A, np.complex64), _chk(B, np.complex64), _chk(C, np.complex64)
    al = np.array([alpha.real, alpha.imag], np.float32)
    be = np.array([beta.real, beta.imag], np.float32)
    fn = ref_lib().ref_cpu_cgemm if reference else lib().orc_cgemm
    fn(transa, transb, m, n, k, _p(al), _p(A), _ld(A), _p(B), _ld(B), _p(be), _p(C), _ld(C))
    return C


def strsm(side, uplo, trans, diag, m, n, alpha, A, B):
    _chk(A, np.float32), _chk(B, np.float32)
    lib().orc_strsm(side, uplo, trans, diag, m, n, alpha, _p(A), _ld(A), _p(B), _ld(B))
    return B


def ctrsm(side, uplo, trans, diag, m, n, alpha, A, B):
    """Ctrsm_naive restated (complex64 in place on B)."""
    _chk(A, np.complex64), _chk(B, np.complex64)
    al = np.array([complex(alpha).real, complex(alpha).imag], np.float32)
    lib().orc_ctrsm(side, uplo, trans, diag, m, n, _p(al), _p(A), _ld(A), _p(B), _ld(B))
    return B


# ---------------------------------------------------------------- layer 2
def dgemm(transa, transb, m, n, k, alpha, A, B, beta=0.0, C=None):
    """Returns (C64, absAB) as dense m x n float64 (C order irrelevant: Fortran view)."""
    _chk(A, np.float32), _chk(B, np.float32)
    C64 = np.zeros(m * n, np.float64)
    ab = np.zeros(m * n, np.float64)
    Cp, ldc = (_p(C), _ld(C)) if C is not None else (None, 1)
    lib().orc_dgemm(transa, transb, m, n, k, float(alpha), _p(A), _ld(A), _p(B), _ld(B), float(beta), Cp, ldc,
                    _p(C64), _p(ab))
    return C64.reshape(n, m).T, ab.reshape(n, m).T


def dgemm_sample(transa, transb, k, alpha, A, B, rows, cols, beta=0.0, C=None):
    _chk(A, np.float32), _chk(B, np.float32)
    rows = np.ascontiguousarray(rows, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    out = np.zeros(len(rows), np.float64)
    ab = np.zeros(len(rows), np.float64)
    Cp, ldc = (_p(C), _ld(C)) if C is not None else (None, 1)
    lib().orc_dgemm_sample(transa, transb, k, float(alpha), _p(A), _ld(A), _p(B), _ld(B), float(beta), Cp, ldc,
                           len(rows), _p(rows), _p(cols), _p(out), _p(ab))
    return out, ab


def zgemm(transa, transb, m, n, k, alpha, A, B, beta=0.0, C=None):
    _chk(A, np.complex64), _chk(B, np.complex64)
    al = np.array([complex(alpha).real, complex(alpha).imag], np.float64)
    be = np.array([complex(beta).real, complex(beta).imag], np.float64)
    out = np.zeros(2 * m * n, np.float64)
    ab = np.zeros(m * n, np.float64)
    Cp, ldc = (_p(C), _ld(C)) if C is not None else (None, 1)
    lib().orc_zgemm(transa, transb, m, n, k, _p(al), _p(A), _ld(A), _p(B), _ld(B), _p(be), Cp, ldc, _p(out), _p(ab))
    z = out[0::2] + 1j * out[1::2]
    return z.reshape(n, m).T, ab.reshape(n, m).T


def zgemm_sample(transa, transb, k, alpha, A, B, rows, cols, beta=0.0, C=None):
    _chk(A, np.complex64), _chk(B, np.complex64)
    rows = np.ascontiguousarray(rows, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    al = np.array([complex(alpha).real, complex(alpha).imag], np.float64)
    be = np.array([complex(beta).real, complex(beta).imag], np.float64)
    out = np.zeros(2 * len(rows), np.float64)
    ab = np.zeros(len(rows), np.float64)
    Cp, ldc = (_p(C), _ld(C)) if C is not None else (None, 1)
    lib().orc_zgemm_sample(transa, transb, k, _p(al), _p(A), _ld(A), _p(B), _ld(B), _p(be), Cp, ldc, len(rows),
                           _p(rows), _p(cols), _p(out), _p(ab))
    return out[0::2] + 1j * out[1::2], ab


def dtrsm(side, uplo, trans, diag, m, n, alpha, A, B):
    """FP64 substitution; B (float32) is read, a float64 m x n result returned."""
    _chk(A, np.float32)
    X = np.asfortranarray(np.array(B[:m, :n], dtype=np.float64))
    lib().orc_dtrsm(side, uplo, trans, diag, m, n, float(alpha), _p(A), _ld(A), _p(X), _ld(X))
    return X


def ztrsm(side, uplo, trans, diag, m, n, alpha, A, B):
    """Complex FP64 substitution; B (complex64) is read, a complex128 m x n result returned."""
    _chk(A, np.complex64)
    X = np.asfortranarray(np.array(B[:m, :n], dtype=np.complex128))
    al = np.array([complex(alpha).real, complex(alpha).imag], np.float64)
    lib().orc_ztrsm(side, uplo, trans, diag, m, n, _p(al), _p(A), _ld(A), _p(X), _ld(X))
    return X


# ---------------------------------------------------------------- level 3 (FP64)
def _z(X) -> np.ndarray:
    """complex128 Fortran copy of a float32 / complex64 matrix."""
    return np.asfortranarray(np.asarray(X, np.complex128))


def _zs(v) -> np.ndarray:
    v = complex(v)
    return np.array([v.real, v.imag], np.float64)


def _ret(Z, mag, m, n, real):
    R = Z[:m, :n]
    return (np.ascontiguousarray(R.real) if real else R), mag.reshape(n, m).T


def zsymm(side, uplo, herm, m, n, alpha, A, B, beta, C):
    """FP64 SYMM / HEMM. Returns (C result m x n, mag m x n); C is only read."""
    real = not np.iscomplexobj(A)
    Az, Bz, Cz = _z(A), _z(B), _z(C)
    mag = np.zeros(m * n, np.float64)
    lib().orc_zsymm(side, uplo, int(herm), m, n, _p(_zs(alpha)), _p(Az), _ld(Az), _p(Bz), _ld(Bz), _p(_zs(beta)),
                    _p(Cz), _ld(Cz), _p(mag))
    return _ret(Cz, mag, m, n, real)


def zsyrk(uplo, trans, herm, n, k, alpha, A, beta, C):
    """FP64 SYRK / HERK on the uplo triangle of C (the rest of C is returned as
    given). Returns (C n x n, mag n x n)."""
    real = not np.iscomplexobj(A)
    Az, Cz = _z(A), _z(C)
    mag = np.zeros(n * n, np.float64)
    lib().orc_zsyrk(uplo, trans, int(herm), n, k, _p(_zs(alpha)), _p(Az), _ld(Az), _p(_zs(beta)), _p(Cz), _ld(Cz),
                    _p(mag))
    return _ret(Cz, mag, n, n, real)


def zsyr2k(uplo, trans, herm, n, k, alpha, A, B, beta, C):
    """FP64 SYR2K / HER2K on the uplo triangle of C. Returns (C n x n, mag n x n)."""
    real = not np.iscomplexobj(A)
    Az, Bz, Cz = _z(A), _z(B), _z(C)
    mag = np.zeros(n * n, np.float64)
    lib().orc_zsyr2k(uplo, trans, int(herm), n, k, _p(_zs(alpha)), _p(Az), _ld(Az), _p(Bz), _ld(Bz), _p(_zs(beta)),
                     _p(Cz), _ld(Cz), _p(mag))
    return _ret(Cz, mag, n, n, real)


def ztrmm(side, uplo, trans, diag, m, n, alpha, A, B):
    """FP64 TRMM (header contract). Returns (C m x n, mag m x n)."""
    real = not np.iscomplexobj(A)
    Az, Bz = _z(A), _z(B)
    Cz = np.asfortranarray(np.zeros((m, n), np.complex128))
    mag = np.zeros(m * n, np.float64)
    lib().orc_ztrmm(side, uplo, trans, diag, m, n, _p(_zs(alpha)), _p(Az), _ld(Az), _p(Bz), _ld(Bz), _p(Cz),
                    _ld(Cz), _p(mag))
    return _ret(Cz, mag, m, n, real)


def zgemv(trans, m, n, alpha, A, x, incx, beta, y, incy):
    """FP64 GEMV (level2.c). A is a float32 / complex64 Fortran matrix, x and y
    1-D arrays read with strides incx / incy. Returns (y result of length len(y)
    as float64 / complex128, mag)."""
    real = not np.iscomplexobj(A)
    Az = _z(A)
    xz = np.ascontiguousarray(np.asarray(x, np.complex128))
    yz = np.ascontiguousarray(np.asarray(y, np.complex128))
    leny = m if trans == 0 else n
    out = np.zeros(leny, np.complex128)
    mag = np.zeros(leny, np.float64)
    lib().orc_zgemv(trans, m, n, _p(_zs(alpha)), _p(Az), _ld(Az), _p(xz), incx, _p(_zs(beta)), _p(yz), incy, _p(out),
                    _p(mag))
    return (out.real.copy() if real else out), mag


# ---------------------------------------------------------------- error metrics
U32 = 2.0 ** -24


def gemm_errors(Cg, C64, absAB, k, beta_absC=None):
    """Max |err|, max relative error and the max ratio to the elementwise bound
    4*k*2^-24*(|A||B|)_ij (reported as err / (k*2^-24*(|A||B|)_ij), must be <= 4).
    beta_absC adds |beta||C|*2^-24 to the bound denominator for beta != 0."""
    err = np.abs(np.asarray(Cg, np.complex128 if np.iscomplexobj(Cg) else np.float64) - C64)
    den = k * U32 * np.asarray(absAB, np.float64)
    if beta_absC is not None:
        den = den + U32 * beta_absC
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = np.where(den > 0, err / den, np.where(err == 0, 0.0, np.inf))
        rel = np.where(np.abs(C64) > 0, err / np.abs(C64), np.where(err == 0, 0.0, np.inf))
    return {"max_abs_err": float(err.max(initial=0.0)), "max_rel_err": float(rel.max(initial=0.0)),
            "max_bound_ratio": float(ratio.max(initial=0.0))}
