"""torch.autograd.Function wrappers over the SGEMM / STRSM adjoint entry points.

Used as the gradient test harness only: forward and backward both run the
library's HIP kernels through the C ABI (iclblasSgemm / iclblasStrsm in device
memory mode, iclblasSgemmBackward / iclblasStrsmBackward).

Matrices are column-major: a rows x cols matrix is a contiguous float32 CUDA
tensor of shape (cols, rows) (leading dimension = rows), i.e. the transpose of
torch's row-major view. `cm(t)` / `rm(t)` convert between the two views.
"""
from __future__ import annotations

import torch

from . import iclblas as ib

_handles: dict = {}


def handle() -> ib.Handle:
    """Device-mode handle bound to torch's current stream on the current device."""
    dev = torch.cuda.current_device()
    stream = torch.cuda.current_stream().cuda_stream
    key = (dev, stream)
    h = _handles.get(key)
    if h is None:
        h = ib.Handle(memory_mode=ib.MEMORY_DEVICE, stream=stream)
        _handles[key] = h
    return h


def cm(t: torch.Tensor) -> torch.Tensor:
    """Row-major (rows, cols) tensor -> column-major storage (cols, rows)."""
    return t.t().contiguous()


def rm(t: torch.Tensor) -> torch.Tensor:
    """Column-major storage (cols, rows) -> row-major (rows, cols) view."""
    return t.t()


def _dims(t: torch.Tensor):
    cols, rows = t.shape
    return rows, cols


class Sgemm(torch.autograd.Function):
    """C = alpha * op(A) * op(B) (beta = 0), column-major storage tensors."""

    @staticmethod
    def forward(ctx, A, B, alpha: float = 1.0, transa: int = 0, transb: int = 0):
        ra, ca = _dims(A)
        rb, cb = _dims(B)
        m, k = (ra, ca) if transa == 0 else (ca, ra)
        k2, n = (rb, cb) if transb == 0 else (cb, rb)
        if k != k2:
            raise ValueError("inner dimensions differ")
        C = torch.empty((n, m), device=A.device, dtype=torch.float32)
        ib.check(ib.iclblasSgemm(handle().value, transa, transb, m, n, k, alpha, A.data_ptr(), ra, B.data_ptr(), rb,
                                 0.0, C.data_ptr(), m), "iclblasSgemm")
        ctx.save_for_backward(A, B)
        ctx.cfg = (alpha, transa, transb, m, n, k)
        return C

    @staticmethod
    def backward(ctx, dC):
        A, B = ctx.saved_tensors
        alpha, transa, transb, m, n, k = ctx.cfg
        dC = dC.contiguous()
        dA = torch.empty_like(A) if ctx.needs_input_grad[0] else None
        dB = torch.empty_like(B) if ctx.needs_input_grad[1] else None
        ra, rb = A.shape[1], B.shape[1]
        ib.check(ib.iclblasSgemmBackward(handle().value, transa, transb, m, n, k, alpha, A.data_ptr(), ra,
                                         B.data_ptr(), rb, dC.data_ptr(), m,
                                         dA.data_ptr() if dA is not None else None, ra,
                                         dB.data_ptr() if dB is not None else None, rb), "iclblasSgemmBackward")
        return dA, dB, None, None, None


class Strsm(torch.autograd.Function):
    """X = alpha * op(A)^-1 * B (left) or alpha * B * op(A)^-1 (right)."""

    @staticmethod
    def forward(ctx, A, B, alpha: float = 1.0, side: int = 0, uplo: int = 1, trans: int = 0, diag: int = 0):
        m, n = _dims(B)
        X = B.detach().clone()
        ib.check(ib.iclblasStrsm(handle().value, side, uplo, trans, diag, m, n, alpha, A.data_ptr(), A.shape[1],
                                 X.data_ptr(), m), "iclblasStrsm")
        ctx.save_for_backward(A, X)
        ctx.cfg = (alpha, side, uplo, trans, diag, m, n)
        return X

    @staticmethod
    def backward(ctx, dX):
        A, X = ctx.saved_tensors
        alpha, side, uplo, trans, diag, m, n = ctx.cfg
        dX = dX.contiguous()
        dA = torch.empty_like(A) if ctx.needs_input_grad[0] else None
        dB = torch.empty_like(X) if ctx.needs_input_grad[1] else None
        ib.check(ib.iclblasStrsmBackward(handle().value, side, uplo, trans, diag, m, n, alpha, A.data_ptr(),
                                         A.shape[1], X.data_ptr(), m, dX.data_ptr(), m,
                                         dA.data_ptr() if dA is not None else None, A.shape[1],
                                         dB.data_ptr() if dB is not None else None, m), "iclblasStrsmBackward")
        return dA, dB, None, None, None, None, None


def sgemm(A, B, alpha=1.0, transa=0, transb=0):
    return Sgemm.apply(A, B, alpha, transa, transb)


def strsm(A, B, alpha=1.0, side=0, uplo=1, trans=0, diag=0):
    return Strsm.apply(A, B, alpha, side, uplo, trans, diag)
