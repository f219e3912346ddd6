"""clgpu_amd: MI355X (gfx950) level-3 BLAS behind the iclBLAS C API.

The product is the C-ABI library libiclBLAS.so (include/iclBLAS.h,
include/iclBLAS_ext.h) built from clgpu_amd/csrc by clgpu_amd.build.
clgpu_amd.iclblas is its ctypes binding; clgpu_amd.autograd wraps the
adjoint entry points in torch.autograd.Function for gradient tests.
"""
from . import iclblas  # noqa: F401

__all__ = ["iclblas"]
