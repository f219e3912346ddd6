"""Seeded synthetic matrices for the level-3 workloads.

Generator: SplitMix64. Stream s of seed `seed` starts from
state0 = mix64(seed ^ (0xA0761D6478BD642F * (s + 1))) and draw i is
x_i = mix64(state0 + (i + 1) * 0x9E3779B97F4A7C15) (mod 2^64), where mix64 is
the SplitMix64 finaliser. Streams: A = 0, B = 1, C = 2, G = 3, H = 4, W = 5.

uniform[-1, 1]:  v = 2u - 1 with u = ((x >> 40) + 0.5) * 2^-24, i.e. an odd
                 multiple of 2^-24 minus 1 -- exactly representable in FP32.
normal(0, 1):    Box-Muller on two consecutive draws, cos branch, in FP64.

Matrices are column-major; all ld * cols entries (padding included) are drawn
in storage order so that a padded matrix has the same values in its leading
rows as an unpadded one only when ld is equal; tests set padding explicitly
where it matters.
"""
from __future__ import annotations

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
STREAM_MUL = 0xA0761D6478BD642F
STREAMS = {"A": 0, "B": 1, "C": 2, "G": 3, "H": 4, "W": 5}


def _mix64(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _state0(seed: int, stream: int) -> np.uint64:
    with np.errstate(over="ignore"):
        s = np.array([(seed ^ ((STREAM_MUL * (stream + 1)) & 0xFFFFFFFFFFFFFFFF)) & 0xFFFFFFFFFFFFFFFF],
                     dtype=np.uint64)
        return _mix64(s)[0]


def draws(count: int, seed: int = 0, stream: int = 0, offset: int = 0) -> np.ndarray:
    """Raw 64-bit draws offset .. offset+count-1 of a stream."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        return _mix64(_state0(seed, stream) + i * GOLDEN)


def uniform_pm1(count: int, seed: int = 0, stream: int = 0) -> np.ndarray:
    """`count` values in (-1, 1), float64 (exact in float32)."""
    x = draws(count, seed, stream) >> np.uint64(40)
    return (x.astype(np.float64) * 2.0 + 1.0) * 2.0 ** -24 - 1.0


def normal(count: int, seed: int = 0, stream: int = 0) -> np.ndarray:
    x = draws(2 * count, seed, stream) >> np.uint64(40)
    u = (x.astype(np.float64) + 0.5) * 2.0 ** -24
    u1, u2 = u[0::2], u[1::2]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def matrix(rows: int, cols: int, ld: int | None = None, name: str = "A", seed: int = 0,
           dtype=np.float32) -> np.ndarray:
    """Column-major rows x cols matrix (uniform[-1,1]) stored in an (ld x cols)
    Fortran-ordered array; returns the full ld x cols storage."""
    ld = rows if ld is None else ld
    v = uniform_pm1(ld * cols, seed, STREAMS[name])
    return np.asfortranarray(v.reshape(cols, ld).T.astype(dtype))


def complex_matrix(rows: int, cols: int, ld: int | None = None, name: str = "A", seed: int = 0) -> np.ndarray:
    """complex64 column-major matrix; real and imaginary parts are consecutive
    draws (interleaved storage order)."""
    ld = rows if ld is None else ld
    v = uniform_pm1(2 * ld * cols, seed, STREAMS[name])
    c = (v[0::2] + 1j * v[1::2]).astype(np.complex64)
    return np.asfortranarray(c.reshape(cols, ld).T)


def triangular(order: int, lower: bool = True, ld: int | None = None, name: str = "A", seed: int = 0,
               poison: bool = True, offdiag_scale: float | None = None) -> np.ndarray:
    """Triangular test matrix: diagonal = sign * U[1, 2] (sign +-1 with equal
    probability, from bit 0 of the draw), strict stored triangle = U[-1, 1] *
    offdiag_scale (default 1/sqrt(order)), the other triangle NaN when `poison`
    (it must never be read) and 0 otherwise."""
    ld = order if ld is None else ld
    scale = 1.0 / np.sqrt(order) if offdiag_scale is None else offdiag_scale
    x = draws(ld * order, seed, STREAMS[name])
    u = (((x >> np.uint64(40)).astype(np.float64) + 0.5) * 2.0 ** -24)
    full = (2.0 * u - 1.0).reshape(order, ld).T  # ld x order, column-major values
    sign = np.where((x & np.uint64(1)) == 1, -1.0, 1.0).reshape(order, ld).T
    out = np.full((ld, order), np.nan if poison else 0.0)
    r = np.arange(ld)[:, None]
    c = np.arange(order)[None, :]
    stored = (r > c) if lower else (r < c)
    stored &= r < order
    out[stored] = (full * scale)[stored]
    d = np.arange(order)
    out[d, d] = sign[d, d] * (1.0 + u.reshape(order, ld).T[d, d])
    out[order:, :] = 0.0  # padding rows
    return np.asfortranarray(out.astype(np.float32))


def complex_triangular(order: int, lower: bool = True, ld: int | None = None, name: str = "A", seed: int = 0,
                       poison: bool = True) -> np.ndarray:
    """complex64 triangular test matrix: the real parts are triangular(...)'s
    values (diagonal sign * U[1, 2], stored strict triangle U[-1, 1] / sqrt(order)),
    the imaginary parts of the stored triangle are an independent draw of the same
    kind scaled like the off-diagonal (so the diagonal dominates); the other
    triangle is NaN + NaN i when `poison`, 0 otherwise."""
    ld = order if ld is None else ld
    re = triangular(order, lower, ld, name, seed, poison).astype(np.float64)
    im = triangular(order, lower, ld, name, seed + 7919, poison).astype(np.float64)
    d = np.arange(order)
    im[d, d] = 0.5 * im[d, d] / np.sqrt(order)
    return np.asfortranarray((re + 1j * im).astype(np.complex64))


def quarter_integers(count: int, seed: int = 0, stream: int = 0) -> np.ndarray:
    """Values in {-1, -0.75, ..., 1} (uniform over the 9 values): the data family
    of the reference's implementation-level tests (tests/blas_funcs/test_helpers.hpp:120-127),
    for which every product and partial sum of a 256-term dot product is exact
    in FP32 and in BF16."""
    x = draws(count, seed, stream) >> np.uint64(40)
    return ((x % np.uint64(9)).astype(np.float64) - 4.0) * 0.25
