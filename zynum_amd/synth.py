"""Seeded synthetic inputs (splitmix64, counter form).

value_i = (z_i >> 11) * 2^-52 - 1 with z_i = mix(seed + (offset + i + 1) * 0x9e3779b97f4a7c15),
uniform on [-1, 1) with 52 random bits, exact in binary64.  A, B and C of a
problem are drawn from one stream in that order, each in column-major order
over its logical rows x cols (padding rows are filled with a sentinel).
"""
from __future__ import annotations

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def uniform(seed: int, offset: int, count: int) -> np.ndarray:
    """``count`` values of the stream starting at ``offset`` (float64)."""
    out = np.empty(count, dtype=np.float64)
    step = 1 << 22
    with np.errstate(over="ignore"):
        for s in range(0, count, step):
            n = min(step, count - s)
            z = np.arange(offset + s + 1, offset + s + 1 + n, dtype=np.uint64)
            z *= GOLDEN
            z += np.uint64(seed)
            z ^= z >> np.uint64(30)
            z *= M1
            z ^= z >> np.uint64(27)
            z *= M2
            z ^= z >> np.uint64(31)
            out[s:s + n] = (z >> np.uint64(11)).astype(np.float64) * 2.0**-52 - 1.0
    return out


def matrix(seed: int, offset: int, rows: int, cols: int, ld: int, dtype=np.float64,
           sentinel: float = -999.0) -> np.ndarray:
    """Column-major rows x cols matrix with leading dimension ld as a flat array."""
    flat = np.full(max(1, ld * cols), sentinel, dtype=dtype)
    if rows > 0 and cols > 0:
        vals = uniform(seed, offset, rows * cols).astype(dtype)
        view = flat[: ld * cols].reshape(cols, ld)
        view[:, :rows] = vals.reshape(cols, rows)
    return flat


def wide_range(seed: int, offset: int, rows: int, cols: int, ld: int, emin: int = -16, emax: int = 16) -> np.ndarray:
    """a_ij = u_ij * 2^e_ij, u uniform [-1, 1), e uniform integer in [emin, emax]."""
    flat = np.full(max(1, ld * cols), -999.0, dtype=np.float64)
    n = rows * cols
    u = uniform(seed, offset, n)
    r = uniform(seed ^ 0x5DEECE66D, offset, n)
    e = np.floor((r + 1.0) * 0.5 * (emax - emin + 1)).astype(np.int64) + emin
    e = np.minimum(e, emax)
    vals = np.ldexp(u, e)
    view = flat[: ld * cols].reshape(cols, ld)
    view[:, :rows] = vals.reshape(cols, rows)
    return flat
