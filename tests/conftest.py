import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (MI355X)")
    config.addinivalue_line("markers", "perf: wall-clock checks on a GPU, kept out of the gpu parity suite")


@pytest.fixture(scope="session")
def oracle():
    import oracle as orc  # noqa: WPS433  (oracle/oracle.py)

    orc.load()
    return orc


@pytest.fixture(scope="session")
def zlib():
    from zynum_amd import _lib

    return _lib.load()


@pytest.fixture(scope="session")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


# ---------------------------------------------------------------------------
# helpers shared by tests: Zynum's own test-input generators
# ---------------------------------------------------------------------------
def sample_value(row: int, col: int, salt: int) -> float:
    """fortran_compat_test.zig:12-16."""
    raw = ((row + 1) * (salt + 3) + (col + 1) * (salt + 7)) % 29
    return (raw - 14.0) / 9.0


def sample_int(row: int, col: int, salt: int) -> float:
    """Integer-valued variant (9 * sample_value) in [-14, 14]: exact products/sums."""
    raw = ((row + 1) * (salt + 3) + (col + 1) * (salt + 7)) % 29
    return float(raw - 14)


def fill_padded(rows: int, cols: int, ld: int, salt: int, sentinel: float, dtype=np.float64, integer=False):
    """fortran_compat_test.zig:196-203: sentinel everywhere, samples in the logical part."""
    buf = np.full(max(1, ld * cols), sentinel, dtype=dtype)
    f = sample_int if integer else sample_value
    i = np.arange(rows)[:, None]
    j = np.arange(cols)[None, :]
    raw = ((i + 1) * (salt + 3) + (j + 1) * (salt + 7)) % 29
    vals = (raw - 14.0) if integer else (raw - 14.0) / 9.0
    if rows and cols:
        view = buf[: ld * cols].reshape(cols, ld)
        view[:, :rows] = vals.T.astype(dtype)
    _ = f
    return buf


class XorShift:
    """test/abi/reference.zig:11-39 (xorshift64, small() in [-2, 2))."""

    def __init__(self, seed: int):
        self.state = seed if seed != 0 else 0x9E3779B97F4A7C15

    def next(self) -> int:
        x = self.state
        x ^= (x << 13) & 0xFFFFFFFFFFFFFFFF
        x ^= x >> 7
        x ^= (x << 17) & 0xFFFFFFFFFFFFFFFF
        self.state = x
        return x

    def small(self) -> float:
        bits = self.next() >> 11
        return (bits / 9007199254740992.0) * 4.0 - 2.0


def fill_row_major(rng: XorShift, rows: int, cols: int, ld: int, dtype=np.float64) -> np.ndarray:
    """reference.zig:141-146."""
    a = np.array([-57.0 - i for i in range(rows * ld)], dtype=dtype)
    for i in range(rows):
        for j in range(cols):
            a[i * ld + j] = rng.small()
    return a


def exact_gemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """Exact result with Python Fractions (small shapes only)."""
    from fractions import Fraction as Fr

    out = c.copy()
    for j in range(n):
        for i in range(m):
            s = Fr(0)
            for p in range(k):
                av = a[i + p * lda] if ta == "N" else a[p + i * lda]
                bv = b[p + j * ldb] if tb == "N" else b[j + p * ldb]
                s += Fr(float(av)) * Fr(float(bv))
            val = Fr(float(alpha)) * s
            if beta != 0:
                val += Fr(float(beta)) * Fr(float(c[i + j * ldc]))
            out[i + j * ldc] = float(val)
    return out


def int_gemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """Exact result for integer-valued inputs and dyadic alpha/beta via int64 numpy."""
    def mat(x, rows, cols, ld):
        return x[: ld * cols].reshape(cols, ld)[:, :rows].T

    A = mat(a, m, k, lda) if ta == "N" else mat(a, k, m, lda).T
    B = mat(b, k, n, ldb) if tb == "N" else mat(b, n, k, ldb).T
    S = A.astype(np.int64) @ B.astype(np.int64)
    out = c.copy()
    view = out[: ldc * n].reshape(n, ldc)
    old = view[:, :m].T.astype(np.float64)
    res = alpha * S.astype(np.float64) + (beta * old if beta != 0 else 0.0)
    view[:, :m] = res.T.astype(out.dtype)
    return out


# ---------------------------------------------------------------------------
# complex helpers (interleaved storage held as numpy complex arrays)
# ---------------------------------------------------------------------------
COMPLEX_SENTINEL = complex(-77, 33)  # fortran_compat_test.zig:476


def fill_padded_complex(rows: int, cols: int, ld: int, salt: int, dtype=np.complex128, integer=False):
    """fortran_compat_test.zig:175-194: re = sample(i, j, salt), im = sample(i, j, salt + 11),
    the sentinel -77 + 33i outside the logical part."""
    re = fill_padded(rows, cols, ld, salt, COMPLEX_SENTINEL.real, integer=integer)
    im = fill_padded(rows, cols, ld, salt + 11, COMPLEX_SENTINEL.imag, integer=integer)
    return (re + 1j * im).astype(dtype)


def cmat(x: np.ndarray, rows: int, cols: int, ld: int) -> np.ndarray:
    """Column-major rows x cols view (copy) of a flat complex buffer."""
    return x[: ld * cols].reshape(cols, ld)[:, :rows].T.copy()


def cop(x: np.ndarray, t: str) -> np.ndarray:
    """op(X) for 'N', 'T', 'C'."""
    return x if t == "N" else (x.T if t == "T" else x.conj().T)


def int_cgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """Exact complex result for integer-valued inputs and integer alpha/beta
    (every partial sum is an integer far below 2^53)."""
    A = cop(cmat(a, k if ta != "N" else m, m if ta != "N" else k, lda), ta)
    B = cop(cmat(b, n if tb != "N" else k, k if tb != "N" else n, ldb), tb)
    S = A.astype(np.complex128) @ B.astype(np.complex128)
    out = c.copy()
    view = out[: ldc * n].reshape(n, ldc)
    old = view[:, :m].T.astype(np.complex128)
    res = complex(alpha) * S + (complex(beta) * old if complex(beta) != 0 else 0)
    view[:, :m] = res.T.astype(out.dtype)
    return out
