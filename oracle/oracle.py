"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle (liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker or as the timed CPU baseline.  See
oracle/zynum_oracle.c for the reference files each function restates.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-B", "-C", str(HERE), "liboracle.so"], check=True)
    return LIB


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        vp, i, d, f, ch = C.c_void_p, C.c_int, C.c_double, C.c_float, C.c_char
        lib.oracle_dgemm.argtypes = [ch, ch, i, i, i, d, vp, i, vp, i, d, vp, i]
        lib.oracle_sgemm.argtypes = [ch, ch, i, i, i, f, vp, i, vp, i, f, vp, i]
        lib.oracle_dtrsm.argtypes = [ch, ch, ch, ch, i, i, d, vp, i, vp, i]
        lib.oracle_dsyrk.argtypes = [ch, ch, i, i, d, vp, i, d, vp, i]
        lib.oracle_dd_gemm_entries.argtypes = [ch, ch, i, d, vp, i, vp, i, d, vp, i, vp, vp, C.c_int64, vp, vp]
        lib.oracle_fill_uniform.argtypes = [C.c_uint64, C.c_uint64, vp, C.c_int64]
        lib.oracle_now.restype = d
        lib.oracle_set_threads.argtypes = [i]
        lib.oracle_force_isa.argtypes = [i]
        _lib = lib
    return _lib


def _p(x: np.ndarray | None):
    return None if x is None else x.ctypes.data


def _c(t) -> bytes:
    return t.encode() if isinstance(t, str) else t


def dgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc) -> None:
    """In-place on ``c`` (float64 flat column-major buffers)."""
    load().oracle_dgemm(_c(ta), _c(tb), m, n, k, alpha, _p(a), lda, _p(b), ldb, beta, _p(c), ldc)


def sgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc) -> None:
    load().oracle_sgemm(_c(ta), _c(tb), m, n, k, alpha, _p(a), lda, _p(b), ldb, beta, _p(c), ldc)


def dtrsm(side, uplo, trans, diag, m, n, alpha, a, lda, b, ldb) -> None:
    load().oracle_dtrsm(_c(side), _c(uplo), _c(trans), _c(diag), m, n, alpha, _p(a), lda, _p(b), ldb)


def dsyrk(uplo, trans, n, k, alpha, a, lda, beta, c, ldc) -> None:
    load().oracle_dsyrk(_c(uplo), _c(trans), n, k, alpha, _p(a), lda, beta, _p(c), ldc)


def dd_gemm_entries(ta, tb, k, alpha, a, lda, b, ldb, beta, c0, ldc, rows, cols):
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    hi = np.empty(rows.size)
    lo = np.empty(rows.size)
    load().oracle_dd_gemm_entries(_c(ta), _c(tb), k, alpha, _p(a), lda, _p(b), ldb, beta, _p(c0), ldc,
                                  _p(rows), _p(cols), rows.size, _p(hi), _p(lo))
    return hi, lo


def fill_uniform(seed: int, offset: int, count: int) -> np.ndarray:
    out = np.empty(count)
    load().oracle_fill_uniform(seed, offset, _p(out), count)
    return out


def set_threads(t: int) -> int:
    return load().oracle_set_threads(t)


def threads() -> int:
    return load().oracle_threads()


def isa() -> int:
    return load().oracle_isa()


def force_isa(v: int) -> int:
    return load().oracle_force_isa(v)


def now() -> float:
    return load().oracle_now()


def _cfuncs():
    lib = load()
    if not getattr(lib, "_complex_bound", False):
        vp, i, ch = C.c_void_p, C.c_int, C.c_char
        for name in ("oracle_zgemm", "oracle_cgemm"):
            getattr(lib, name).argtypes = [ch, ch, i, i, i, vp, vp, i, vp, i, vp, vp, i]
        for name in ("oracle_zgemm_entries", "oracle_cgemm_entries", "oracle_zdd_gemm_entries"):
            extra = [vp] if name == "oracle_zdd_gemm_entries" else []
            getattr(lib, name).argtypes = [ch, ch, i, vp, vp, i, vp, i, vp, vp, i, vp, vp, C.c_int64, vp] + extra
        lib._complex_bound = True
    return lib


def _scal(v, dtype) -> np.ndarray:
    v = complex(v)
    return np.array([v.real, v.imag], dtype=dtype)


def zgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc) -> None:
    """In-place on ``c``; a, b, c are complex128 flat column-major buffers,
    leading dimensions in complex elements (gemm.zig:1234-1246)."""
    al, be = _scal(alpha, np.float64), _scal(beta, np.float64)
    _cfuncs().oracle_zgemm(_c(ta), _c(tb), m, n, k, _p(al), _p(a), lda, _p(b), ldb, _p(be), _p(c), ldc)


def cgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc) -> None:
    """complex64 form of :func:`zgemm`."""
    al, be = _scal(alpha, np.float32), _scal(beta, np.float32)
    _cfuncs().oracle_cgemm(_c(ta), _c(tb), m, n, k, _p(al), _p(a), lda, _p(b), ldb, _p(be), _p(c), ldc)


def zgemm_entries(ta, tb, k, alpha, a, lda, b, ldb, beta, c0, ldc, rows, cols) -> np.ndarray:
    """Oracle value of selected entries of a complex128 product (c0 may be None when beta == 0)."""
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    out = np.empty(rows.size, dtype=np.complex128)
    al, be = _scal(alpha, np.float64), _scal(beta, np.float64)
    _cfuncs().oracle_zgemm_entries(_c(ta), _c(tb), k, _p(al), _p(a), lda, _p(b), ldb, _p(be), _p(c0), ldc,
                                   _p(rows), _p(cols), rows.size, _p(out))
    return out


def zdd_gemm_entries(ta, tb, k, alpha, a, lda, b, ldb, beta, c0, ldc, rows, cols):
    """Double-double value (hi, lo as complex128 arrays) of selected entries of a complex product."""
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    hi = np.empty(rows.size, dtype=np.complex128)
    lo = np.empty(rows.size, dtype=np.complex128)
    al, be = _scal(alpha, np.float64), _scal(beta, np.float64)
    _cfuncs().oracle_zdd_gemm_entries(_c(ta), _c(tb), k, _p(al), _p(a), lda, _p(b), ldb, _p(be), _p(c0), ldc,
                                      _p(rows), _p(cols), rows.size, _p(hi), _p(lo))
    return hi, lo


def _l3funcs():
    lib = load()
    if not getattr(lib, "_l3_bound", False):
        vp, i, d, ch = C.c_void_p, C.c_int, C.c_double, C.c_char
        lib.oracle_dsymm.argtypes = [ch, ch, i, i, d, vp, i, vp, i, d, vp, i]
        lib.oracle_dsyr2k.argtypes = [ch, ch, i, i, d, vp, i, vp, i, d, vp, i]
        lib.oracle_dtrmm.argtypes = [ch, ch, ch, ch, i, i, d, vp, i, vp, i]
        lib._l3_bound = True
    return lib


def dsymm(side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc) -> None:
    """In-place on ``c`` (symmetric.zig:56-70)."""
    _l3funcs().oracle_dsymm(_c(side), _c(uplo), m, n, alpha, _p(a), lda, _p(b), ldb, beta, _p(c), ldc)


def dsyr2k(uplo, trans, n, k, alpha, a, lda, b, ldb, beta, c, ldc) -> None:
    """In-place on the stored triangle of ``c`` (symmetric.zig:301-329)."""
    _l3funcs().oracle_dsyr2k(_c(uplo), _c(trans), n, k, alpha, _p(a), lda, _p(b), ldb, beta, _p(c), ldc)


def dtrmm(side, uplo, trans, diag, m, n, alpha, a, lda, b, ldb) -> None:
    """In-place on ``b`` (triangular.zig:202-231)."""
    _l3funcs().oracle_dtrmm(_c(side), _c(uplo), _c(trans), _c(diag), m, n, alpha, _p(a), lda, _p(b), ldb)
