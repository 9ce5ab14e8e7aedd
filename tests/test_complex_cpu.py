"""CPU checks of the complex GEMM (CGEMM / ZGEMM) path: the oracle pinned
against the reference's known answers, exact arithmetic and the rounding rules
it restates, and the C ABI's argument handling without a device."""
from fractions import Fraction as Fr

import ctypes as C
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from conftest import cmat, cop, fill_padded_complex, int_cgemm

ROOT = Path(__file__).resolve().parent.parent
PAIRS = [(a, b) for a in "NTC" for b in "NTC"]


def _small_case(dtype):
    """fortran_compat_test.zig:732-781 (m = n = 2, k = 3, lda 2, ldb 3, ldc 2)."""
    a = np.array([1 + 1j, 2 - 1j, -3 + 0.5j, 4 + 2j, 0.25 - 2j, -1 + 1j], dtype=dtype)
    b = np.array([2, -1 + 1j, 0.5 - 0.5j, 1 - 2j, 3 + 1j, -2 + 0.25j], dtype=dtype)
    c = np.array([1 + 0.5j, -2 + 1j, -1j, 3 + 2j], dtype=dtype)
    return a, b, c, 0.75 - 0.25j, -0.5 + 0.25j


def _exact_complex(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """Exact value with Fractions (re, im pairs); small shapes only."""
    def fr(z):
        return Fr(float(z.real)), Fr(float(z.imag))

    def mul(x, y):
        return x[0] * y[0] - x[1] * y[1], x[0] * y[1] + x[1] * y[0]

    def at(x, ld, r, col, t):
        v = x[r + col * ld] if t == "N" else x[col + r * ld]
        re, im = fr(v)
        return (re, -im) if t == "C" else (re, im)

    al, be = fr(complex(alpha)), fr(complex(beta))
    out = []
    for j in range(n):
        for i in range(m):
            s = (Fr(0), Fr(0))
            for p in range(k):
                pr = mul(at(a, lda, i, p, ta), at(b, ldb, p, j, tb))
                s = (s[0] + pr[0], s[1] + pr[1])
            v = mul(al, s)
            if be != (0, 0):
                bc = mul(be, fr(c[i + j * ldc]))
                v = (v[0] + bc[0], v[1] + bc[1])
            out.append(v)
    return out


@pytest.mark.parametrize("dtype,fn", [(np.complex128, "zgemm"), (np.complex64, "cgemm")])
def test_oracle_known_answer_no_trans(oracle, dtype, fn):
    """The reference's small CGEMM/ZGEMM case: every intermediate is a short
    dyadic fraction, so the result is the exact value bit for bit."""
    a, b, c, alpha, beta = _small_case(dtype)
    exact = _exact_complex("N", "N", 2, 2, 3, alpha, a, 2, b, 3, beta, c, 2)
    getattr(oracle, fn)("N", "N", 2, 2, 3, alpha, a, 2, b, 3, beta, c, 2)
    for got, (re, im) in zip(c, exact):
        assert got.real == float(re) and got.imag == float(im)


@pytest.mark.parametrize("ta,tb", PAIRS)
@pytest.mark.parametrize("dtype,fn", [(np.complex128, "zgemm"), (np.complex64, "cgemm")])
def test_oracle_integer_exact_all_transposes(oracle, ta, tb, dtype, fn):
    """Integer-valued inputs, padded storage (lda = rows + 4, ldb = rows + 5,
    ldc = m + 6 as fortran_compat_test.zig:497-504): the oracle equals exact
    arithmetic on op(A) op(B) (numpy's own conjugate / transpose), and padding
    keeps its sentinel."""
    m, n, k = 7, 5, 9
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    lda, ldb, ldc = ar + 4, br + 5, m + 6
    a = fill_padded_complex(ar, ac, lda, 37, dtype, integer=True)
    b = fill_padded_complex(br, bc, ldb, 41, dtype, integer=True)
    c = fill_padded_complex(m, n, ldc, 43, dtype, integer=True)
    alpha, beta = 2 - 3j, -1 + 2j
    want = int_cgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
    getattr(oracle, fn)(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
    assert np.array_equal(c, want)


def test_oracle_products_are_not_fused(oracle):
    """scalar.zig:77-82 rounds each of the four products: with
    a = b = (1 + 2^-30) + 1i the real part a.re*b.re - a.im*b.im loses the
    2^-60 term a fused form would keep."""
    x = 1 + 2.0 ** -30
    a = np.array([x + 1j])
    b = np.array([x + 1j])
    c = np.zeros(1, dtype=np.complex128)
    oracle.zgemm("N", "N", 1, 1, 1, 1.0, a, 1, b, 1, 0.0, c, 1)
    assert c[0].real == 2.0 ** -29          # fused would give 2^-29 + 2^-60
    assert c[0].real != 2.0 ** -29 + 2.0 ** -60
    assert c[0].imag == 2 * x


def test_oracle_summation_order_is_ascending(oracle):
    """gemm.zig:1241: sum = sum + a(i, p) b(p, j) for p = 0, 1, 2: with terms
    1, 2^-53, 2^-53 the ascending sum keeps 1 (each addition ties to even),
    while summing the small terms first would give 1 + 2^-52."""
    a = np.array([1 + 0j, 2.0 ** -53, 2.0 ** -53])
    b = np.ones(3, dtype=np.complex128)
    c = np.zeros(1, dtype=np.complex128)
    oracle.zgemm("N", "N", 1, 1, 3, 1.0, a, 1, b, 3, 0.0, c, 1)
    assert c[0].real == 1.0


def test_oracle_inactive_nans(oracle):
    """fortran_compat_test.zig:1047-1074: beta == 0 never reads C; alpha == 0
    never reads A or B."""
    nan = float("nan")
    a = np.array([2 + 1j]); b = np.array([3 - 1j]); c = np.array([complex(nan, nan)])
    oracle.zgemm("N", "N", 1, 1, 1, 1.0, a, 1, b, 1, 0.0, c, 1)
    assert c[0] == 7 + 1j
    a = np.array([complex(nan, nan)]); b = np.array([complex(nan, nan)]); c = np.array([4 + 3j])
    oracle.zgemm("N", "N", 1, 1, 1, 0.0, a, 1, b, 1, 2 - 1j, c, 1)
    assert c[0] == 11 + 2j


def test_oracle_k_zero_scales_c(oracle):
    """gemm.zig:1407-1414: k == 0 gives beta * C, beta == 0 gives exact zeros."""
    c = np.array([1 + 2j, float("nan")])
    oracle.zgemm("N", "N", 2, 1, 0, 1.0, None, 2, None, 1, 0.0, c, 2)
    assert np.array_equal(c, np.zeros(2, dtype=np.complex128))
    c = np.array([1 + 2j, 3 - 1j])
    oracle.zgemm("N", "N", 2, 1, 0, 1.0, None, 2, None, 1, 1j, c, 2)
    assert np.array_equal(c, np.array([-2 + 1j, 1 + 3j]))


def test_oracle_entries_match_full(oracle):
    m, n, k = 9, 7, 11
    rng = np.random.default_rng(5)
    a = rng.standard_normal(m * k) + 1j * rng.standard_normal(m * k)
    b = rng.standard_normal(k * n) + 1j * rng.standard_normal(k * n)
    c0 = rng.standard_normal(m * n) + 1j * rng.standard_normal(m * n)
    for ta, tb in (("N", "N"), ("C", "T")):
        full = c0.copy()
        oracle.zgemm(ta, tb, m, n, k, 0.5 - 1j, a, m if ta == "N" else k, b, k if tb == "N" else n, 2 + 1j,
                     full, m)
        rows = np.array([0, 3, 8, 5]); cols = np.array([0, 6, 2, 4])
        got = oracle.zgemm_entries(ta, tb, k, 0.5 - 1j, a, m if ta == "N" else k, b, k if tb == "N" else n,
                                   2 + 1j, c0, m, rows, cols)
        assert np.array_equal(got, full[rows + cols * m])


def test_dd_complex_entries_against_fractions(oracle):
    """The double-double complex checker is within 2^-100 (relative to the
    magnitude sum) of the exact value."""
    m, n, k = 4, 3, 25
    rng = np.random.default_rng(9)
    a = rng.standard_normal(m * k) + 1j * rng.standard_normal(m * k)
    b = rng.standard_normal(k * n) + 1j * rng.standard_normal(k * n)
    c0 = rng.standard_normal(m * n) + 1j * rng.standard_normal(m * n)
    alpha, beta = 0.6 - 0.4j, -0.2 + 0.3j
    for ta, tb in (("N", "N"), ("C", "N"), ("T", "C")):
        lda = m if ta == "N" else k
        ldb = k if tb == "N" else n
        exact = _exact_complex(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c0, m)
        rows = np.repeat(np.arange(m), n)
        cols = np.tile(np.arange(n), m)
        hi, lo = oracle.zdd_gemm_entries(ta, tb, k, alpha, a, lda, b, ldb, beta, c0, m, rows, cols)
        scale = float(np.abs(a).sum() * np.abs(b).sum() + np.abs(c0).sum())
        for e, (i, j) in enumerate(zip(rows, cols)):
            re, im = exact[j * m + i]
            err_re = abs(Fr(hi[e].real) + Fr(lo[e].real) - re)
            err_im = abs(Fr(hi[e].imag) + Fr(lo[e].imag) - im)
            assert err_re <= Fr(scale) * Fr(2) ** -100
            assert err_im <= Fr(scale) * Fr(2) ** -100


def test_int_cgemm_helper_against_fractions():
    """The exact integer helper used by the GPU tests agrees with Fractions."""
    m, n, k = 3, 4, 5
    for ta, tb in (("N", "C"), ("T", "N"), ("C", "T")):
        ar, ac = (m, k) if ta == "N" else (k, m)
        br, bc = (k, n) if tb == "N" else (n, k)
        a = fill_padded_complex(ar, ac, ar + 1, 3, integer=True)
        b = fill_padded_complex(br, bc, br + 2, 7, integer=True)
        c = fill_padded_complex(m, n, m + 1, 11, integer=True)
        got = int_cgemm(ta, tb, m, n, k, 2 - 1j, a, ar + 1, b, br + 2, 1 + 1j, c, m + 1)
        exact = _exact_complex(ta, tb, m, n, k, 2 - 1j, a, ar + 1, b, br + 2, 1 + 1j, c, m + 1)
        for j in range(n):
            for i in range(m):
                re, im = exact[j * m + i]
                assert got[i + j * (m + 1)] == complex(float(re), float(im))
        _ = cmat, cop


# ---------------------------------------------------------------------------
# C ABI without a device
# ---------------------------------------------------------------------------
def _child(code: str) -> subprocess.CompletedProcess:
    return subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=str(ROOT),
                          env={"PYTHONPATH": str(ROOT), "HIP_VISIBLE_DEVICES": "", "PATH": "/usr/bin:/bin"})


@pytest.mark.parametrize("routine,dt", [("zgemm_", "complex128"), ("cgemm_", "complex64")])
@pytest.mark.parametrize("args,expected", [
    (("X", "N", 1, 1, 1, 1, 1, 1), 1), (("N", "Q", 1, 1, 1, 1, 1, 1), 2), (("N", "N", -1, 1, 1, 1, 1, 1), 3),
    (("N", "N", 1, -1, 1, 1, 1, 1), 4), (("N", "N", 1, 1, -1, 1, 1, 1), 5), (("C", "N", 1, 1, 3, 2, 3, 1), 8),
    (("N", "C", 1, 3, 1, 1, 2, 1), 10), (("N", "N", 4, 1, 1, 4, 1, 3), 13),
])
def test_complex_parameter_numbers_and_untouched_c(routine, dt, args, expected):
    """fortran.zig:768-775 with gemmError (fortran.zig:88-103): the parameter
    number goes to xerbla with the routine's name and C is left unchanged."""
    ta, tb, m, n, k, lda, ldb, ldc = args
    code = f"""
import ctypes as C, numpy as np
from zynum_amd import _lib
lib = _lib.load()
I = C.c_int32
a = np.full(16, 2 + 1j, dtype=np.{dt}); b = np.full(16, 3 - 1j, dtype=np.{dt}); c = np.full(16, 7 + 7j, dtype=np.{dt})
one = np.array([1 + 0j], dtype=np.{dt}); zero = np.array([0j], dtype=np.{dt})
lib.{routine}(b"{ta}", b"{tb}", C.byref(I({m})), C.byref(I({n})), C.byref(I({k})), one.ctypes.data,
           a.ctypes.data, C.byref(I({lda})), b.ctypes.data, C.byref(I({ldb})), zero.ctypes.data,
           c.ctypes.data, C.byref(I({ldc})))
assert (c == 7 + 7j).all(), c
"""
    r = _child(code)
    assert r.returncode == 0, r.stderr
    name = routine[:-1].upper()
    assert f" ** On entry to {name} parameter number {expected} had an illegal value" in r.stderr


def test_complex_decline_without_device_leaves_c_unchanged():
    code = """
import ctypes as C, numpy as np
from zynum_amd import _lib, blas
lib = _lib.load()
a = np.ones(4, dtype=np.complex128); c = np.full(4, 7 + 1j)
st = blas.gemm_request("C", "N", 2, 2, 2, 1 - 1j, a.ctypes.data, 2, a.ctypes.data, 2, 0.5j, c.ctypes.data, 2,
                       _lib.C64)
assert st == _lib.DECLINE_NO_DEVICE, st
assert (c == 7 + 1j).all()
one = np.array([1 + 0j]); zero = np.array([0j])
I = C.c_int32
lib.zgemm_(b"N", b"N", C.byref(I(2)), C.byref(I(2)), C.byref(I(2)), one.ctypes.data, a.ctypes.data,
           C.byref(I(2)), a.ctypes.data, C.byref(I(2)), zero.ctypes.data, c.ctypes.data, C.byref(I(2)))
assert lib.zynum_gpu_last_status() == _lib.DECLINE_NO_DEVICE
assert (c == 7 + 1j).all()
"""
    r = _child(code)
    assert r.returncode == 0, r.stderr


def test_complex_request_checks(zlib):
    from zynum_amd import _lib, blas

    c = np.full(4, 7 + 0j)
    # lda < m
    st = blas.gemm_request("N", "C", 2, 2, 2, 1.0, 0, 1, 0, 2, 0.0, c.ctypes.data, 2, _lib.C64)
    assert st == _lib.DECLINE_ARGUMENT
    # quick return m == 0 needs no device
    st = blas.gemm_request("N", "N", 0, 2, 2, 1j, 0, 1, 0, 2, 0.0, c.ctypes.data, 1, _lib.C64)
    assert st == _lib.OK
    # an imaginary part on a real request is refused before any device work
    st = blas.gemm_request(False, False, 2, 2, 2, 1j, 0, 2, 0, 2, 0.0, c.ctypes.data, 2, _lib.F64)
    assert st == _lib.DECLINE_UNSUPPORTED
    assert (c == 7).all()


def test_cblas_complex_invalid_enums_return_silently(zlib):
    """cblas.zig:1441-1465: invalid layout / transpose codes return with C unchanged."""
    c = np.full(4, 5 + 5j)
    one = np.array([1 + 0j])
    for lay, t in ((99, 111), (102, 99)):
        zlib.cblas_zgemm(lay, t, 111, 2, 2, 2, one.ctypes.data, 0, 2, 0, 2, one.ctypes.data, c.ctypes.data, 2)
    assert (c == 5 + 5j).all()
    _ = C
