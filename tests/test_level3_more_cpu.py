"""CPU checks of DSYMM, DTRMM and DSYR2K: the oracle pinned against the
reference's known answers, exact arithmetic, closed forms and the rounding
rules it restates, and the ABI's argument checks without a device."""
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from conftest import fill_padded

ROOT = Path(__file__).resolve().parent.parent


def mat(x, rows, cols, ld):
    return x[: ld * cols].reshape(cols, ld)[:, :rows].T.copy()


def sym_full(a, n, lda, uplo):
    A = mat(a, n, n, lda)
    low = np.tril(A) if uplo == "L" else np.triu(A)
    return low + low.T - np.diag(np.diag(A))


def tri_full(a, n, lda, uplo, diag):
    A = mat(a, n, n, lda)
    T = np.tril(A) if uplo == "L" else np.triu(A)
    if diag == "U":
        np.fill_diagonal(T, 1.0)
    return T


# ---------------------------------------------------------------------------
# known answers (fortran_compat_test.zig:1676-1798, cblas_compat_test.zig:595-640)
# ---------------------------------------------------------------------------
def test_known_answers(oracle):
    a = np.array([1, 777, 2, 3.0])
    b = np.array([4, 6, 5, 7.0])
    c = np.zeros(4)
    oracle.dsymm("L", "U", 2, 2, 1.0, a, 2, b, 2, 0.0, c, 2)
    assert c.tolist() == [16, 26, 19, 31]
    ra, rb = np.array([1, 3, 2, 4.0]), np.array([5, 7, 6, 8.0])
    rc = np.array([0, 777, 0, 0.0])
    oracle.dsyr2k("U", "N", 2, 2, 1.0, ra, 2, rb, 2, 0.0, rc, 2)
    assert rc.tolist() == [34, 777, 62, 106]
    tb = b.copy()
    oracle.dtrmm("L", "U", "N", "N", 2, 2, 1.0, a, 2, tb, 2)
    assert tb.tolist() == [16, 18, 19, 21]


# ---------------------------------------------------------------------------
# exact arithmetic on integer inputs, every form
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("side", "LR")
@pytest.mark.parametrize("uplo", "UL")
def test_symm_integer_exact(oracle, side, uplo):
    m, n = 9, 7
    na = m if side == "L" else n
    lda, ldb, ldc = na + 3, m + 2, m + 1
    a = fill_padded(na, na, lda, 5, -999.0, integer=True)
    # the unstored triangle holds a sentinel that must never be read
    A = a[: lda * na].reshape(na, lda)
    for j in range(na):
        for i in range(na):
            if (uplo == "U" and i > j) or (uplo == "L" and i < j):
                A[j, i] = 777.0
    b = fill_padded(m, n, ldb, 7, -999.0, integer=True)
    c = fill_padded(m, n, ldc, 9, -777.0, integer=True)
    S = sym_full(a, na, lda, uplo)
    B, C0 = mat(b, m, n, ldb), mat(c, m, n, ldc)
    want = 2.0 * (S @ B if side == "L" else B @ S) - 3.0 * C0
    got = c.copy()
    oracle.dsymm(side, uplo, m, n, 2.0, a, lda, b, ldb, -3.0, got, ldc)
    assert np.array_equal(mat(got, m, n, ldc), want)
    pad = np.ones(ldc * n, dtype=bool)
    pad.reshape(n, ldc)[:, :m] = False
    assert np.array_equal(got[pad], c[pad])


@pytest.mark.parametrize("uplo", "UL")
@pytest.mark.parametrize("trans", "NT")
def test_syr2k_integer_exact(oracle, uplo, trans):
    n, k = 8, 6
    rows, cols = (n, k) if trans == "N" else (k, n)
    lda, ldb, ldc = rows + 2, rows + 3, n + 1
    a = fill_padded(rows, cols, lda, 3, -999.0, integer=True)
    b = fill_padded(rows, cols, ldb, 11, -999.0, integer=True)
    c = fill_padded(n, n, ldc, 13, -777.0, integer=True)
    A = mat(a, rows, cols, lda)
    B = mat(b, rows, cols, ldb)
    if trans == "T":
        A, B = A.T, B.T
    full = 2.0 * (A @ B.T + B @ A.T) + 0.5 * mat(c, n, n, ldc)
    got = c.copy()
    oracle.dsyr2k(uplo, trans, n, k, 2.0, a, lda, b, ldb, 0.5, got, ldc)
    G = mat(got, n, n, ldc)
    C0 = mat(c, n, n, ldc)
    stored = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    assert np.array_equal(G[stored], full[stored])
    assert np.array_equal(G[~stored], C0[~stored])


@pytest.mark.parametrize("side", "LR")
@pytest.mark.parametrize("uplo", "UL")
@pytest.mark.parametrize("trans", "NT")
@pytest.mark.parametrize("diag", "NU")
def test_trmm_integer_exact(oracle, side, uplo, trans, diag):
    m, n = 7, 9
    na = m if side == "L" else n
    lda, ldb = na + 2, m + 3
    a = fill_padded(na, na, lda, 17, -999.0, integer=True)
    b = fill_padded(m, n, ldb, 19, -999.0, integer=True)
    T = tri_full(a, na, lda, uplo, diag)
    opT = T.T if trans == "T" else T
    B = mat(b, m, n, ldb)
    want = -2.0 * (opT @ B if side == "L" else B @ opT)
    got = b.copy()
    oracle.dtrmm(side, uplo, trans, diag, m, n, -2.0, a, lda, got, ldb)
    assert np.array_equal(mat(got, m, n, ldb), want)
    pad = np.ones(ldb * n, dtype=bool)
    pad.reshape(n, ldb)[:, :m] = False
    assert np.array_equal(got[pad], b[pad])


def test_trmm_closed_forms(oracle):
    """Identity triangle: B = alpha B; diagonal triangle: rows (L) or columns
    (R) scaled by the diagonal."""
    n = 6
    rng = np.random.default_rng(1)
    b = rng.standard_normal(n * n)
    eye = np.eye(n).T.ravel().copy()
    got = b.copy()
    oracle.dtrmm("L", "U", "N", "N", n, n, 0.5, eye, n, got, n)
    assert np.array_equal(got, 0.5 * b)
    d = np.diag(np.arange(1.0, n + 1)).T.ravel().copy()
    got = b.copy()
    oracle.dtrmm("R", "L", "T", "N", n, n, 1.0, d, n, got, n)
    assert np.array_equal(mat(got, n, n, n), mat(b, n, n, n) * np.arange(1.0, n + 1)[None, :])
    got = b.copy()
    oracle.dtrmm("L", "L", "N", "U", n, n, 1.0, d, n, got, n)  # unit diagonal ignores d
    assert np.array_equal(got, b)


# ---------------------------------------------------------------------------
# rounding rules the oracle restates
# ---------------------------------------------------------------------------
def test_symm_products_not_fused(oracle):
    """symmetric.zig:62: sum = add(sum, mul(S(i,p), B(p,j))): with
    S = [[-1, x], [x, 0]], B = [1, x] and x = 1 + 2^-30, C(0,0) is
    -1 + RN(x^2) = 2^-29; a fused step would keep the 2^-60 term."""
    x = 1 + 2.0 ** -30
    a = np.array([-1.0, x, x, 0.0])
    b = np.array([1.0, x])
    c = np.zeros(2)
    oracle.dsymm("L", "U", 2, 1, 1.0, a, 2, b, 2, 0.0, c, 2)
    assert c[0] == 2.0 ** -29


def test_syr2k_pair_grouping(oracle):
    """symmetric.zig:320: sum = sum + (a_i b_j + b_i a_j): on the diagonal with
    p0 giving 1 and p1 giving two terms of 2^-53, the pair is added first
    (1 + 2^-52); adding the terms one by one would round back to 1."""
    a = np.array([1.0, 2.0 ** -26])  # n = 1, k = 2, trans N: a(0, p)
    b = np.array([0.5, 2.0 ** -27])
    c = np.zeros(1)
    oracle.dsyr2k("L", "N", 1, 2, 1.0, a, 1, b, 1, 0.0, c, 1)
    assert c[0] == 1 + 2.0 ** -52


def test_symm_no_alpha_quick_return(oracle):
    """The reference forms alpha * sum even for alpha == 0 (no quick return),
    so a NaN in A still reaches C; beta == 0 never reads C."""
    a = np.array([np.nan, 1.0, 1.0, 1.0])
    b = np.ones(2)
    c = np.array([np.nan, np.nan])
    oracle.dsymm("L", "L", 2, 1, 0.0, a, 2, b, 2, 0.0, c, 2)
    assert np.isnan(c[0]) and c[1] == 0.0


# ---------------------------------------------------------------------------
# ABI checks without a device
# ---------------------------------------------------------------------------
def _child(code: str) -> subprocess.CompletedProcess:
    return subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=str(ROOT),
                          env={"PYTHONPATH": str(ROOT), "HIP_VISIBLE_DEVICES": "", "PATH": "/usr/bin:/bin"})


@pytest.mark.parametrize("args,expected", [
    (("X", "U", 2, 2, 2, 2, 2), 1), (("L", "X", 2, 2, 2, 2, 2), 2), (("L", "U", -1, 2, 2, 2, 2), 3),
    (("L", "U", 2, -1, 2, 2, 2), 4), (("L", "U", 3, 2, 2, 3, 3), 7), (("R", "U", 2, 3, 2, 2, 2), 7),
    (("L", "U", 3, 2, 3, 2, 3), 9), (("L", "U", 3, 2, 3, 3, 2), 12),
])
def test_dsymm_parameter_numbers(args, expected):
    """fortran.zig:239-250 through dsymm_; C untouched (fortran_compat_test.zig:1184-1203)."""
    side, uplo, m, n, lda, ldb, ldc = args
    code = f"""
import ctypes as C, numpy as np
from zynum_amd import _lib
lib = _lib.load(); I = C.c_int32
a = np.ones(16); b = np.ones(16); c = np.full(16, 9.0)
lib.dsymm_(b"{side}", b"{uplo}", C.byref(I({m})), C.byref(I({n})), C.byref(C.c_double(1)), a.ctypes.data,
           C.byref(I({lda})), b.ctypes.data, C.byref(I({ldb})), C.byref(C.c_double(0)), c.ctypes.data, C.byref(I({ldc})))
assert (c == 9.0).all()
"""
    r = _child(code)
    assert r.returncode == 0, r.stderr
    assert f"On entry to DSYMM parameter number {expected} had an illegal value" in r.stderr


@pytest.mark.parametrize("args,expected", [
    (("X", "N", 2, 2, 2, 2, 2), 1), (("U", "X", 2, 2, 2, 2, 2), 2), (("U", "N", -1, 2, 2, 2, 2), 3),
    (("U", "N", 2, -1, 2, 2, 2), 4), (("U", "N", 3, 2, 2, 3, 3), 7), (("U", "T", 2, 3, 2, 3, 2), 7),
    (("U", "N", 2, 2, 2, 1, 2), 9), (("U", "N", 3, 2, 3, 3, 2), 12),
])
def test_dsyr2k_parameter_numbers(args, expected):
    """fortran.zig:264-275 through dsyr2k_ (fortran_compat_test.zig:1208-1214: ldb = 1 gives 9)."""
    uplo, trans, n, k, lda, ldb, ldc = args
    code = f"""
import ctypes as C, numpy as np
from zynum_amd import _lib
lib = _lib.load(); I = C.c_int32
a = np.ones(16); b = np.ones(16); c = np.full(16, 17.0)
lib.dsyr2k_(b"{uplo}", b"{trans}", C.byref(I({n})), C.byref(I({k})), C.byref(C.c_double(1)), a.ctypes.data,
            C.byref(I({lda})), b.ctypes.data, C.byref(I({ldb})), C.byref(C.c_double(0)), c.ctypes.data,
            C.byref(I({ldc})))
assert (c == 17.0).all()
"""
    r = _child(code)
    assert r.returncode == 0, r.stderr
    assert f"On entry to DSYR2K parameter number {expected} had an illegal value" in r.stderr


def test_dtrmm_invalid_side_and_cblas_enums():
    """fortran_compat_test.zig:1121-1135 (side 'X' gives 1, B untouched) and
    cblas_compat_test.zig:57-60,106-122 (invalid enums return silently)."""
    code = """
import ctypes as C, numpy as np
from zynum_amd import _lib
lib = _lib.load(); I = C.c_int32
tri = np.array([1, 0, 2, 3.0]); b = np.array([4, 5, 6, 7.0]); orig = b.copy()
lib.dtrmm_(b"X", b"U", b"N", b"N", C.byref(I(2)), C.byref(I(2)), C.byref(C.c_double(1)), tri.ctypes.data,
           C.byref(I(2)), b.ctypes.data, C.byref(I(2)))
assert (b == orig).all()
lib.cblas_dtrmm(102, 141, 121, 999, 131, 2, 2, 1.0, tri.ctypes.data, 2, b.ctypes.data, 2)
assert (b == orig).all()
c = np.array([17, 18, 19, 20.0]); oc = c.copy()
lib.cblas_dsyr2k(99, 121, 111, 2, 2, 1.0, tri.ctypes.data, 2, b.ctypes.data, 2, 0.0, c.ctypes.data, 2)
lib.cblas_dsymm(99, 141, 121, 2, 2, 1.0, tri.ctypes.data, 2, b.ctypes.data, 2, 0.0, c.ctypes.data, 2)
assert (c == oc).all()
"""
    r = _child(code)
    assert r.returncode == 0, r.stderr
    assert "On entry to DTRMM parameter number 1 had an illegal value" in r.stderr
    assert r.stderr.count("illegal value") == 1


def test_level3_more_decline_without_device():
    code = """
import ctypes as C, numpy as np
from zynum_amd import _lib
lib = _lib.load()
a = np.ones(4); b = np.full(4, 3.0); c = np.full(4, 7.0)
assert lib.zynum_gpu_dsymm(b"L", b"U", 2, 2, 1.0, a.ctypes.data, 2, b.ctypes.data, 2, 0.0, c.ctypes.data, 2, None, 0) == _lib.DECLINE_NO_DEVICE
assert lib.zynum_gpu_dtrmm(b"L", b"U", b"N", b"N", 2, 2, 1.0, a.ctypes.data, 2, b.ctypes.data, 2, None, 0) == _lib.DECLINE_NO_DEVICE
assert lib.zynum_gpu_dsyr2k(b"U", b"N", 2, 2, 1.0, a.ctypes.data, 2, b.ctypes.data, 2, 0.0, c.ctypes.data, 2, None, 0) == _lib.DECLINE_NO_DEVICE
assert lib.zynum_gpu_dsymm(b"L", b"U", 2, 2, 1.0, a.ctypes.data, 1, b.ctypes.data, 2, 0.0, c.ctypes.data, 2, None, 0) == _lib.DECLINE_ARGUMENT
assert (b == 3.0).all() and (c == 7.0).all()
"""
    r = _child(code)
    assert r.returncode == 0, r.stderr
