"""GPU parity of DSYMM, DTRMM and DSYR2K on the GEMM core (and their FP32 forms).

Bars: bitwise on integer-valued inputs (every summation order is exact there)
for every side / uplo / trans / diag form, with padding and the untouched
triangle checked bit for bit; the reference's known answers; on random inputs
the componentwise bound |C_gpu - C_oracle| <= 2 (K + 2) u (|alpha| |X| |Y| + |beta| |C0|)
with K the inner dimension (u = 2^-53).
"""
import ctypes as C

import numpy as np
import pytest

from conftest import fill_padded
from test_level3_more_cpu import mat, sym_full, tri_full

pytestmark = pytest.mark.gpu
I = C.c_int32
U64 = 2.0 ** -53


def dev(torch, *xs):
    return [torch.from_numpy(x.copy()).cuda() for x in xs]


def run_symm(torch, side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc, single=False):
    from zynum_amd import _lib

    lib = _lib.load()
    da, db, dc = dev(torch, a, b, c)
    S = C.c_float if single else C.c_double
    fn = lib.ssymm_ if single else lib.dsymm_
    fn(side.encode(), uplo.encode(), C.byref(I(m)), C.byref(I(n)), C.byref(S(alpha)), da.data_ptr(), C.byref(I(lda)),
       db.data_ptr(), C.byref(I(ldb)), C.byref(S(beta)), dc.data_ptr(), C.byref(I(ldc)))
    assert lib.zynum_gpu_last_status() == 0
    torch.cuda.synchronize()
    return dc.cpu().numpy()


def run_trmm(torch, side, uplo, trans, diag, m, n, alpha, a, lda, b, ldb, single=False):
    from zynum_amd import _lib

    lib = _lib.load()
    da, db = dev(torch, a, b)
    S = C.c_float if single else C.c_double
    fn = lib.strmm_ if single else lib.dtrmm_
    fn(side.encode(), uplo.encode(), trans.encode(), diag.encode(), C.byref(I(m)), C.byref(I(n)), C.byref(S(alpha)),
       da.data_ptr(), C.byref(I(lda)), db.data_ptr(), C.byref(I(ldb)))
    assert lib.zynum_gpu_last_status() == 0
    torch.cuda.synchronize()
    return db.cpu().numpy()


def run_syr2k(torch, uplo, trans, n, k, alpha, a, lda, b, ldb, beta, c, ldc, single=False):
    from zynum_amd import _lib

    lib = _lib.load()
    da, db, dc = dev(torch, a, b, c)
    S = C.c_float if single else C.c_double
    fn = lib.ssyr2k_ if single else lib.dsyr2k_
    fn(uplo.encode(), trans.encode(), C.byref(I(n)), C.byref(I(k)), C.byref(S(alpha)), da.data_ptr(), C.byref(I(lda)),
       db.data_ptr(), C.byref(I(ldb)), C.byref(S(beta)), dc.data_ptr(), C.byref(I(ldc)))
    assert lib.zynum_gpu_last_status() == 0
    torch.cuda.synchronize()
    return dc.cpu().numpy()


def test_known_answers(torch_cuda):
    """fortran_compat_test.zig:1676-1798."""
    a = np.array([1, 777, 2, 3.0])
    b = np.array([4, 6, 5, 7.0])
    assert run_symm(torch_cuda, "L", "U", 2, 2, 1.0, a, 2, b, 2, 0.0, np.zeros(4), 2).tolist() == [16, 26, 19, 31]
    ra, rb = np.array([1, 3, 2, 4.0]), np.array([5, 7, 6, 8.0])
    got = run_syr2k(torch_cuda, "U", "N", 2, 2, 1.0, ra, 2, rb, 2, 0.0, np.array([0, 777, 0, 0.0]), 2)
    assert got.tolist() == [34, 777, 62, 106]
    assert run_trmm(torch_cuda, "L", "U", "N", "N", 2, 2, 1.0, a, 2, b, 2).tolist() == [16, 18, 19, 21]


def test_cblas_row_major_known_answers(torch_cuda, zlib):
    """cblas_compat_test.zig:595-640: row-major DTRMM and DSYR2K."""
    tri = np.array([1, 2, 777, 3.0])
    b = np.array([4, 5, 6, 7.0])
    dt, db = dev(torch_cuda, tri, b)
    zlib.cblas_dtrmm(101, 141, 121, 111, 131, 2, 2, 1.0, dt.data_ptr(), 2, db.data_ptr(), 2)
    torch_cuda.cuda.synchronize()
    assert db.cpu().numpy().tolist() == [16, 19, 18, 21]
    a = np.array([1, 2, 3, 4.0])
    bb = np.array([5, 6, 7, 8.0])
    c = np.array([0, 0, 777, 0.0])
    da, dbb, dc = dev(torch_cuda, a, bb, c)
    zlib.cblas_dsyr2k(101, 121, 111, 2, 2, 1.0, da.data_ptr(), 2, dbb.data_ptr(), 2, 0.0, dc.data_ptr(), 2)
    torch_cuda.cuda.synchronize()
    assert dc.cpu().numpy().tolist() == [34, 62, 777, 106]


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("side", "LR")
@pytest.mark.parametrize("uplo", "UL")
@pytest.mark.parametrize("shape", [(9, 7), (200, 130), (131, 257)])
def test_symm_integer_exact(torch_cuda, oracle, single, side, uplo, shape):
    m, n = shape
    na = m if side == "L" else n
    lda, ldb, ldc = na + 3, m + 2, m + 1
    dt = np.float32 if single else np.float64
    a = fill_padded(na, na, lda, 5, -999.0, dtype=dt, integer=True)
    A = a[: lda * na].reshape(na, lda)
    # A[j, i] is element (i, j); the unstored triangle (i < j for L, i > j for U) is never read
    mask = np.tril(np.ones((na, na), bool), -1) if uplo == "L" else np.triu(np.ones((na, na), bool), 1)
    A[:, :na][mask] = 777.0
    b = fill_padded(m, n, ldb, 7, -999.0, dtype=dt, integer=True)
    c = fill_padded(m, n, ldc, 9, -777.0, dtype=dt, integer=True)
    S = sym_full(a.astype(np.float64), na, lda, uplo)
    B, C0 = mat(b, m, n, ldb).astype(np.float64), mat(c, m, n, ldc).astype(np.float64)
    want = 2.0 * (S @ B if side == "L" else B @ S) - 3.0 * C0
    got = run_symm(torch_cuda, side, uplo, m, n, 2.0, a, lda, b, ldb, -3.0, c, ldc, single)
    assert np.array_equal(mat(got, m, n, ldc).astype(np.float64), want)
    pad = np.ones(ldc * n, dtype=bool)
    pad.reshape(n, ldc)[:, :m] = False
    assert np.array_equal(got[pad], c[pad])
    if not single:
        ref = c.copy()
        oracle.dsymm(side, uplo, m, n, 2.0, a, lda, b, ldb, -3.0, ref, ldc)
        assert np.array_equal(got, ref)


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("side", "LR")
@pytest.mark.parametrize("uplo", "UL")
@pytest.mark.parametrize("trans", "NT")
@pytest.mark.parametrize("diag", "NU")
@pytest.mark.parametrize("shape", [(7, 9), (300, 130), (129, 700)])
def test_trmm_integer_exact(torch_cuda, oracle, single, side, uplo, trans, diag, shape):
    """Every form; the 300- and 700-wide triangles span several K-range panels."""
    m, n = shape
    na = m if side == "L" else n
    lda, ldb = na + 2, m + 3
    dt = np.float32 if single else np.float64
    a = fill_padded(na, na, lda, 17, -999.0, dtype=dt, integer=True)
    b = fill_padded(m, n, ldb, 19, -999.0, dtype=dt, integer=True)
    if single:  # keep FP32 partial sums exact
        a = np.where(np.abs(a) > 3, np.sign(a) * 3, a).astype(dt)
    T = tri_full(a.astype(np.float64), na, lda, uplo, diag)
    opT = T.T if trans == "T" else T
    B = mat(b, m, n, ldb).astype(np.float64)
    want = -2.0 * (opT @ B if side == "L" else B @ opT)
    got = run_trmm(torch_cuda, side, uplo, trans, diag, m, n, -2.0, a, lda, b, ldb, single)
    assert np.array_equal(mat(got, m, n, ldb).astype(np.float64), want)
    pad = np.ones(ldb * n, dtype=bool)
    pad.reshape(n, ldb)[:, :m] = False
    assert np.array_equal(got[pad], b[pad])


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("uplo", "UL")
@pytest.mark.parametrize("trans", "NT")
@pytest.mark.parametrize("shape", [(8, 6), (300, 77), (129, 260), (300, 1100)])
def test_syr2k_integer_exact(torch_cuda, oracle, single, uplo, trans, shape):
    n, k = shape
    rows, cols = (n, k) if trans == "N" else (k, n)
    lda, ldb, ldc = rows + 2, rows + 3, n + 1
    dt = np.float32 if single else np.float64
    a = fill_padded(rows, cols, lda, 3, -999.0, dtype=dt, integer=True)
    b = fill_padded(rows, cols, ldb, 11, -999.0, dtype=dt, integer=True)
    c = fill_padded(n, n, ldc, 13, -777.0, dtype=dt, integer=True)
    A, B = mat(a, rows, cols, lda).astype(np.float64), mat(b, rows, cols, ldb).astype(np.float64)
    if trans == "T":
        A, B = A.T, B.T
    full = 2.0 * (A @ B.T + B @ A.T) + 0.5 * mat(c, n, n, ldc)
    got = run_syr2k(torch_cuda, uplo, trans, n, k, 2.0, a, lda, b, ldb, 0.5, c, ldc, single)
    G, C0 = mat(got, n, n, ldc), mat(c, n, n, ldc)
    stored = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    assert np.array_equal(G[stored].astype(np.float64), full[stored])
    assert np.array_equal(G[~stored], C0[~stored])


@pytest.mark.parametrize("uplo", "UL")
@pytest.mark.parametrize("trans", "NT")
def test_dsyr2k_direct_load_tiles_integer_exact(torch_cuda, uplo, trans):
    """Full 128x128 triangle tiles, k a multiple of 8 and even leading
    dimensions: both passes take the direct-to-LDS tile; exact on integers."""
    n, k = 256, 136
    rows, cols = (n, k) if trans == "N" else (k, n)
    lda, ldb, ldc = rows + 2, rows + 4, n + 1
    a = fill_padded(rows, cols, lda, 5, -999.0, dtype=np.float64, integer=True)
    b = fill_padded(rows, cols, ldb, 7, -999.0, dtype=np.float64, integer=True)
    c = fill_padded(n, n, ldc, 9, -777.0, dtype=np.float64, integer=True)
    A, B = mat(a, rows, cols, lda), mat(b, rows, cols, ldb)
    if trans == "T":
        A, B = A.T, B.T
    full = 2.0 * (A @ B.T + B @ A.T) + 0.5 * mat(c, n, n, ldc)
    got = run_syr2k(torch_cuda, uplo, trans, n, k, 2.0, a, lda, b, ldb, 0.5, c, ldc, False)
    G, C0 = mat(got, n, n, ldc), mat(c, n, n, ldc)
    stored = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    assert np.array_equal(G[stored], full[stored])
    assert np.array_equal(G[~stored], C0[~stored])


@pytest.mark.parametrize("side", "LR")
def test_symm_random_bound(torch_cuda, oracle, side):
    m, n = 257, 190
    na = m if side == "L" else n
    rng = np.random.default_rng(3)
    a, b, c = rng.uniform(-1, 1, na * na), rng.uniform(-1, 1, m * n), rng.uniform(-1, 1, m * n)
    ref = c.copy()
    oracle.dsymm(side, "L", m, n, 0.7, a, na, b, m, 0.3, ref, m)
    got = run_symm(torch_cuda, side, "L", m, n, 0.7, a, na, b, m, 0.3, c, m)
    S = np.abs(sym_full(a, na, na, "L"))
    B = np.abs(mat(b, m, n, m))
    mag = 0.7 * (S @ B if side == "L" else B @ S) + 0.3 * np.abs(mat(c, m, n, m))
    assert (np.abs(mat(got, m, n, m) - mat(ref, m, n, m)) <= 2 * (na + 2) * U64 * mag).all()


@pytest.mark.parametrize("side,uplo,trans", [("L", "L", "N"), ("L", "U", "T"), ("R", "U", "N"), ("R", "L", "T")])
def test_trmm_random_bound(torch_cuda, oracle, side, uplo, trans):
    m, n = 300, 260
    na = m if side == "L" else n
    rng = np.random.default_rng(4)
    a, b = rng.uniform(-1, 1, na * na), rng.uniform(-1, 1, m * n)
    ref = b.copy()
    oracle.dtrmm(side, uplo, trans, "N", m, n, 1.5, a, na, ref, m)
    got = run_trmm(torch_cuda, side, uplo, trans, "N", m, n, 1.5, a, na, b, m)
    T = np.abs(tri_full(a, na, na, uplo, "N"))
    T = T.T if trans == "T" else T
    B = np.abs(mat(b, m, n, m))
    mag = 1.5 * (T @ B if side == "L" else B @ T)
    assert (np.abs(mat(got, m, n, m) - mat(ref, m, n, m)) <= 2 * (na + 2) * U64 * mag).all()


@pytest.mark.parametrize("uplo,trans", [("L", "N"), ("U", "T")])
def test_syr2k_random_bound(torch_cuda, oracle, uplo, trans):
    n, k = 270, 150
    rows, cols = (n, k) if trans == "N" else (k, n)
    rng = np.random.default_rng(5)
    a, b, c = rng.uniform(-1, 1, rows * cols), rng.uniform(-1, 1, rows * cols), rng.uniform(-1, 1, n * n)
    ref = c.copy()
    oracle.dsyr2k(uplo, trans, n, k, 0.9, a, rows, b, rows, -0.4, ref, n)
    got = run_syr2k(torch_cuda, uplo, trans, n, k, 0.9, a, rows, b, rows, -0.4, c, n)
    A, B = np.abs(mat(a, rows, cols, rows)), np.abs(mat(b, rows, cols, rows))
    if trans == "T":
        A, B = A.T, B.T
    mag = 0.9 * (A @ B.T + B @ A.T) + 0.4 * np.abs(mat(c, n, n, n))
    stored = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    err = np.abs(mat(got, n, n, n) - mat(ref, n, n, n))
    assert (err[stored] <= (2 * (2 * k + 2) * U64 * mag)[stored]).all()
    assert np.array_equal(mat(got, n, n, n)[~stored], mat(c, n, n, n)[~stored])


def test_alpha_zero_and_k_zero(torch_cuda):
    """alpha == 0: SYMM gives beta C, TRMM gives zeros, SYR2K gives beta C on
    the stored triangle (k == 0 too); beta == 0 writes exact zeros."""
    a = np.ones(9)
    b = np.arange(1.0, 10.0)
    c = np.arange(10.0, 19.0)
    got = run_symm(torch_cuda, "L", "U", 3, 3, 0.0, a, 3, b, 3, 2.0, c, 3)
    assert np.array_equal(got, 2.0 * c)
    got = run_trmm(torch_cuda, "R", "L", "N", "N", 3, 3, 0.0, a, 3, b, 3)
    assert np.array_equal(got, np.zeros(9))
    got = run_syr2k(torch_cuda, "L", "N", 3, 0, 1.0, a, 3, a, 3, 0.0, c, 3)
    assert np.array_equal(mat(got, 3, 3, 3)[np.tril(np.ones((3, 3), bool))], np.zeros(6))
    assert np.array_equal(mat(got, 3, 3, 3)[~np.tril(np.ones((3, 3), bool))],
                          mat(c, 3, 3, 3)[~np.tril(np.ones((3, 3), bool))])


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("side", ["L", "R"])
def test_symm_alpha_zero_forms_the_product(torch_cuda, oracle, single, side):
    """symmetric.zig:56-70 has no alpha == 0 quick return: 0 * sum is formed, so a
    NaN or an infinity in the stored triangle of A or in B reaches C, and finite
    entries give beta C (exact 0 for beta == 0, C not read).  Same as the oracle."""
    dt = np.float32 if single else np.float64
    m, n = 37, 29
    na = m if side == "L" else n
    rng = np.random.default_rng(3)
    a = rng.uniform(-1, 1, na * na).astype(dt)
    a[5 + 7 * na] = np.nan   # stored (lower) triangle, row 5... column 7 lies above the diagonal: unstored
    a[9 + 2 * na] = np.nan   # row 9, column 2: in the lower triangle
    b = rng.uniform(-1, 1, m * n).astype(dt)
    b[3 + 4 * m] = np.inf
    for beta in (0.0, 2.0):
        c = rng.uniform(-1, 1, m * n).astype(dt)
        c[0] = np.nan  # read only when beta != 0
        got = run_symm(torch_cuda, side, "L", m, n, 0.0, a, na, b, m, beta, c, m, single=single)
        want = c.astype(np.float64).copy()
        oracle.dsymm(side, "L", m, n, 0.0, a.astype(np.float64), na, b.astype(np.float64), m, beta, want, m)
        assert np.array_equal(np.isnan(got), np.isnan(want))
        fin = ~np.isnan(want)
        assert np.array_equal(got[fin].astype(np.float64), want[fin])
        assert np.isnan(got).sum() > 0


def test_host_operands_are_staged(zlib):
    """Host (numpy) operands run through the staging path of the same kernels."""
    I_ = C.c_int32
    a = np.array([1, 777, 2, 3.0])
    b = np.array([4, 6, 5, 7.0])
    zlib.dtrmm_(b"L", b"U", b"N", b"N", C.byref(I_(2)), C.byref(I_(2)), C.byref(C.c_double(1)), a.ctypes.data,
                C.byref(I_(2)), b.ctypes.data, C.byref(I_(2)))
    assert zlib.zynum_gpu_last_status() == 0
    assert b.tolist() == [16, 18, 19, 21]
    c = np.zeros(4)
    zlib.dsymm_(b"L", b"U", C.byref(I_(2)), C.byref(I_(2)), C.byref(C.c_double(1)), a.ctypes.data, C.byref(I_(2)),
                np.array([4, 6, 5, 7.0]).ctypes.data, C.byref(I_(2)), C.byref(C.c_double(0)), c.ctypes.data,
                C.byref(I_(2)))
    assert c.tolist() == [16, 26, 19, 31]
