"""FP32 Level-3 symbols (strsm_, ssyrk_) and the valid paths of cblas_dtrsm /
cblas_dsyrk on the GEMM core.

Bars: the reference's row-major known answer (cblas_compat_test.zig:596-617:
DTRMM gives [16, 19, 18, 21], then DTRSM restores [4, 5, 6, 7]); exact
recovery of integer solutions of unit-diagonal integer systems (every
substitution step and GEMM update is exact, in FP32 too while every partial
sum stays below 2^24); the componentwise residual bound
|op(A) X - alpha B| <= 8 n u (|op(A)||X| + |alpha||B|) on random systems
(u = 2^-24); SYRK bitwise on integer inputs with the other triangle and the
padding untouched, and within 2 (K + 2) u (|alpha||A||A^T| + |beta||C|) on
random inputs.  Row-major CBLAS results are read back as row-major matrices
and compared with the row-major definition directly, so the argument mapping
(cblas.zig:1550-1559,1742-1753) is checked, not restated.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

I = C.c_int32
COMBOS = [(s, u, t, d) for s in "LR" for u in "LU" for t in "NT" for d in "NU"]


def dev(torch, *xs):
    out = tuple(torch.from_numpy(np.ascontiguousarray(x).copy()).cuda() for x in xs)
    return out if len(out) > 1 else out[0]


def seed_of(*parts):
    return sum(ord(ch) * 31**i for i, ch in enumerate("".join(parts)))


def strsm(torch, side, uplo, trans, diag, m, n, alpha, a, lda, b, ldb):
    from zynum_amd import _lib

    lib = _lib.load()
    da, db = dev(torch, a, b)
    lib.strsm_(side.encode(), uplo.encode(), trans.encode(), diag.encode(), C.byref(I(m)), C.byref(I(n)),
               C.byref(C.c_float(alpha)), da.data_ptr(), C.byref(I(lda)), db.data_ptr(), C.byref(I(ldb)))
    assert lib.zynum_gpu_last_status() == 0
    torch.cuda.synchronize()
    return db.cpu().numpy()


def ssyrk(torch, uplo, trans, n, k, alpha, a, lda, beta, c, ldc):
    from zynum_amd import _lib

    lib = _lib.load()
    da, dc = dev(torch, a, c)
    lib.ssyrk_(uplo.encode(), trans.encode(), C.byref(I(n)), C.byref(I(k)), C.byref(C.c_float(alpha)),
               da.data_ptr(), C.byref(I(lda)), C.byref(C.c_float(beta)), dc.data_ptr(), C.byref(I(ldc)))
    assert lib.zynum_gpu_last_status() == 0
    torch.cuda.synchronize()
    return dc.cpu().numpy()


def triangle_storage(T, uplo, lda, dtype, poison=np.nan):
    na = T.shape[0]
    st = np.full((na, lda), -999.0)
    M = T.astype(np.float64).copy()
    other = np.triu_indices(na, 1) if uplo == "L" else np.tril_indices(na, -1)
    M[other] = poison
    st[:, :na] = M.T
    return st.ravel().astype(dtype)


def test_strsm_known_answer(torch_cuda):
    """fortran_compat_test.zig:1779-1798 in FP32: dtrsm_ on the DTRMM result restores [4, 6, 5, 7]."""
    tri = np.array([1, 777, 2, 3], dtype=np.float32)
    b = np.array([16, 18, 19, 21], dtype=np.float32)
    assert strsm(torch_cuda, "L", "U", "N", "N", 2, 2, 1.0, tri, 2, b, 2).tolist() == [4, 6, 5, 7]


@pytest.mark.parametrize("side,uplo,trans,diag", COMBOS)
def test_strsm_integer_exact_recovery(torch_cuda, side, uplo, trans, diag):
    rng = np.random.default_rng(seed_of(side, uplo, trans, diag, "s"))
    m, n = 173, 91
    na = m if side == "L" else n
    T = (np.tril if uplo == "L" else np.triu)(rng.integers(-1, 2, (na, na)), 1 if uplo == "U" else -1) + np.eye(na)
    store_T = T.copy()
    if diag == "U":
        store_T[np.diag_indices(na)] = 5.0  # ignored for a unit diagonal
    opA = T.T if trans == "T" else T
    X = rng.integers(-3, 4, (m, n)).astype(np.float64)
    Bm = opA @ X if side == "L" else X @ opA
    # every partial sum of any substitution / update order is bounded by sum |op(A)||X|
    scale = (np.abs(opA) @ np.abs(X)) if side == "L" else (np.abs(X) @ np.abs(opA))
    assert np.max(scale) < 2**24
    lda, ldb = na + 2, m + 3
    a = triangle_storage(store_T, uplo, lda, np.float32)
    bb = np.full((n, ldb), -777.0, dtype=np.float32)
    bb[:, :m] = Bm.T
    got = strsm(torch_cuda, side, uplo, trans, diag, m, n, 1.0, a, lda, bb.ravel(), ldb).reshape(n, ldb)
    assert np.array_equal(got[:, :m].T, X.astype(np.float32))
    assert np.all(got[:, m:] == -777.0)


@pytest.mark.parametrize("side,uplo,trans,diag", COMBOS)
def test_strsm_residual_random(torch_cuda, side, uplo, trans, diag):
    rng = np.random.default_rng(17)
    m, n = 200, 150
    na = m if side == "L" else n
    A = rng.uniform(-1, 1, (na, na))
    A[np.diag_indices(na)] = 8 + np.abs(rng.uniform(-1, 1, na))
    A = A.astype(np.float32).astype(np.float64)
    T = np.tril(A) if uplo == "L" else np.triu(A)
    Tsolve = T.copy()
    if diag == "U":
        Tsolve[np.diag_indices(na)] = 1.0
    opA = Tsolve.T if trans == "T" else Tsolve
    lda, ldb = na + 1, m + 2
    a = triangle_storage(T, uplo, lda, np.float32)
    B = rng.uniform(-1, 1, (m, n)).astype(np.float32).astype(np.float64)
    bb = np.full((n, ldb), -777.0, dtype=np.float32)
    bb[:, :m] = B.T
    alpha = -1.5
    got = strsm(torch_cuda, side, uplo, trans, diag, m, n, alpha, a, lda, bb.ravel(), ldb).reshape(n, ldb)
    X = got[:, :m].T.astype(np.float64)
    R = (opA @ X if side == "L" else X @ opA) - alpha * B
    scale = (np.abs(opA) @ np.abs(X)) if side == "L" else (np.abs(X) @ np.abs(opA))
    assert np.all(np.abs(R) <= 8 * na * 2.0**-24 * (scale + abs(alpha) * np.abs(B)))
    assert np.all(got[:, m:] == -777.0)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("n,k", [(301, 77), (301, 1100), (1030, 200)])
def test_ssyrk_integer_exact_and_triangle_untouched(torch_cuda, uplo, trans, n, k):
    rng = np.random.default_rng(seed_of(uplo, trans, str(n), str(k)))
    ar, ac = (n, k) if trans == "N" else (k, n)
    lda, ldc = ar + 3, n + 2
    A = rng.integers(-9, 10, (ar, ac)).astype(np.float64)
    aa = np.full((ac, lda), -555.0, dtype=np.float32)
    aa[:, :ar] = A.T
    Cm = rng.integers(-9, 10, (n, n)).astype(np.float64)
    cc = np.full((n, ldc), -777.0, dtype=np.float32)
    cc[:, :n] = Cm.T
    X = A if trans == "N" else A.T
    assert 81 * k < 2**24
    got = ssyrk(torch_cuda, uplo, trans, n, k, 0.75, aa.ravel(), lda, -0.25, cc.ravel(), ldc).reshape(n, ldc)
    want = 0.75 * (X @ X.T) - 0.25 * Cm
    tri = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    GT = got[:, :n].T.astype(np.float64)
    assert np.array_equal(GT[tri], want[tri])
    assert np.array_equal(GT[~tri], Cm[~tri])
    assert np.all(got[:, n:] == -777.0)


def test_ssyrk_known_answer(torch_cuda):
    """fortran_compat_test.zig:1742-1749 in FP32."""
    a = np.array([1, 3, 2, 4], dtype=np.float32)
    c = np.array([0, 777, 0, 0], dtype=np.float32)
    assert ssyrk(torch_cuda, "U", "N", 2, 2, 1.0, a, 2, 0.0, c, 2).tolist() == [5, 777, 11, 25]


@pytest.mark.parametrize("uplo,trans", [("L", "N"), ("U", "T")])
def test_ssyrk_random_within_bound(torch_cuda, uplo, trans):
    rng = np.random.default_rng(23)
    n, k = 520, 700
    ar, ac = (n, k) if trans == "N" else (k, n)
    A = rng.uniform(-1, 1, (ar, ac)).astype(np.float32)
    Cm = rng.uniform(-1, 1, (n, n)).astype(np.float32)
    got = ssyrk(torch_cuda, uplo, trans, n, k, 1.5, A.T.ravel(), ar, 0.5, Cm.T.ravel(), n).reshape(n, n).T
    X = (A if trans == "N" else A.T).astype(np.float64)
    exact = 1.5 * (X @ X.T) + 0.5 * Cm.astype(np.float64)
    bound = 2 * (k + 2) * 2.0**-24 * (1.5 * (np.abs(X) @ np.abs(X).T) + 0.5 * np.abs(Cm))
    tri = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    assert np.all(np.abs(got.astype(np.float64) - exact)[tri] <= bound[tri])
    assert np.array_equal(got[~tri], Cm[~tri])


def test_cblas_row_major_trmm_then_trsm_restores(torch_cuda, zlib):
    """cblas_compat_test.zig:596-617: row-major DTRMM gives [16, 19, 18, 21], DTRSM restores [4, 5, 6, 7]."""
    tri = np.array([1, 2, 777, 3.0])
    b = np.array([4, 5, 6, 7.0])
    dt, db = dev(torch_cuda, tri, b)
    zlib.cblas_dtrmm(101, 141, 121, 111, 131, 2, 2, 1.0, dt.data_ptr(), 2, db.data_ptr(), 2)
    torch_cuda.cuda.synchronize()
    assert db.cpu().numpy().tolist() == [16, 19, 18, 21]
    zlib.cblas_dtrsm(101, 141, 121, 111, 131, 2, 2, 1.0, dt.data_ptr(), 2, db.data_ptr(), 2)
    torch_cuda.cuda.synchronize()
    assert zlib.zynum_gpu_last_status() == 0
    np.testing.assert_allclose(db.cpu().numpy(), [4, 5, 6, 7], atol=1e-12)
    assert db.cpu().numpy().tolist() == [4, 5, 6, 7]


@pytest.mark.parametrize("layout", [101, 102])
@pytest.mark.parametrize("side,uplo,trans,diag", COMBOS)
def test_cblas_dtrsm_integer_exact(torch_cuda, zlib, layout, side, uplo, trans, diag):
    """Every side/uplo/trans/diag in both layouts: the matrices are built in the
    caller's layout and the solution must come back bitwise."""
    rng = np.random.default_rng(seed_of(str(layout), side, uplo, trans, diag))
    m, n = 75, 130
    na = m if side == "L" else n
    T = (np.tril if uplo == "L" else np.triu)(rng.integers(-1, 2, (na, na)), 1 if uplo == "U" else -1) + np.eye(na)
    store_T = T.copy()
    other = np.triu_indices(na, 1) if uplo == "L" else np.tril_indices(na, -1)
    store_T[other] = np.nan
    if diag == "U":
        store_T[np.diag_indices(na)] = 5.0
    opA = T.T if trans == "T" else T
    X = rng.integers(-3, 4, (m, n)).astype(np.float64)
    Bm = opA @ X if side == "L" else X @ opA
    lda, ldb = na + 2, (n if layout == 101 else m) + 3

    def store(M, ld):  # caller's layout with ld padding of -999
        rows, cols = M.shape
        if layout == 101:
            buf = np.full((rows, ld), -999.0)
            buf[:, :cols] = M
        else:
            buf = np.full((cols, ld), -999.0)
            buf[:, :rows] = M.T
        return buf.ravel()

    sd = 141 if side == "L" else 142
    ul = 122 if uplo == "L" else 121
    tr = 111 if trans == "N" else 112
    dg = 131 if diag == "N" else 132
    da, db = dev(torch_cuda, store(store_T, lda), store(Bm, ldb))
    zlib.cblas_dtrsm(layout, sd, ul, tr, dg, m, n, 1.0, da.data_ptr(), lda, db.data_ptr(), ldb)
    torch_cuda.cuda.synchronize()
    assert zlib.zynum_gpu_last_status() == 0
    assert np.array_equal(db.cpu().numpy(), store(X, ldb))


@pytest.mark.parametrize("layout", [101, 102])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_cblas_dsyrk_integer_exact(torch_cuda, zlib, layout, uplo, trans):
    """C := alpha op(A) op(A)^T + beta C on the stored triangle of the caller's layout, other triangle untouched."""
    rng = np.random.default_rng(seed_of(str(layout), uplo, trans, "syrk"))
    n, k = 141, 66
    A = rng.integers(-9, 10, (n, k) if trans == "N" else (k, n)).astype(np.float64)
    Cm = rng.integers(-9, 10, (n, n)).astype(np.float64)
    X = A if trans == "N" else A.T
    want = 2.0 * (X @ X.T) - 1.0 * Cm
    ar, ac = A.shape
    lda = (ac if layout == 101 else ar) + 1
    ldc = n + 4

    def store(M, ld):
        rows, cols = M.shape
        if layout == 101:
            buf = np.full((rows, ld), -999.0)
            buf[:, :cols] = M
        else:
            buf = np.full((cols, ld), -999.0)
            buf[:, :rows] = M.T
        return buf

    da, dc = dev(torch_cuda, store(A, lda).ravel(), store(Cm, ldc).ravel())
    zlib.cblas_dsyrk(layout, 122 if uplo == "L" else 121, 111 if trans == "N" else 112, n, k, 2.0, da.data_ptr(),
                     lda, -1.0, dc.data_ptr(), ldc)
    torch_cuda.cuda.synchronize()
    assert zlib.zynum_gpu_last_status() == 0
    buf = dc.cpu().numpy().reshape(-1, ldc)
    G = buf[:, :n] if layout == 101 else buf[:, :n].T
    tri = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    assert np.array_equal(G[tri], want[tri])
    assert np.array_equal(G[~tri], Cm[~tri])
    assert np.all(buf[:, n:] == -999.0)
