"""DTRSM / DSYRK on the GEMM core against the oracle and exact arithmetic.

Bars: the reference's known answers; exact recovery (bitwise) of integer
solutions of unit-triangular integer systems, where every step of any blocking
is exact; residual |op(A) X - alpha B| <= c n u |op(A)||X| on random systems;
SYRK bitwise on integer inputs with the unstored triangle untouched.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def dev(torch, x):
    return torch.from_numpy(np.ascontiguousarray(x).copy()).cuda()


def trsm(torch, side, uplo, trans, diag, m, n, alpha, a, lda, b, ldb, mode=0, param=0, via="request"):
    from zynum_amd import _lib

    lib = _lib.load()
    da, db = dev(torch, a), dev(torch, b)
    if via == "fortran":
        I = C.c_int32
        lib.dtrsm_(side.encode(), uplo.encode(), trans.encode(), diag.encode(), C.byref(I(m)), C.byref(I(n)),
                   C.byref(C.c_double(alpha)), da.data_ptr(), C.byref(I(lda)), db.data_ptr(), C.byref(I(ldb)))
        assert lib.zynum_gpu_last_status() == 0
    else:
        st = lib.zynum_gpu_dtrsm(side.encode(), uplo.encode(), trans.encode(), diag.encode(), m, n, alpha,
                                 da.data_ptr(), lda, db.data_ptr(), ldb, None, mode, param, 0)
        _lib.check(st, "dtrsm")
    torch.cuda.synchronize()
    return db.cpu().numpy()


def test_dtrsm_known_answer(torch_cuda):
    # fortran_compat_test.zig:1779-1798
    tri = np.array([1, 777, 2, 3.0])
    b = np.array([16, 18, 19, 21.0])
    got = trsm(torch_cuda, "L", "U", "N", "N", 2, 2, 1.0, tri, 2, b, 2, via="fortran")
    np.testing.assert_allclose(got, [4, 6, 5, 7], atol=1e-12)


def triangle_storage(T, uplo, lda, poison=np.nan):
    """Column-major storage of T's stored triangle; the other triangle holds poison."""
    na = T.shape[0]
    st = np.full((na, lda), -999.0)
    M = T.copy()
    other = np.triu_indices(na, 1) if uplo == "L" else np.tril_indices(na, -1)
    M[other] = poison
    st[:, :na] = M.T
    return st.ravel()


COMBOS = [(s, u, t, d) for s in "LR" for u in "LU" for t in "NT" for d in "NU"]


@pytest.mark.parametrize("m,n", [(173, 91), (460, 300), (200, 1), (1, 250)])
@pytest.mark.parametrize("side,uplo,trans,diag", COMBOS)
def test_dtrsm_integer_exact_recovery(torch_cuda, side, uplo, trans, diag, m, n):
    """Unit-diagonal integer triangle, integer X: B = op(A) X (or X op(A)) is
    exact and so is every substitution and GEMM update, so X comes back bitwise
    (m, n not multiples of the 64-row leaf; unstored triangle is NaN).  A
    triangle of 257-512 unknowns is two block solves with a thin GEMM update
    between them."""
    rng = np.random.default_rng(sum(ord(ch) * 31**i for i, ch in enumerate(side + uplo + trans + diag)) + m)
    na = m if side == "L" else n
    A = rng.integers(-1, 2, (na, na)).astype(np.float64)
    T = np.tril(A, -1) if uplo == "L" else np.triu(A, 1)
    T = T + np.eye(na)
    if diag == "N":
        store_T = T  # diagonal stored as 1, solved as non-unit (division by 1 is exact)
    else:
        store_T = T.copy()
        store_T[np.diag_indices(na)] = 5.0  # must be ignored for unit diagonal
    opA = T.T if trans == "T" else T
    X = rng.integers(-3, 4, (m, n)).astype(np.float64)
    Bm = opA @ X if side == "L" else X @ opA
    assert np.max(np.abs(Bm)) < 2**40
    lda, ldb = na + 2, m + 3
    a = triangle_storage(store_T, uplo, lda)
    bb = np.full((n, ldb), -777.0)
    bb[:, :m] = Bm.T
    got = trsm(torch_cuda, side, uplo, trans, diag, m, n, 1.0, a, lda, bb.ravel(), ldb)
    G = got.reshape(n, ldb)
    assert np.array_equal(G[:, :m].T, X)
    assert np.all(G[:, m:] == -777.0)


@pytest.mark.parametrize("m,n", [(200, 150), (460, 300)])
@pytest.mark.parametrize("side,uplo,trans,diag", COMBOS)
def test_dtrsm_residual_random(torch_cuda, oracle, side, uplo, trans, diag, m, n):
    rng = np.random.default_rng(7)
    na = m if side == "L" else n
    A = rng.uniform(-1, 1, (na, na))
    A[np.diag_indices(na)] = 8 + np.abs(rng.uniform(-1, 1, na))
    T = np.tril(A) if uplo == "L" else np.triu(A)
    Tsolve = T.copy()
    if diag == "U":
        Tsolve[np.diag_indices(na)] = 1.0
    opA = Tsolve.T if trans == "T" else Tsolve
    lda, ldb = na + 1, m + 2
    a = triangle_storage(T, uplo, lda)
    B = rng.uniform(-1, 1, (m, n))
    bb = np.full((n, ldb), -777.0)
    bb[:, :m] = B.T
    alpha = -1.5
    got = trsm(torch_cuda, side, uplo, trans, diag, m, n, alpha, a, lda, bb.ravel(), ldb)
    X = got.reshape(n, ldb)[:, :m].T
    R = (opA @ X if side == "L" else X @ opA) - alpha * B
    scale = (np.abs(opA) @ np.abs(X)) if side == "L" else (np.abs(X) @ np.abs(opA))
    assert np.all(np.abs(R) <= 8 * na * 2.0**-53 * (scale + abs(alpha) * np.abs(B)))
    # the oracle (the reference's column substitution) agrees to the same order
    a_orc = np.nan_to_num(a, nan=0.0)
    ref = bb.ravel().copy()
    oracle.dtrsm(side, uplo, trans, diag, m, n, alpha, a_orc, lda, ref, ldb)
    Xr = ref.reshape(n, ldb)[:, :m].T
    cond_scale = np.max(np.abs(X)) + 1
    assert np.max(np.abs(X - Xr)) <= 64 * na * 2.0**-53 * cond_scale


@pytest.mark.parametrize("side,uplo,trans", [(s, u, t) for s in "LR" for u in "LU" for t in "NT"])
def test_dtrsm_lookahead_integer_exact_recovery(torch_cuda, side, uplo, trans):
    """A triangle of 4160 (second halves of 2048 and more: the plain form
    overlaps the far part of each such update with the near part's solve on a
    second stream): integer systems recover X bitwise in every side, uplo and
    transpose form, so no update runs before the unknowns it reads are final."""
    rng = np.random.default_rng(sum(ord(ch) * 31**i for i, ch in enumerate("la" + side + uplo + trans)))
    na, other = 4160, 72
    m, n = (na, other) if side == "L" else (other, na)
    A = rng.integers(-1, 2, (na, na)).astype(np.float64)
    T = (np.tril(A, -1) if uplo == "L" else np.triu(A, 1)) + np.eye(na)
    opA = T.T if trans == "T" else T
    X = rng.integers(-3, 4, (m, n)).astype(np.float64)
    Bm = opA @ X if side == "L" else X @ opA
    assert np.max(np.abs(Bm)) < 2**40
    lda, ldb = na + 1, m + 3
    a = triangle_storage(T, uplo, lda)
    bb = np.full((n, ldb), -777.0)
    bb[:, :m] = Bm.T
    got = trsm(torch_cuda, side, uplo, trans, "N", m, n, 1.0, a, lda, bb.ravel(), ldb).reshape(n, ldb)
    assert np.array_equal(got[:, :m].T, X)
    assert np.all(got[:, m:] == -777.0)


LEAF_CASES = ["plain", "wide_range", "tiny_column", "huge_diagonal", "zeros", "padded"]


@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("case", LEAF_CASES)
def test_dtrsm_single_leaf_matches_oracle_exactly(torch_cuda, oracle, uplo, case):
    """One leaf solve (m <= 64, side L, no transpose) performs the reference's
    column substitution x_k /= a_kk; x_i = fma(a_ik, -x_k, x_i) in the
    oracle's order, so X equals the oracle's to the last bit, on the
    reciprocal-based division's range (2^-400 <= |x|, |a_kk| <= 2^400) and
    outside it (tiny right-hand sides, a huge diagonal: the leaf solves again
    with the division).  Zeros compare by value: the oracle skips the updates
    of a zero x_k, which can only change the sign of a zero."""
    rng = np.random.default_rng(LEAF_CASES.index(case) * 2 + (uplo == "U"))
    m = 37 if case == "padded" else 64
    n = 200
    A = rng.uniform(-1, 1, (m, m))
    d = rng.uniform(0.5, 2.0, m) * rng.choice([-1.0, 1.0], m)
    if case == "wide_range":
        d *= 2.0 ** rng.integers(0, 40, m)
    if case == "huge_diagonal":
        d[m // 2] = 1e125  # above 2^400
    A[np.diag_indices(m)] = d
    T = np.tril(A) if uplo == "L" else np.triu(A)
    B = rng.uniform(-1, 1, (m, n))
    if case == "wide_range":
        B *= 2.0 ** rng.integers(-350, 350, (m, n))
    if case == "tiny_column":
        B[:, 70] *= 1e-150  # below 2^-400: the wave holding column 70 divides
    if case == "zeros":
        B[rng.random((m, n)) < 0.3] = 0.0
        B[:, 3] = -0.0
    lda, ldb = m + 1, m + 2
    a = triangle_storage(T, uplo, lda)
    bb = np.full((n, ldb), -777.0)
    bb[:, :m] = B.T
    got = trsm(torch_cuda, "L", uplo, "N", "N", m, n, 1.0, a, lda, bb.ravel(), ldb).reshape(n, ldb)
    ref = bb.ravel().copy()
    oracle.dtrsm("L", uplo, "N", "N", m, n, 1.0, np.nan_to_num(a, nan=0.0), lda, ref, ldb)
    ref = ref.reshape(n, ldb)
    assert np.all(np.isfinite(ref[:, :m]))
    assert np.array_equal(got[:, :m], ref[:, :m])
    if case != "zeros":
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    assert np.all(got[:, m:] == -777.0)


@pytest.mark.parametrize("mode,param", [(3, 14), (1, 8)])
def test_dtrsm_split_precision_updates(torch_cuda, mode, param):
    """Off-diagonal updates in split precision (the updates whose m, n and k are
    all at least 1024; smaller ones run on FP64 MFMA): exact on integer systems."""
    rng = np.random.default_rng(3)
    m, n = 2200, 1100
    T = np.tril(rng.integers(-1, 2, (m, m)).astype(np.float64), -1) + np.eye(m)
    X = rng.integers(-3, 4, (m, n)).astype(np.float64)
    B = T @ X
    got = trsm(torch_cuda, "L", "L", "N", "N", m, n, 1.0, triangle_storage(T, "L", m), m, B.T.ravel(), m,
               mode=mode, param=param)
    assert np.array_equal(got.reshape(n, m).T, X)


def syrk(torch, uplo, trans, n, k, alpha, a, lda, beta, c, ldc, mode=0, param=0, via="request"):
    from zynum_amd import _lib

    lib = _lib.load()
    da, dc = dev(torch, a), dev(torch, c)
    if via == "fortran":
        I = C.c_int32
        lib.dsyrk_(uplo.encode(), trans.encode(), C.byref(I(n)), C.byref(I(k)), C.byref(C.c_double(alpha)),
                   da.data_ptr(), C.byref(I(lda)), C.byref(C.c_double(beta)), dc.data_ptr(), C.byref(I(ldc)))
        assert lib.zynum_gpu_last_status() == 0
    else:
        st = lib.zynum_gpu_dsyrk(uplo.encode(), trans.encode(), n, k, alpha, da.data_ptr(), lda, beta,
                                 dc.data_ptr(), ldc, None, mode, param, 0)
        _lib.check(st, "dsyrk")
    torch.cuda.synchronize()
    return dc.cpu().numpy()


def test_dsyrk_known_answer(torch_cuda):
    # fortran_compat_test.zig:1742-1749
    a = np.array([1, 3, 2, 4.0])
    c = np.array([0, 777, 0, 0.0])
    got = syrk(torch_cuda, "U", "N", 2, 2, 1.0, a, 2, 0.0, c, 2, via="fortran")
    assert got.tolist() == [5, 777, 11, 25]


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("mode,param", [(0, 0), (3, 14)])
def test_dsyrk_integer_exact_and_triangle_untouched(torch_cuda, uplo, trans, mode, param):
    rng = np.random.default_rng(5)
    n, k = 301, 77
    ar, ac = (n, k) if trans == "N" else (k, n)
    lda, ldc = ar + 3, n + 2
    A = rng.integers(-9, 10, (ar, ac)).astype(np.float64)
    aa = np.full((ac, lda), -555.0)
    aa[:, :ar] = A.T
    Cm = rng.integers(-9, 10, (n, n)).astype(np.float64)
    cc = np.full((n, ldc), -777.0)
    cc[:, :n] = Cm.T
    got = syrk(torch_cuda, uplo, trans, n, k, 0.75, aa.ravel(), lda, -0.25, cc.ravel(), ldc, mode, param)
    G = got.reshape(n, ldc)
    X = A if trans == "N" else A.T
    S = X.astype(np.int64) @ X.T.astype(np.int64)
    want = 0.75 * S + -0.25 * Cm
    tri = np.tril_indices(n) if uplo == "L" else np.triu_indices(n)
    other = np.triu_indices(n, 1) if uplo == "L" else np.tril_indices(n, -1)
    GT = G[:, :n].T
    assert np.array_equal(GT[tri], want[tri])
    assert np.array_equal(GT[other], Cm[other])
    assert np.all(G[:, n:] == -777.0)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("n,k,pad", [(384, 264, 2), (384, 264, 1)])
def test_dsyrk_direct_load_tiles_integer_exact(torch_cuda, uplo, trans, n, k, pad):
    """Full 128x128 triangle tiles with k a multiple of 8 and an even lda take
    the direct-to-LDS tile (pad 2), an odd lda the register-staged one (pad 1):
    both exact on integer inputs, the other triangle and the padding untouched."""
    rng = np.random.default_rng(11)
    ar, ac = (n, k) if trans == "N" else (k, n)
    lda, ldc = ar + pad, n + 2
    A = rng.integers(-9, 10, (ar, ac)).astype(np.float64)
    aa = np.full((ac, lda), -555.0)
    aa[:, :ar] = A.T
    Cm = rng.integers(-9, 10, (n, n)).astype(np.float64)
    cc = np.full((n, ldc), -777.0)
    cc[:, :n] = Cm.T
    got = syrk(torch_cuda, uplo, trans, n, k, 0.75, aa.ravel(), lda, -0.25, cc.ravel(), ldc)
    G = got.reshape(n, ldc)
    X = A if trans == "N" else A.T
    want = 0.75 * (X.astype(np.int64) @ X.T.astype(np.int64)) + -0.25 * Cm
    tri = np.tril_indices(n) if uplo == "L" else np.triu_indices(n)
    other = np.triu_indices(n, 1) if uplo == "L" else np.tril_indices(n, -1)
    GT = G[:, :n].T
    assert np.array_equal(GT[tri], want[tri])
    assert np.array_equal(GT[other], Cm[other])
    assert np.all(G[:, n:] == -777.0)


def test_dsyrk_scale_only(torch_cuda):
    n = 70
    c = np.arange(n * n, dtype=np.float64)
    got = syrk(torch_cuda, "L", "N", n, 0, 1.0, np.zeros(n), n, 0.0, c, n)
    G = got.reshape(n, n).T
    C0 = c.reshape(n, n).T
    assert np.all(G[np.tril_indices(n)] == 0.0)
    assert np.array_equal(G[np.triu_indices(n, 1)], C0[np.triu_indices(n, 1)])


def test_dsyrk_random_vs_oracle(torch_cuda, oracle):
    rng = np.random.default_rng(8)
    n, k = 260, 190
    a = rng.uniform(-1, 1, n * k)
    c = rng.uniform(-1, 1, n * n)
    got = syrk(torch_cuda, "L", "N", n, k, 1.0, a, n, 1.0, c, n)
    ref = c.copy()
    oracle.dsyrk("L", "N", n, k, 1.0, a, n, 1.0, ref, n)
    A = a.reshape(k, n).T
    bound = 2 * (k + 2) * 2.0**-53 * (np.abs(A) @ np.abs(A).T + np.abs(c.reshape(n, n).T))
    G, R = got.reshape(n, n).T, ref.reshape(n, n).T
    tri = np.tril_indices(n)
    assert np.all(np.abs(G[tri] - R[tri]) <= bound[tri])


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("n,k", [(301, 1100), (4096, 1024)])
def test_dsyrk_long_k_integer_exact(torch_cuda, uplo, trans, n, k):
    """Long K over few and many triangle tiles (6 for n = 301; 528 for
    n = 4096, walked in 8 x 8 super-tiles with a partial last super-row).
    Integer inputs: exact, the other triangle and the padding untouched."""
    rng = np.random.default_rng(n + k)
    ar, ac = (n, k) if trans == "N" else (k, n)
    lda, ldc = ar + 2, n + 1
    A = rng.integers(-9, 10, (ar, ac)).astype(np.float64)
    aa = np.full((ac, lda), -555.0)
    aa[:, :ar] = A.T
    Cm = rng.integers(-9, 10, (n, n)).astype(np.float64)
    cc = np.full((n, ldc), -777.0)
    cc[:, :n] = Cm.T
    got = syrk(torch_cuda, uplo, trans, n, k, -0.5, aa.ravel(), lda, 2.0, cc.ravel(), ldc)
    G = got.reshape(n, ldc)
    X = A if trans == "N" else A.T
    want = -0.5 * (X @ X.T) + 2.0 * Cm  # every partial sum is an integer below 2^53
    GT = G[:, :n].T
    lower = np.tril(np.ones((n, n), dtype=bool))
    tri = lower if uplo == "L" else lower.T
    assert np.array_equal(GT[tri], want[tri])
    assert np.array_equal(GT[~tri], Cm[~tri])
    assert np.all(G[:, n:] == -777.0)


def test_dsyrk_long_k_random_within_bound(torch_cuda, oracle):
    rng = np.random.default_rng(21)
    n, k = 700, 2000
    a = rng.uniform(-1, 1, n * k)
    c = rng.uniform(-1, 1, n * n)
    got = syrk(torch_cuda, "U", "N", n, k, 1.5, a, n, 0.5, c, n)
    A = a.reshape(k, n).T
    exact = 1.5 * (A @ A.T) + 0.5 * c.reshape(n, n).T
    bound = 2 * (k + 2) * 2.0**-53 * (1.5 * (np.abs(A) @ np.abs(A).T) + 0.5 * np.abs(c.reshape(n, n).T))
    G = got.reshape(n, n).T
    tri = np.triu_indices(n)
    assert np.all(np.abs(G[tri] - exact[tri]) <= 2 * bound[tri])
