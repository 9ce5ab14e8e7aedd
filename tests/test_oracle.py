"""Pins the CPU oracle (a restatement of Zynum's portable CPU path) against
the reference's own known-answer tests, exact arithmetic and closed forms.

Each check is chosen so that a plausible mistake in the oracle (transposed
operand, wrong leading dimension, dropped beta term, non-fused accumulation,
wrong triangle) fails it.
"""
from fractions import Fraction as Fr

import numpy as np
import pytest

from conftest import XorShift, exact_gemm, fill_padded, fill_row_major, int_gemm


def test_dgemm_known_answer_column_major(oracle):
    # fortran_compat_test.zig:537-563
    a = np.array([1, 2, 3, 4, 5, 6], dtype=np.float64)
    b = np.array([7, 8, 9, 10, 11, 12], dtype=np.float64)
    c = np.zeros(4)
    oracle.dgemm("N", "N", 2, 2, 3, 1.0, a, 2, b, 3, 0.0, c, 2)
    assert c.tolist() == [76, 100, 103, 136]


def test_sgemm_known_answer(oracle):
    a = np.array([1, 2, 3, 4, 5, 6], dtype=np.float32)
    b = np.array([7, 8, 9, 10, 11, 12], dtype=np.float32)
    c = np.zeros(4, dtype=np.float32)
    oracle.sgemm("N", "N", 2, 2, 3, 1.0, a, 2, b, 3, 0.0, c, 2)
    assert c.tolist() == [76, 100, 103, 136]


def test_gemm_nan_isolation(oracle):
    # fortran_compat_test.zig:1012-1045
    nan = float("nan")
    a = np.array([1, 2, 3, 4.0])
    b = np.array([5, 6, 7, 8.0])
    c = np.full(4, nan)
    oracle.dgemm("N", "N", 2, 2, 2, 1.0, a, 2, b, 2, 0.0, c, 2)
    assert c.tolist() == [23, 34, 31, 46]
    a[:] = nan
    b[:] = nan
    c = np.array([1, 2, 3, 4.0])
    oracle.dgemm("N", "N", 2, 2, 2, 0.0, a, 2, b, 2, 2.0, c, 2)
    assert c.tolist() == [2, 4, 6, 8]


# Padded-case shapes of fortran_compat_test.zig:638-730,973-1010,1077-1091
# (m, n, k, lda, ldb, ldc, alpha, beta) per transpose pair.
NN_CASES = [
    (5, 5, 3, 7, 5, 8, -0.5, 2.0), (7, 5, 3, 9, 5, 10, -0.5, 2.0), (35, 37, 19, 37, 23, 39, 1.0, 0.0),
    (70, 97, 31, 73, 37, 75, 1.0, 0.0), (64, 129, 128, 67, 132, 69, 0.001, 0.0),
    (130, 161, 67, 133, 71, 135, 1.0, 0.0), (96, 97, 96, 96, 96, 96, 0.75, -0.25),
    (512, 16, 1024, 512, 1024, 512, 1.0, 0.0), (127, 129, 65, 131, 69, 137, 1.0, 0.0),
]
NT_CASES = [
    (35, 37, 67, 39, 41, 43, -0.5, 2.0), (37, 35, 71, 41, 39, 43, 1.0, 0.0), (128, 128, 128, 131, 137, 139, 1.0, 0.0),
    (17, 19, 2049, 23, 29, 31, 0.75, -0.25), (1, 257, 131, 3, 263, 5, -0.5, 2.0),
    (257, 1, 131, 263, 3, 269, -0.5, 2.0), (193, 65, 577, 197, 69, 199, 0.75, -0.25),
    (127, 129, 65, 131, 133, 137, 1.0, 0.0),
]
TN_CASES = [  # (transposed_b, m, n, k, lda, ldb, ldc, alpha, beta)
    (False, 35, 37, 67, 71, 73, 43, -0.5, 2.0), (True, 35, 37, 67, 71, 41, 43, -0.5, 2.0),
    (False, 37, 35, 71, 73, 79, 43, 0.75, -0.25), (True, 37, 35, 71, 73, 39, 43, 0.75, -0.25),
    (False, 127, 129, 65, 69, 71, 137, 1.0, 0.0), (True, 127, 129, 65, 69, 133, 137, 1.0, 0.0),
    (False, 17, 19, 2049, 2053, 2057, 23, 0.75, -0.25), (True, 17, 19, 2049, 2053, 23, 23, 0.75, -0.25),
    (False, 193, 65, 577, 581, 583, 199, 0.75, -0.25), (True, 193, 65, 577, 581, 69, 199, 0.75, -0.25),
    (False, 5, 7, 11, 13, 13, 11, 0.0, -0.25), (True, 5, 7, 11, 13, 9, 11, 0.0, -0.25),
    (False, 5, 7, 0, 3, 3, 11, 1.0, 2.0), (True, 5, 7, 0, 3, 9, 11, 1.0, 2.0),
    (False, 5, 7, 3, 5, 5, 11, -0.5, 2.0), (True, 5, 7, 3, 5, 9, 11, -0.5, 2.0),
]


def all_cases():
    out = []
    for (m, n, k, lda, ldb, ldc, al, be) in NN_CASES:
        out.append(("N", "N", m, n, k, lda, ldb, ldc, al, be))
    for (m, n, k, lda, ldb, ldc, al, be) in NT_CASES:
        out.append(("N", "T", m, n, k, lda, ldb, ldc, al, be))
    for (tb, m, n, k, lda, ldb, ldc, al, be) in TN_CASES:
        out.append(("T", "T" if tb else "N", m, n, k, lda, ldb, ldc, al, be))
    return out


def operands(ta, tb, m, n, k, lda, ldb, ldc, dtype, integer, salts=(1, 5, 9)):
    sent = -777.0 if dtype == np.float32 else -999.0
    a = fill_padded(m, k, lda, salts[0], sent, dtype, integer) if ta == "N" else fill_padded(k, m, lda, salts[0], sent, dtype, integer)
    b = fill_padded(k, n, ldb, salts[1], sent, dtype, integer) if tb == "N" else fill_padded(n, k, ldb, salts[1], sent, dtype, integer)
    c = fill_padded(m, n, ldc, salts[2], sent, dtype, integer)
    return a, b, c


@pytest.mark.parametrize("case", all_cases())
def test_integer_inputs_bitwise_exact(oracle, case):
    """Integer-valued inputs: every summation order is exact, so the oracle must
    equal exact integer arithmetic bitwise, and padding rows must be untouched."""
    ta, tb, m, n, k, lda, ldb, ldc, al, be = case
    if al == 0.001:
        al = 0.125  # keep alpha dyadic so the exact result is representable
    a, b, c = operands(ta, tb, m, n, k, lda, ldb, ldc, np.float64, True)
    want = int_gemm(ta, tb, m, n, k, al, a, lda, b, ldb, be, c, ldc)
    got = c.copy()
    oracle.dgemm(ta, tb, m, n, k, al, a, lda, b, ldb, be, got, ldc)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


@pytest.mark.parametrize("case", [c for c in all_cases() if c[4] <= 600])
def test_sample_inputs_within_reference_tolerance(oracle, case):
    """The reference's own tolerance (1e-10 for f64, fortran_compat_test.zig:375)
    against exact rationals, on the reference's non-integer sample values."""
    ta, tb, m, n, k, lda, ldb, ldc, al, be = case
    if m * n * k > 300_000:
        pytest.skip("exact rational reference too slow for this shape")
    a, b, c = operands(ta, tb, m, n, k, lda, ldb, ldc, np.float64, False)
    want = exact_gemm(ta, tb, m, n, k, al, a, lda, b, ldb, be, c, ldc)
    got = c.copy()
    oracle.dgemm(ta, tb, m, n, k, al, a, lda, b, ldb, be, got, ldc)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-10)
    # padding untouched bitwise
    mask = np.ones_like(got, dtype=bool)
    mask[: ldc * n].reshape(n, ldc)[:, :m] = False
    assert np.array_equal(got[mask], c[mask])


def test_fma_chain_order_pinned(oracle):
    """The packed path accumulates with a fused multiply-add per step
    (packed_simd.zig:140-162).  With a = [1, 1 + 2^-27], b = [-1, 1 - 2^-27]
    the fused chain gives fma(1 + 2^-27, 1 - 2^-27, -1) = -2^-54 exactly while
    separately rounded products give 0.  The oracle must give the fused value
    on the packed route and the non-fused value on the portable loop
    (gemm.zig:1422-1429) that the reference takes for small TN shapes."""
    e = 2.0**-27
    a = np.array([1.0, 1.0 + e])
    b = np.array([-1.0, 1.0 - e])
    c = np.zeros(1)
    oracle.dgemm("N", "N", 1, 1, 2, 1.0, a, 1, b, 2, 0.0, c, 1)
    assert c[0] == -(2.0**-54)
    c2 = np.zeros(1)
    oracle.dgemm("T", "N", 1, 1, 2, 1.0, a, 2, b, 2, 0.0, c2, 1)
    assert c2[0] == 0.0


def test_simd_tiers_agree_bitwise(oracle):
    """AVX-512 (16x8 tile), AVX2 (4x8 tile) and scalar rows restate the same
    per-element FMA order; their outputs must agree bitwise."""
    rng = np.random.default_rng(7)
    m, n, k = 70, 45, 333
    a = rng.uniform(-1, 1, m * k)
    b = rng.uniform(-1, 1, k * n)
    c0 = rng.uniform(-1, 1, m * n)
    native = oracle.isa()
    outs = []
    for isa in (2, 1, 0):
        if isa > native:
            continue
        oracle.force_isa(isa)
        c = c0.copy()
        oracle.dgemm("N", "N", m, n, k, 1.5, a, m, b, k, 0.5, c, m)
        outs.append(c)
    oracle.force_isa(-1)
    assert len(outs) >= 2
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


def test_threads_split_matches_serial(oracle):
    """planner.zig:296-325: column-panel tasks own disjoint C columns, so any
    thread count gives the same bits (cf. triangular.zig:376-406)."""
    rng = np.random.default_rng(3)
    m, n, k = 97, 131, 257
    a = rng.uniform(-1, 1, m * k)
    b = rng.uniform(-1, 1, k * n)
    c0 = rng.uniform(-1, 1, m * n)
    outs = []
    for t in (1, 3, 8):
        oracle.set_threads(t)
        c = c0.copy()
        oracle.dgemm("N", "N", m, n, k, 0.75, a, m, b, k, -0.25, c, m)
        outs.append(c)
    oracle.set_threads(0)
    assert all(np.array_equal(o, outs[0]) for o in outs)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_cblas_row_major_cases_against_exact(oracle, ta, tb):
    """cblas_compat_test.zig:326-364: 35x37x67 with padded ld, alpha=-0.5,
    beta=2, xorshift inputs; the row-major call is the column-major product
    with A/B and m/n swapped (cblas.zig:1428-1439)."""
    m, n, k = 35, 37, 67
    a_rows, a_cols = (m, k) if ta == "N" else (k, m)
    b_rows, b_cols = (k, n) if tb == "N" else (n, k)
    lda, ldb, ldc = a_cols + 4, b_cols + 4, n + 5
    rng = XorShift(0x74A3119D)
    a = fill_row_major(rng, a_rows, a_cols, lda)
    b = fill_row_major(rng, b_rows, b_cols, ldb)
    c = fill_row_major(rng, m, n, ldc)
    got = c.copy()
    oracle.dgemm(tb, ta, n, m, k, -0.5, b, ldb, a, lda, 2.0, got, ldc)
    want = exact_gemm(tb, ta, n, m, k, -0.5, b, ldb, a, lda, 2.0, c, ldc)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-10)


def test_dsyrk_known_answer_and_triangle(oracle):
    # fortran_compat_test.zig:1742-1749: upper, A = [[1,2],[3,4]] column-major
    a = np.array([1, 3, 2, 4.0])
    c = np.array([0, 777, 0, 0.0])
    oracle.dsyrk("U", "N", 2, 2, 1.0, a, 2, 0.0, c, 2)
    assert c.tolist() == [5, 777, 11, 25]


def test_dsyrk_integer_exact_lower(oracle):
    n, k, lda, ldc = 37, 23, 41, 39
    a = fill_padded(n, k, lda, 3, -999.0, np.float64, True)
    c = fill_padded(n, n, ldc, 7, -999.0, np.float64, True)
    got = c.copy()
    oracle.dsyrk("L", "N", n, k, 0.75, a, lda, -0.25, got, ldc)
    A = a[: lda * k].reshape(k, lda)[:, :n].T.astype(np.int64)
    S = A @ A.T
    C0 = c[: ldc * n].reshape(n, ldc)[:, :n].T
    G = got[: ldc * n].reshape(n, ldc)[:, :n].T
    low = np.tril_indices(n)
    assert np.array_equal(G[low], 0.75 * S[low] + -0.25 * C0[low])
    up = np.triu_indices(n, 1)
    assert np.array_equal(G[up], C0[up])


def test_dtrsm_known_answer(oracle):
    # fortran_compat_test.zig:1779-1798: dtrmm gives [16,18,19,21]; dtrsm restores [4,6,5,7]
    tri = np.array([1, 777, 2, 3.0])
    b = np.array([16, 18, 19, 21.0])
    oracle.dtrsm("L", "U", "N", "N", 2, 2, 1.0, tri, 2, b, 2)
    np.testing.assert_allclose(b, [4, 6, 5, 7], atol=1e-12)


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_dtrsm_residual_all_forms(oracle, side, uplo, trans, diag):
    """Residual |op(A) X - alpha B| (or |X op(A) - alpha B|) small relative to
    |op(A)||X| for every side/uplo/trans/diag; the unstored triangle of A holds
    poison that must never be read."""
    rng = np.random.default_rng(11)
    m, n = 29, 17
    na = m if side == "L" else n
    lda, ldb = na + 3, m + 2
    A = rng.uniform(-1, 1, (na, na))
    A[np.diag_indices(na)] = 4 + np.abs(rng.uniform(-1, 1, na))
    T = np.tril(A) if uplo == "L" else np.triu(A)
    if diag == "U":
        T = T.copy()
        T[np.diag_indices(na)] = 1.0
    store = np.full((na, lda), 1e300)
    store[:, :na] = (np.tril(A) if uplo == "L" else np.triu(A)).T  # column j -> row j of store
    poison_tri = np.triu_indices(na, 1) if uplo == "L" else np.tril_indices(na, -1)
    st = store[:, :na].T.copy()
    st[poison_tri] = np.nan
    store[:, :na] = st.T
    a = store.ravel()
    B = rng.uniform(-1, 1, (m, n))
    bb = np.full((n, ldb), -999.0)
    bb[:, :m] = B.T
    b = bb.ravel()
    alpha = -1.5
    oracle.dtrsm(side, uplo, trans, diag, m, n, alpha, a, lda, b, ldb)
    X = b.reshape(n, ldb)[:, :m].T
    opA = T.T if trans == "T" else T
    R = (opA @ X if side == "L" else X @ opA) - alpha * B
    scale = (np.abs(opA) @ np.abs(X)) if side == "L" else (np.abs(X) @ np.abs(opA))
    assert np.all(np.abs(R) <= 64 * na * 2.0**-53 * (scale + abs(alpha) * np.abs(B)))
    assert np.all(b.reshape(n, ldb)[:, m:] == -999.0)


def test_dtrsm_closed_forms(oracle):
    """Identity L gives X = alpha*B exactly; diagonal L gives B*alpha/d."""
    m, n = 8, 5
    eye = np.eye(m).ravel(order="F")
    B = np.arange(m * n, dtype=np.float64)
    X = B.copy()
    oracle.dtrsm("L", "L", "N", "N", m, n, 0.5, eye, m, X, m)
    assert np.array_equal(X, 0.5 * B)
    d = np.diag([2.0, 4.0, 8.0, 0.5, 1.0, 16.0, 2.0, 4.0]).ravel(order="F")
    X = B.copy()
    oracle.dtrsm("L", "U", "N", "N", m, n, 1.0, d, m, X, m)
    want = (B.reshape(n, m) / np.array([2.0, 4.0, 8.0, 0.5, 1.0, 16.0, 2.0, 4.0])).ravel()
    assert np.array_equal(X, want)


def test_double_double_against_fractions(oracle):
    """dd GEMM entries agree with exact rationals to ~2^-100 relative."""
    rng = np.random.default_rng(5)
    m, n, k = 6, 5, 300
    a = rng.uniform(-1, 1, m * k) * np.exp2(rng.integers(-20, 20, m * k))
    b = rng.uniform(-1, 1, k * n) * np.exp2(rng.integers(-20, 20, k * n))
    c0 = rng.uniform(-1, 1, m * n)
    rows, cols = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    hi, lo = oracle.dd_gemm_entries("N", "N", k, 1.5, a, m, b, k, 0.25, c0, m, rows.ravel(), cols.ravel())
    for e, (i, j) in enumerate(zip(rows.ravel(), cols.ravel())):
        s = sum(Fr(a[i + p * m]) * Fr(b[p + j * k]) for p in range(k))
        exact = Fr(1.5) * s + Fr(0.25) * Fr(c0[i + j * m])
        mag = sum(abs(Fr(a[i + p * m]) * Fr(b[p + j * k])) for p in range(k))
        err = abs(Fr(hi[e]) + Fr(lo[e]) - exact)
        assert err <= mag * Fr(2) ** -100


def test_splitmix_generator_matches_numpy(oracle):
    from zynum_amd import synth

    c = oracle.fill_uniform(12345, 1000, 5000)
    p = synth.uniform(12345, 1000, 5000)
    assert np.array_equal(c, p)
    assert p.min() >= -1.0 and p.max() < 1.0


# ---------------------------------------------------------------------------
# FP32 (sgemm) pins: the same restatement in single precision
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("case", all_cases())
def test_sgemm_integer_inputs_bitwise_exact(oracle, case):
    """Integer inputs in [-14, 14] whose sums stay below 2^24 (every order
    exact in FP32): the oracle's SGEMM equals exact arithmetic bitwise and
    leaves the padding (sentinel -777) untouched."""
    ta, tb, m, n, k, lda, ldb, ldc, al, be = case
    if 14 * 14 * k + 2 * 14 >= 2**24 / 8:
        pytest.skip("sums may exceed 2^24 in FP32")
    if al == 0.001:
        al = 0.125
    a, b, c = operands(ta, tb, m, n, k, lda, ldb, ldc, np.float32, True)
    want = int_gemm(ta, tb, m, n, k, al, a, lda, b, ldb, be, c, ldc)
    got = c.copy()
    oracle.sgemm(ta, tb, m, n, k, al, a, lda, b, ldb, be, got, ldc)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_sgemm_fma_chain_order_pinned(oracle):
    """packed_simd.zig:140-162 in FP32: a = [1, 1 + 2^-13], b = [-1, 1 - 2^-13]:
    the fused chain gives fma(1 + 2^-13, 1 - 2^-13, -1) = -2^-26 exactly, while
    the separately rounded product is 1 (the float below 1 is 1 - 2^-24), so
    the small TN portable loop (gemm.zig:1422-1429) gives 0."""
    e = np.float32(2.0**-13)
    a = np.array([1.0, 1.0 + e], dtype=np.float32)
    b = np.array([-1.0, 1.0 - e], dtype=np.float32)
    c = np.zeros(1, dtype=np.float32)
    oracle.sgemm("N", "N", 1, 1, 2, 1.0, a, 1, b, 2, 0.0, c, 1)
    assert c[0] == np.float32(-(2.0**-26))
    c2 = np.zeros(1, dtype=np.float32)
    oracle.sgemm("T", "N", 1, 1, 2, 1.0, a, 2, b, 2, 0.0, c2, 1)
    assert c2[0] == 0.0


def test_sgemm_epilogue_rounding_pinned(oracle):
    """epilogue.zig:10-39 in FP32: for beta not 0 or 1 the result is
    fma(acc, alpha, RN(beta C)).  acc = alpha = 1 + 2^-23, C = 1,
    beta = -(1 + 2^-22): alpha acc = 1 + 2^-22 + 2^-46 exactly, so the fused
    form gives 2^-46 and a separately rounded alpha acc gives 0."""
    u = 2.0**-23
    a = np.array([1.0 + u], dtype=np.float32)
    b = np.array([1.0], dtype=np.float32)
    c = np.array([1.0], dtype=np.float32)
    got = c.copy()
    oracle.sgemm("N", "N", 1, 1, 1, 1.0 + u, a, 1, b, 1, -(1.0 + 2 * u), got, 1)
    assert got[0] == np.float32(2.0**-46)


def test_sgemm_simd_tiers_and_threads_agree(oracle):
    """AVX-512, AVX2 and scalar FP32 tiles, and any thread count, give the same bits."""
    rng = np.random.default_rng(17)
    m, n, k = 70, 45, 333
    a = rng.uniform(-1, 1, m * k).astype(np.float32)
    b = rng.uniform(-1, 1, k * n).astype(np.float32)
    c0 = rng.uniform(-1, 1, m * n).astype(np.float32)
    native = oracle.isa()
    outs = []
    for isa in (2, 1, 0):
        if isa > native:
            continue
        oracle.force_isa(isa)
        c = c0.copy()
        oracle.sgemm("N", "N", m, n, k, 1.5, a, m, b, k, 0.5, c, m)
        outs.append(c)
    oracle.force_isa(-1)
    for t in (1, 3):
        oracle.set_threads(t)
        c = c0.copy()
        oracle.sgemm("N", "N", m, n, k, 1.5, a, m, b, k, 0.5, c, m)
        outs.append(c)
    oracle.set_threads(0)
    assert len(outs) >= 3
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "T")])
def test_sgemm_random_within_bound_of_float64(oracle, ta, tb):
    """|C32 - C_exact| <= 2 (K + 2) u32 (|alpha||A||B| + |beta||C0|) against the
    float64 product of the same float32 inputs (exact to far below u32)."""
    rng = np.random.default_rng(23)
    m, n, k = 90, 70, 500
    lda = (m if ta == "N" else k) + 2
    ldb = (k if tb == "N" else n) + 3
    a = rng.uniform(-1, 1, lda * (k if ta == "N" else m)).astype(np.float32)
    b = rng.uniform(-1, 1, ldb * (n if tb == "N" else k)).astype(np.float32)
    c = rng.uniform(-1, 1, m * n).astype(np.float32)
    got = c.copy()
    oracle.sgemm(ta, tb, m, n, k, 1.25, a, lda, b, ldb, -0.5, got, m)

    def mat(x, rows, cols, ld):
        return x[: ld * cols].reshape(cols, ld)[:, :rows].T.astype(np.float64)

    A = mat(a, m, k, lda) if ta == "N" else mat(a, k, m, lda).T
    B = mat(b, k, n, ldb) if tb == "N" else mat(b, n, k, ldb).T
    C0 = mat(c, m, n, m)
    exact = 1.25 * (A @ B) - 0.5 * C0
    bound = 2 * (k + 2) * 2.0**-24 * (1.25 * np.abs(A) @ np.abs(B) + 0.5 * np.abs(C0))
    assert np.all(np.abs(mat(got, m, n, m) - exact) <= bound)
