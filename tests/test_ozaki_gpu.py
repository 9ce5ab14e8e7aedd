"""Split-precision DGEMM (Ozaki scheme I with INT8 slices, Ozaki scheme II
with INT8 residues + CRT) and the on-device double-double reference.

Bars: bitwise on integer-valued inputs (scaling, slicing and residues are
exact there, and the CRT recovers the exact integer product); on random
inputs, the method's own error bound against the double-double value, and
"no worse than plain FP64 DGEMM" on the wide-range workload.
"""
import math

import numpy as np
import pytest

from conftest import int_gemm
from test_oracle import operands

pytestmark = pytest.mark.gpu

MODULI = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193, 191, 181, 179, 173]


def run(torch, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, mode, param, tile=0):
    from zynum_amd import _lib, blas

    da, db, dc = (torch.from_numpy(x.copy()).cuda() for x in (a, b, c))
    st = blas.gemm_request(ta != "N", tb != "N", m, n, k, alpha, da.data_ptr(), lda, db.data_ptr(), ldb, beta,
                           dc.data_ptr(), ldc, _lib.F64, mode=mode, mode_param=param, tile=tile)
    _lib.check(st, f"mode {mode}/{param}")
    torch.cuda.synchronize()
    return dc.cpu().numpy()


def dd_ref(torch, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    from zynum_amd import _lib

    lib = _lib.load()
    da, db, dc = (torch.from_numpy(x.copy()).cuda() for x in (a, b, c))
    hi = torch.zeros(ldc * n, dtype=torch.float64, device="cuda")
    lo = torch.zeros(ldc * n, dtype=torch.float64, device="cuda")
    st = lib.zynum_gpu_dd_gemm(ta.encode(), tb.encode(), m, n, k, alpha, da.data_ptr(), lda, db.data_ptr(), ldb,
                               beta, dc.data_ptr(), ldc, hi.data_ptr(), lo.data_ptr(), ldc, None)
    _lib.check(st, "dd")
    return hi.cpu().numpy(), lo.cpu().numpy()


def logical(x, rows, cols, ld):
    return x[: ld * cols].reshape(cols, ld)[:, :rows].T


def test_dd_gemm_matches_cpu_double_double(torch_cuda, oracle):
    """The device double-double product equals the CPU double-double oracle to
    ~2^-100 relative on sampled entries (both use TwoProd/TwoSum)."""
    rng = np.random.default_rng(9)
    m, n, k = 70, 53, 600
    a = rng.uniform(-1, 1, m * k) * np.exp2(rng.integers(-16, 17, m * k))
    b = rng.uniform(-1, 1, k * n) * np.exp2(rng.integers(-16, 17, k * n))
    c = rng.uniform(-1, 1, m * n)
    hi, lo = dd_ref(torch_cuda, "N", "N", m, n, k, 1.5, a, m, b, k, 0.25, c, m)
    rows = rng.integers(0, m, 200)
    cols = rng.integers(0, n, 200)
    chi, clo = oracle.dd_gemm_entries("N", "N", k, 1.5, a, m, b, k, 0.25, c, m, rows, cols)
    idx = rows + cols * m
    mag = np.abs(logical(a.reshape(-1), m, k, m)) @ np.abs(logical(b.reshape(-1), k, n, k))
    diff = np.abs((hi[idx] - chi) + (lo[idx] - clo))
    assert np.all(diff <= 2.0**-98 * (1.5 * mag[rows, cols] + 1))


@pytest.mark.parametrize("mode,param", [(3, 14), (3, 8), (3, 20), (1, 2), (1, 4), (1, 7), (2, 3), (2, 7)])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_integer_inputs_exact(torch_cuda, mode, param, ta, tb):
    """Integer inputs in [-14, 14]: power-of-two scaling, slicing and residues
    are exact, the CRT recovers the exact integer product, so the result must
    equal exact arithmetic bitwise (odd sizes, padded ld, alpha/beta)."""
    m, n, k = 131, 77, 203
    lda = (m if ta == "N" else k) + 5
    ldb = (k if tb == "N" else n) + 3
    ldc = m + 2
    a, b, c = operands(ta, tb, m, n, k, lda, ldb, ldc, np.float64, True, salts=(3, 8, 13))
    want = int_gemm(ta, tb, m, n, k, 0.75, a, lda, b, ldb, -0.25, c, ldc)
    got = run(torch_cuda, ta, tb, m, n, k, 0.75, a, lda, b, ldb, -0.25, c, ldc, mode, param)
    assert np.array_equal(got, want)


def crt_alpha(nmod: int, k: int) -> int:
    log2p = sum(math.log2(x) for x in MODULI[:nmod])
    alpha = math.floor((log2p - 1) / 2)
    while 2 * math.log2(2.0**alpha + 0.5 * math.sqrt(k)) > log2p - 1 - 1e-6:
        alpha -= 1
    return alpha


@pytest.mark.parametrize("nmod", [10, 14, 16])
def test_crt_error_bound_random(torch_cuda, nmod):
    """Ozaki II error bound: |C - C_exact| <= 4 sqrt(K) 2^-alpha ||a_i|| ||b_j||
    (rounding a, b to integers after scaling their norms to 2^alpha) plus one
    rounding of the result, checked against the double-double product."""
    rng = np.random.default_rng(nmod)
    m, n, k = 192, 160, 1000
    a = rng.uniform(-1, 1, m * k) * np.exp2(rng.integers(-16, 17, m * k))
    b = rng.uniform(-1, 1, k * n) * np.exp2(rng.integers(-16, 17, k * n))
    c = np.zeros(m * n)
    got = run(torch_cuda, "N", "N", m, n, k, 1.0, a, m, b, k, 0.0, c, m, 3, nmod)
    hi, lo = dd_ref(torch_cuda, "N", "N", m, n, k, 1.0, a, m, b, k, 0.0, c, m)
    err = np.abs((got - hi) - lo)
    A = a.reshape(k, m).T
    B = b.reshape(n, k).T
    na = np.linalg.norm(A, axis=1)
    nb = np.linalg.norm(B, axis=0)
    alpha = crt_alpha(nmod, k)
    bound = 4 * math.sqrt(k) * 2.0**-alpha * np.outer(na, nb).ravel(order="F") + 2.0**-52 * np.abs(hi)
    assert np.all(err <= bound)


def test_wide_range_no_worse_than_plain_fp64(torch_cuda):
    """On the wide-range workload shape (a = u 2^e, e in [-16, 16]) the 14-residue
    CRT form's and the 8-slice Ozaki I form's max errors against double-double
    are no larger than the plain FP64 MFMA DGEMM's (7 slices are not enough on
    this input: max-exponent row scaling leaves ~6 bits of range unused)."""
    from zynum_amd import synth

    n = 1024
    a = synth.wide_range(7, 0, n, n, n)
    b = synth.wide_range(7, n * n, n, n, n)
    c = np.zeros(n * n)
    hi, lo = dd_ref(torch_cuda, "N", "N", n, n, n, 1.0, a, n, b, n, 0.0, c, n)
    plain = run(torch_cuda, "N", "N", n, n, n, 1.0, a, n, b, n, 0.0, c, n, 0, 0)
    crt = run(torch_cuda, "N", "N", n, n, n, 1.0, a, n, b, n, 0.0, c, n, 3, 14)
    sl8 = run(torch_cuda, "N", "N", n, n, n, 1.0, a, n, b, n, 0.0, c, n, 1, 8)
    e_plain = np.max(np.abs((plain - hi) - lo))
    e_crt = np.max(np.abs((crt - hi) - lo))
    e_sl8 = np.max(np.abs((sl8 - hi) - lo))
    print(f"max err vs dd: plain {e_plain:.3e}  crt14 {e_crt:.3e}  slices8 {e_sl8:.3e}")
    assert e_crt <= e_plain
    assert e_sl8 <= e_plain


def test_slices_error_decreases_with_slices(torch_cuda):
    """Ozaki I: each extra 7-bit slice cuts the error by roughly 2^7."""
    rng = np.random.default_rng(4)
    m = n = 128
    k = 512
    a = rng.uniform(-1, 1, m * k)
    b = rng.uniform(-1, 1, k * n)
    c = np.zeros(m * n)
    hi, lo = dd_ref(torch_cuda, "N", "N", m, n, k, 1.0, a, m, b, k, 0.0, c, m)
    errs = []
    for s in (2, 3, 4, 5):
        got = run(torch_cuda, "N", "N", m, n, k, 1.0, a, m, b, k, 0.0, c, m, 1, s)
        errs.append(np.max(np.abs((got - hi) - lo)))
    for e0, e1 in zip(errs, errs[1:]):
        assert e1 < e0 / 16


def mismatch_report(got, want):
    bad = np.argwhere(got != want)
    if bad.size == 0:
        return ""
    idx = [tuple(x) for x in bad[:8]]
    vals = [(got[i], want[i]) for i in idx]
    return f"{len(bad)} mismatches of {got.size}; rows {sorted(set(int(x[0]) for x in bad))[:12]} " \
           f"cols {sorted(set(int(x[1]) for x in bad))[:12]}; first {idx} {vals}"


@pytest.mark.parametrize("shape", [(301, 301, 77), (108, 64, 192), (64, 64, 128), (300, 64, 64), (257, 129, 130),
                                   (512, 384, 77)])
@pytest.mark.parametrize("ta,tb", [("N", "T"), ("N", "N"), ("T", "N")])
def test_crt_integer_exact_shapes(torch_cuda, shape, ta, tb):
    """CRT exactness across tile counts, K below/above one 128-deep slice and
    ragged edges (diagnostic message lists the mismatching rows/columns)."""
    m, n, k = shape
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    A = rng.integers(-9, 10, (m, k)).astype(np.float64)
    B = rng.integers(-9, 10, (k, n)).astype(np.float64)
    a = (A if ta == "N" else A.T).T.ravel()  # column-major storage of the stored matrix
    b = (B if tb == "N" else B.T).T.ravel()
    lda = m if ta == "N" else k
    ldb = k if tb == "N" else n
    c = np.zeros(m * n)
    got = run(torch_cuda, ta, tb, m, n, k, 1.0, a, lda, b, ldb, 0.0, c, m, 3, 14).reshape(n, m).T
    want = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
    assert np.array_equal(got, want), mismatch_report(got, want)


@pytest.mark.parametrize("ta,tb,beta", [("N", "N", 0.0), ("T", "N", 2.0), ("N", "T", -1.0)])
def test_crt_parts_integer_exact(torch_cuda, ta, tb, beta):
    """N of at least 4096: the CRT product runs in column parts on two
    streams (the second part's operand preparation beside the first part's
    products, each part's reconstruction beside the next part's products),
    each part at its own offsets; a ragged last part, transposes, and beta
    reading C."""
    m, n, k = 4400, 4300, 150
    rng = np.random.default_rng(4400 + int(beta))
    A = rng.integers(-9, 10, (m, k)).astype(np.float64)
    B = rng.integers(-9, 10, (k, n)).astype(np.float64)
    C0 = rng.integers(-50, 51, (m, n)).astype(np.float64)
    a = (A if ta == "N" else A.T).T.ravel()
    b = (B if tb == "N" else B.T).T.ravel()
    lda = m if ta == "N" else k
    ldb = k if tb == "N" else n
    ldc = m + 3
    c = np.zeros((n, ldc))
    c[:, :m] = C0.T
    got = run(torch_cuda, ta, tb, m, n, k, 1.0, a, lda, b, ldb, beta, c.ravel(), ldc, 3, 14).reshape(n, ldc)
    want = A @ B + beta * C0  # integers below 2^53: exact
    assert np.array_equal(got[:, :m].T, want), mismatch_report(got[:, :m].T, want)
    assert np.array_equal(got[:, m:], c[:, m:])


def test_workspace_is_stream_ordered(torch_cuda):
    """Two asynchronous split-precision calls on different streams share the
    library workspace: the second waits on the device for the first (no host
    synchronisation in between), so both results are exact on integer inputs;
    then a third call on the first stream after the second."""
    from zynum_amd import _lib, blas

    torch = torch_cuda
    m = n = k = 2048
    g = torch.Generator(device="cuda").manual_seed(11)
    ops = [(torch.randint(-9, 10, (k, m), generator=g, device="cuda").double(),
            torch.randint(-9, 10, (n, k), generator=g, device="cuda").double()) for _ in range(3)]
    outs = [torch.zeros(n * m, dtype=torch.float64, device="cuda") for _ in range(3)]
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    for (a, b), c, s in zip(ops, outs, (s1, s2, s1)):
        # column-major A (m x k) is a.T's storage: a is (k, m) row-major
        st = blas.gemm_request(False, False, m, n, k, 1.0, a.data_ptr(), m, b.data_ptr(), k, 0.0, c.data_ptr(), m,
                               _lib.F64, stream=s.cuda_stream, mode=3, mode_param=14, flags=_lib.FLAG_ASYNC)
        _lib.check(st, "crt async")
    torch.cuda.synchronize()
    for (a, b), c in zip(ops, outs):
        want = (b @ a).reshape(-1)  # (n, m) row-major = column-major C
        assert torch.equal(c, want)


I8_TILES = 26
BF16_TILES = 4


@pytest.mark.parametrize("tile", range(1, I8_TILES + 1))
@pytest.mark.parametrize("mode,param", [(3, 14), (1, 4)])
def test_every_int8_tile_exact(torch_cuda, tile, mode, param):
    """Every INT8 tile configuration (16x16x64 and 32x32x32 MFMA forms, BK 64
    and 128) gives the exact integer product on ragged shapes spanning several
    tiles in both directions; a configuration past the last one is declined."""
    m, n, k = 523, 389, 300
    rng = np.random.default_rng(tile * 10 + mode)
    A = rng.integers(-9, 10, (m, k)).astype(np.float64)
    B = rng.integers(-9, 10, (k, n)).astype(np.float64)
    c = rng.integers(-9, 10, m * n).astype(np.float64)
    got = run(torch_cuda, "N", "N", m, n, k, 1.0, A.T.ravel(), m, B.T.ravel(), k, 2.0, c, m, mode, param,
              tile).reshape(n, m).T
    want = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64) + 2.0 * c.reshape(n, m).T
    assert np.array_equal(got, want), mismatch_report(got, want)


def test_int8_tile_past_last_declined(torch_cuda):
    from zynum_amd import _lib, blas

    a = torch_cuda.ones(64 * 64, dtype=torch_cuda.float64, device="cuda")
    st = blas.gemm_request(False, False, 64, 64, 64, 1.0, a.data_ptr(), 64, a.data_ptr(), 64, 0.0, a.data_ptr(), 64,
                           _lib.F64, mode=3, mode_param=14, tile=I8_TILES + 1)
    assert st == _lib.DECLINE_UNSUPPORTED


@pytest.mark.parametrize("tile", range(0, BF16_TILES + 1))
def test_bf16_slices_bitwise_equal_int8_slices(torch_cuda, tile):
    """BF16 slices hold the same integers as INT8 slices, FP32 sums of up to
    4096 slice products are exact and are flushed into FP64, so the BF16 form
    equals the INT8 form bitwise on any input (K = 4500 crosses a flush)."""
    m, n, k = 300, 260, 4500
    rng = np.random.default_rng(40 + tile)
    a = rng.uniform(-1, 1, m * k) * np.exp2(rng.integers(-16, 17, m * k))
    b = rng.uniform(-1, 1, k * n) * np.exp2(rng.integers(-16, 17, k * n))
    c = rng.uniform(-1, 1, m * n)
    want = run(torch_cuda, "N", "T", m, n, k, 1.25, a, m, b, n, 0.5, c, m, 1, 6)
    got = run(torch_cuda, "N", "T", m, n, k, 1.25, a, m, b, n, 0.5, c, m, 2, 6, tile)
    assert np.array_equal(got, want), mismatch_report(got.reshape(n, m), want.reshape(n, m))


def test_bf16_tile_past_last_declined(torch_cuda):
    from zynum_amd import _lib, blas

    a = torch_cuda.ones(64 * 64, dtype=torch_cuda.float64, device="cuda")
    st = blas.gemm_request(False, False, 64, 64, 64, 1.0, a.data_ptr(), 64, a.data_ptr(), 64, 0.0, a.data_ptr(), 64,
                           _lib.F64, mode=2, mode_param=4, tile=BF16_TILES + 1)
    assert st == _lib.DECLINE_UNSUPPORTED


def test_k_limits_of_integer_accumulation(torch_cuda):
    """INT32 headroom: CRT declines K >= 2^17 and s-slice INT8 declines
    s K >= 2^19 (before writing C); BF16 slices flush FP32 into FP64 and stay
    exact at K = 2^16 + 3 with every slice product at its maximum."""
    from zynum_amd import _lib

    m, n = 4, 3
    for mode, param, k in ((3, 14, 1 << 17), (1, 8, 1 << 16)):
        a = np.ones(m * k)
        b = np.ones(k * n)
        c = np.full(m * n, 7.0)
        with pytest.raises(_lib.ZynumGpuError) as ei:
            run(torch_cuda, "N", "N", m, n, k, 1.0, a, m, b, k, 0.0, c, m, mode, param)
        assert ei.value.code == _lib.DECLINE_UNSUPPORTED
    k = (1 << 16) + 3
    x = 127.0 / 128.0  # slices 64, -64: every slice product at the 2^12 maximum magnitude
    a = np.full(m * k, x)
    b = np.full(k * n, x)
    want = float(k) * x * x  # exact: k * 16129 / 2^14
    for mode, param in ((2, 8), (1, 7)):
        got = run(torch_cuda, "N", "N", m, n, k, 1.0, a, m, b, k, 0.0, np.zeros(m * n), m, mode, param)
        assert np.all(got == want), (mode, got[:3], want)


@pytest.mark.parametrize("mode,param", [(3, 14), (1, 7), (2, 7)])
def test_non_finite_rows_and_columns_give_nan(torch_cuda, mode, param):
    """A NaN in a row of A or an infinity in a column of B makes that row /
    column of the result NaN (FP64 would give NaN or +-inf there); every other
    entry is unaffected and, on integer inputs, exact."""
    m, n, k = 70, 50, 90
    rng = np.random.default_rng(21)
    A = rng.integers(-9, 10, (m, k)).astype(np.float64)
    B = rng.integers(-9, 10, (k, n)).astype(np.float64)
    A[3, 5] = np.nan
    B[7, 11] = np.inf
    c = np.zeros(m * n)
    got = run(torch_cuda, "N", "N", m, n, k, 1.0, A.T.ravel(), m, B.T.ravel(), k, 0.0, c, m, mode, param).reshape(n, m).T
    assert np.all(np.isnan(got[3, :])) and np.all(np.isnan(got[:, 11]))
    mask = np.ones((m, n), bool)
    mask[3, :] = False
    mask[:, 11] = False
    A[3, 5] = 0
    B[7, 11] = 0
    want = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
    assert np.array_equal(got[mask], want[mask])
