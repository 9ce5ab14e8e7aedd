"""GPU parity of the complex GEMM (CGEMM / ZGEMM on the real MFMA core).

Bars (stated per test): bitwise on integer-valued inputs (every summation order
is exact there), the reference's own tolerances on its sample values
(fortran_compat_test.zig: 1e-12 / 1e-4 on the small case, 1e-9 / 4e-2 on the
padded cases), and the componentwise bound
|C_gpu - C_oracle| <= 4 (2K + 4) u (|alpha| |op(A)| |op(B)| + |beta| |C0|)
on random inputs (u = 2^-53 for ZGEMM, 2^-24 for CGEMM): both sides are a
rounded sum of the same 4K real products, regrouped.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import XorShift, cmat, cop, fill_padded_complex, int_cgemm

pytestmark = pytest.mark.gpu
PAIRS = [(a, b) for a in "NTC" for b in "NTC"]
I = C.c_int32


def gpu_cgemm(torch, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, via="fortran", tile=0, flags=0):
    """One complex GEMM on device copies of the numpy operands; returns C."""
    from zynum_amd import _lib, blas

    dt = a.dtype
    da, db, dc = (torch.from_numpy(x.copy()).cuda() for x in (a, b, c))
    lib = _lib.load()
    if via == "request":
        st = blas.gemm_request(ta, tb, m, n, k, alpha, da.data_ptr(), lda, db.data_ptr(), ldb, beta, dc.data_ptr(),
                               ldc, _lib.C64 if dt == np.complex128 else _lib.C32, tile=tile, flags=flags)
        _lib.check(st, "complex gemm")
    else:
        al = np.array([alpha], dtype=dt)
        be = np.array([beta], dtype=dt)
        fn = lib.zgemm_ if dt == np.complex128 else lib.cgemm_
        fn(ta.encode(), tb.encode(), C.byref(I(m)), C.byref(I(n)), C.byref(I(k)), al.ctypes.data, da.data_ptr(),
           C.byref(I(lda)), db.data_ptr(), C.byref(I(ldb)), be.ctypes.data, dc.data_ptr(), C.byref(I(ldc)))
        assert lib.zynum_gpu_last_status() == 0
    torch.cuda.synchronize()
    return dc.cpu().numpy()


def _dims(ta, tb, m, n, k):
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    return ar, ac, br, bc


def _bound(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, u):
    ar, ac, br, bc = _dims(ta, tb, m, n, k)
    A = np.abs(cop(cmat(a, ar, ac, lda), ta)).astype(np.float64)
    B = np.abs(cop(cmat(b, br, bc, ldb), tb)).astype(np.float64)
    mag = abs(alpha) * (A @ B)
    if beta != 0:
        mag = mag + abs(beta) * np.abs(cmat(c, m, n, ldc)).astype(np.float64)
    return 4.0 * (2 * k + 4) * u * mag


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 1e-4)])
@pytest.mark.parametrize("via", ["fortran", "request"])
def test_known_answer_no_trans(torch_cuda, oracle, dtype, tol, via):
    """fortran_compat_test.zig:732-781: complex alpha and beta, 2 x 2 x 3."""
    a = np.array([1 + 1j, 2 - 1j, -3 + 0.5j, 4 + 2j, 0.25 - 2j, -1 + 1j], dtype=dtype)
    b = np.array([2, -1 + 1j, 0.5 - 0.5j, 1 - 2j, 3 + 1j, -2 + 0.25j], dtype=dtype)
    c = np.array([1 + 0.5j, -2 + 1j, -1j, 3 + 2j], dtype=dtype)
    alpha, beta = 0.75 - 0.25j, -0.5 + 0.25j
    want = c.copy()
    (oracle.zgemm if dtype == np.complex128 else oracle.cgemm)("N", "N", 2, 2, 3, alpha, a, 2, b, 3, beta, want, 2)
    got = gpu_cgemm(torch_cuda, "N", "N", 2, 2, 3, alpha, a, 2, b, 3, beta, c, 2, via=via)
    assert np.max(np.abs(got.real - want.real)) <= tol and np.max(np.abs(got.imag - want.imag)) <= tol
    assert np.array_equal(got, want)  # dyadic inputs: exact on both sides


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("ta,tb", PAIRS)
@pytest.mark.parametrize("shape", [(33, 35, 128), (127, 129, 31), (127, 129, 33), (1, 1024, 128), (1024, 1, 128),
                                   (16, 16, 1)])
def test_integer_exact_all_transposes(torch_cuda, oracle, dtype, ta, tb, shape):
    """The reference's transpose-pair, tail and vector-edge shapes
    (fortran_compat_test.zig:864-996, padded storage lda = rows + 4,
    ldb = rows + 5, ldc = m + 6) on integer-valued inputs with integer complex
    alpha / beta: GPU == oracle == exact bit for bit, padding untouched."""
    m, n, k = shape
    ar, ac, br, bc = _dims(ta, tb, m, n, k)
    lda, ldb, ldc = ar + 4, br + 5, m + 6
    a = fill_padded_complex(ar, ac, lda, 37, dtype, integer=True)
    b = fill_padded_complex(br, bc, ldb, 41, dtype, integer=True)
    c = fill_padded_complex(m, n, ldc, 43, dtype, integer=True)
    alpha, beta = 2 - 1j, -1 + 3j
    if dtype == np.complex64 and k > 64:
        alpha = 1 - 1j  # keep every FP32 partial sum below 2^24
    want = int_cgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
    ref = c.copy()
    (oracle.zgemm if dtype == np.complex128 else oracle.cgemm)(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, ref,
                                                               ldc)
    assert np.array_equal(ref, want)
    got = gpu_cgemm(torch_cuda, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-9), (np.complex64, 4e-2)])
@pytest.mark.parametrize("ta,tb", PAIRS)
def test_reference_padded_cases_within_tolerance(torch_cuda, oracle, dtype, tol, ta, tb):
    """fortran_compat_test.zig:834-891 on the reference's sample values
    (alpha 1, beta 0, m 33, n 35, k 128): the reference's tolerance and the
    componentwise bound; padding keeps its sentinel bit for bit."""
    m, n, k = 33, 35, 128
    ar, ac, br, bc = _dims(ta, tb, m, n, k)
    lda, ldb, ldc = ar + 4, br + 5, m + 6
    a = fill_padded_complex(ar, ac, lda, 37, dtype)
    b = fill_padded_complex(br, bc, ldb, 41, dtype)
    c = fill_padded_complex(m, n, ldc, 43, dtype)
    want = c.copy()
    (oracle.zgemm if dtype == np.complex128 else oracle.cgemm)(ta, tb, m, n, k, 1.0, a, lda, b, ldb, 0.0, want, ldc)
    got = gpu_cgemm(torch_cuda, ta, tb, m, n, k, 1.0, a, lda, b, ldb, 0.0, c, ldc)
    g, w = cmat(got, m, n, ldc), cmat(want, m, n, ldc)
    assert np.max(np.abs(g.real - w.real)) <= tol and np.max(np.abs(g.imag - w.imag)) <= tol
    u = 2.0 ** -53 if dtype == np.complex128 else 2.0 ** -24
    bound = _bound(ta, tb, m, n, k, 1.0, a, lda, b, ldb, 0.0, c, ldc, u)
    assert (np.abs(g.real - w.real) <= bound).all() and (np.abs(g.imag - w.imag) <= bound).all()
    pad = np.ones(ldc * n, dtype=bool)
    pad.reshape(n, ldc)[:, :m] = False
    assert np.array_equal(got[pad], c[pad])


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("C", "N"), ("N", "C"), ("T", "C")])
@pytest.mark.parametrize("beta", [0.0, 1.0, -0.5 + 0.25j])
def test_random_complex_scalars_bound(torch_cuda, oracle, dtype, ta, tb, beta):
    """Random inputs, complex alpha, real and complex beta (the two epilogue
    routes): within the componentwise bound of the oracle."""
    m, n, k = 150, 97, 203
    rng = np.random.default_rng(17)
    ar, ac, br, bc = _dims(ta, tb, m, n, k)
    lda, ldb, ldc = ar + 3, br + 1, m + 2

    def rnd(size):
        return (rng.uniform(-1, 1, size) + 1j * rng.uniform(-1, 1, size)).astype(dtype)

    a, b, c = rnd(lda * ac), rnd(ldb * bc), rnd(ldc * n)
    alpha = 0.6 - 0.4j
    want = c.copy()
    (oracle.zgemm if dtype == np.complex128 else oracle.cgemm)(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, want,
                                                               ldc)
    got = gpu_cgemm(torch_cuda, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, via="request")
    u = 2.0 ** -53 if dtype == np.complex128 else 2.0 ** -24
    bound = _bound(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, u)
    g, w = cmat(got, m, n, ldc), cmat(want, m, n, ldc)
    assert (np.abs(g.real - w.real) <= bound).all() and (np.abs(g.imag - w.imag) <= bound).all()


def test_inactive_nans_isolated(torch_cuda):
    """fortran_compat_test.zig:1047-1074 through zgemm_."""
    nan = float("nan")
    got = gpu_cgemm(torch_cuda, "N", "N", 1, 1, 1, 1.0, np.array([2 + 1j]), 1, np.array([3 - 1j]), 1, 0.0,
                    np.array([complex(nan, nan)]), 1)
    assert got[0] == 7 + 1j
    got = gpu_cgemm(torch_cuda, "N", "N", 1, 1, 1, 0.0, np.array([complex(nan, nan)]), 1,
                    np.array([complex(nan, nan)]), 1, 2 - 1j, np.array([4 + 3j]), 1)
    assert got[0] == 11 + 2j


def test_k_zero_scales_c(torch_cuda):
    c = np.array([1 + 2j, 3 - 1j, 0.5j])
    got = gpu_cgemm(torch_cuda, "N", "N", 3, 1, 0, 1.0, np.zeros(1, dtype=np.complex128), 3,
                    np.zeros(1, dtype=np.complex128), 1, 1j, c, 3)
    assert np.array_equal(got, 1j * c)
    nan = float("nan")
    got = gpu_cgemm(torch_cuda, "N", "N", 3, 1, 0, 1.0, np.zeros(1, dtype=np.complex128), 3,
                    np.zeros(1, dtype=np.complex128), 1, 0.0, np.full(3, complex(nan, nan)), 3)
    assert np.array_equal(got, np.zeros(3, dtype=np.complex128))


def test_cblas_row_major_cgemm_known_answer(torch_cuda, zlib):
    """cblas_compat_test.zig:461-488: row-major 2 x 3 times 3 x 2, complex scalars."""
    left = np.array([1 + 1j, -3 + 0.5j, 0.25 - 2j, 2 - 1j, 4 + 2j, -1 + 1j], dtype=np.complex64)
    right = np.array([2, 1 - 2j, -1 + 1j, 3 + 1j, 0.5 - 0.5j, -2 + 0.25j], dtype=np.complex64)
    res = np.array([1 + 0.5j, -1j, -2 + 1j, 3 + 2j], dtype=np.complex64)
    alpha = np.array([0.75 - 0.25j], dtype=np.complex64)
    beta = np.array([-0.5 + 0.25j], dtype=np.complex64)
    want = alpha[0] * (left.reshape(2, 3).astype(np.complex128) @ right.reshape(3, 2)) + beta[0] * res.reshape(2, 2)
    d = [torch_cuda.from_numpy(x.copy()).cuda() for x in (left, right, res)]
    zlib.cblas_cgemm(101, 111, 111, 2, 2, 3, alpha.ctypes.data, d[0].data_ptr(), 3, d[1].data_ptr(), 2,
                     beta.ctypes.data, d[2].data_ptr(), 2)
    torch_cuda.cuda.synchronize()
    got = d[2].cpu().numpy().reshape(2, 2)
    assert np.max(np.abs(got - want)) <= 1e-4


def _fill_row_major(rng, rows, cols, ld):
    """reference.zig:36-38,141-146: sentinel -57 - i + 29i, then rng.scalar()
    (re, im drawn in that order) row by row."""
    out = np.array([complex(-57.0 - i, 29.0) for i in range(rows * ld)])
    for i in range(rows):
        for j in range(cols):
            re = rng.small()
            out[i * ld + j] = complex(re, rng.small())
    return out


def _row_major_case(torch, zlib, transa, transb, m, n, k, lda, ldb, ldc, seed, alpha, beta, tol):
    ta = {111: "N", 112: "T", 113: "C"}[transa]
    tb = {111: "N", 112: "T", 113: "C"}[transb]
    a_rows, a_cols = (m, k) if ta == "N" else (k, m)
    b_rows, b_cols = (k, n) if tb == "N" else (n, k)
    rng = XorShift(seed)
    a = _fill_row_major(rng, a_rows, a_cols, lda)
    b = _fill_row_major(rng, b_rows, b_cols, ldb)
    c = _fill_row_major(rng, m, n, ldc)
    al, be = np.array([alpha]), np.array([beta])
    A = cop(a.reshape(a_rows, lda)[:, :a_cols], ta)
    B = cop(b.reshape(b_rows, ldb)[:, :b_cols], tb)
    want = c.copy().reshape(m, ldc)
    want[:, :n] = alpha * (A @ B) + beta * want[:, :n]
    d = [torch.from_numpy(x.copy()).cuda() for x in (a, b, c)]
    zlib.cblas_zgemm(101, transa, transb, m, n, k, al.ctypes.data, d[0].data_ptr(), lda, d[1].data_ptr(), ldb,
                     be.ctypes.data, d[2].data_ptr(), ldc)
    torch.cuda.synchronize()
    got = d[2].cpu().numpy().reshape(m, ldc)
    assert np.max(np.abs(got[:, :n] - want[:, :n])) <= tol
    assert np.array_equal(got[:, n:], want[:, n:])


def test_cblas_row_major_zgemm_conj_trans_reference_case(torch_cuda, zlib):
    """cblas_compat_test.zig:790-813: ConjTrans x NoTrans, m 2, n 3, k 4,
    lda 4, ldb 5, ldc 5, seed 0x5eed0112, alpha 0.6 - 0.4i, beta -0.2 + 0.3i."""
    _row_major_case(torch_cuda, zlib, 113, 111, 2, 3, 4, 4, 5, 5, 0x5EED0112, 0.6 - 0.4j, -0.2 + 0.3j, 1e-12)


@pytest.mark.parametrize("transa,transb", [(a, b) for a in (111, 112, 113) for b in (111, 112, 113)])
def test_cblas_row_major_zgemm_padded(torch_cuda, zlib, transa, transb):
    """cblas_compat_test.zig:368-400: row-major with lda = a_cols + 4,
    ldb = b_cols + 5, ldc = n + 6 and the reference's xorshift inputs; the
    padding of C comes back unchanged."""
    m, n, k = 5, 7, 9
    a_cols = k if transa == 111 else m
    b_cols = n if transb == 111 else k
    _row_major_case(torch_cuda, zlib, transa, transb, m, n, k, a_cols + 4, b_cols + 5, n + 6, 0xC31A5EED,
                    0.6 - 0.4j, -0.2 + 0.3j, 1e-12)


def test_host_operands_are_staged(torch_cuda, oracle):
    """Host (numpy) operands go through the staging path of the same kernels."""
    from zynum_amd import _lib

    lib = _lib.load()
    m, n, k = 40, 30, 20
    a = fill_padded_complex(m, k, m, 3, integer=True)
    b = fill_padded_complex(k, n, k, 7, integer=True)
    c = fill_padded_complex(m, n, m, 11, integer=True)
    want = int_cgemm("N", "N", m, n, k, 1 + 1j, a, m, b, k, 2.0, c, m)
    al, be = np.array([1 + 1j]), np.array([2 + 0j])
    lib.zgemm_(b"N", b"N", C.byref(I(m)), C.byref(I(n)), C.byref(I(k)), al.ctypes.data, a.ctypes.data,
               C.byref(I(m)), b.ctypes.data, C.byref(I(k)), be.ctypes.data, c.ctypes.data, C.byref(I(m)))
    assert lib.zynum_gpu_last_status() == 0
    assert np.array_equal(c, want)


def test_every_tile_config_bitwise(torch_cuda):
    """Every plain tile configuration of the real core gives the same bits on
    the expanded complex product."""
    m, n, k = 70, 90, 50
    rng = np.random.default_rng(3)
    a = rng.standard_normal(m * k) + 1j * rng.standard_normal(m * k)
    b = rng.standard_normal(k * n) + 1j * rng.standard_normal(k * n)
    c = np.zeros(m * n, dtype=np.complex128)
    from zynum_amd.autotune import PLAIN_TILES

    first = gpu_cgemm(torch_cuda, "N", "C", m, n, k, 0.5 + 1j, a, m, b, n, 0.0, c, m, via="request", tile=1)
    for tile in range(2, PLAIN_TILES + 1):
        got = gpu_cgemm(torch_cuda, "N", "C", m, n, k, 0.5 + 1j, a, m, b, n, 0.0, c, m, via="request", tile=tile)
        assert np.array_equal(got, first), tile


def test_complex_modes_other_than_plain_declined(torch_cuda):
    from zynum_amd import _lib, blas

    c = torch_cuda.zeros(4, dtype=torch_cuda.complex128, device="cuda")
    a = torch_cuda.ones(4, dtype=torch_cuda.complex128, device="cuda")
    st = blas.gemm_request("N", "N", 2, 2, 2, 1.0, a.data_ptr(), 2, a.data_ptr(), 2, 0.0, c.data_ptr(), 2, _lib.C64,
                           mode=_lib.MODE_OZAKI_CRT)
    assert st == _lib.DECLINE_UNSUPPORTED


def test_row_chunks_of_a_large_expansion(torch_cuda):
    """op(A) whose 2x2 expansion exceeds the 4 GiB pack budget is packed and
    multiplied in row chunks (16320 rows for k = 8200); rows on both sides of
    the chunk boundary and the last row are exact on integer inputs."""
    torch = torch_cuda
    m, n, k = 16400, 16, 8200
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    a = torch.complex(torch.randint(-3, 4, (k, m), device="cuda", generator=g).double(),
                      torch.randint(-3, 4, (k, m), device="cuda", generator=g).double())  # column-major m x k
    b = torch.complex(torch.randint(-3, 4, (n, k), device="cuda", generator=g).double(),
                      torch.randint(-3, 4, (n, k), device="cuda", generator=g).double())  # column-major k x n
    c = torch.zeros((n, m), dtype=torch.complex128, device="cuda")
    from zynum_amd import _lib, blas

    st = blas.gemm_request("N", "N", m, n, k, 1.0, a.data_ptr(), m, b.data_ptr(), k, 0.0, c.data_ptr(), m, _lib.C64)
    _lib.check(st, "large complex gemm")
    torch.cuda.synchronize()
    rows = [0, 16319, 16320, 16321, m - 1]
    A = a[:, rows].T.cpu().numpy()           # rows x k
    B = b.T.cpu().numpy()                     # k x n
    want = A @ B                              # integers below 2^53: exact
    got = c[:, rows].T.cpu().numpy()
    assert np.array_equal(got, want)
    del a, b, c
