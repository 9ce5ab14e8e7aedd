"""GPU parity of the plain FP64/FP32 MFMA GEMM against the oracle.

Bars (stated per test): bitwise on integer-valued inputs (every summation
order is exact there), the reference's own tolerances on its sample values,
and the normwise bound |C_gpu - C_ref| <= 2 (K + 2) u (|alpha| |A||B| +
|beta| |C0|) elementwise on random inputs (u = 2^-53 for FP64, 2^-24 for FP32).
"""
import ctypes as C
import json
from pathlib import Path

import numpy as np
import pytest

from conftest import XorShift, fill_padded, fill_row_major, int_gemm
from test_oracle import all_cases, operands

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def gpu_gemm(torch, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, mode=0, mode_param=0, via="request",
             tile=0):
    """Run one GEMM on device copies of the numpy operands; returns C (numpy)."""
    from zynum_amd import _lib, blas

    dt = a.dtype
    da, db, dc = (torch.from_numpy(x.copy()).cuda() for x in (a, b, c))
    scalar = _lib.F64 if dt == np.float64 else _lib.F32
    if via == "request":
        st = blas.gemm_request(ta != "N", tb != "N", m, n, k, alpha, da.data_ptr(), lda, db.data_ptr(), ldb, beta,
                               dc.data_ptr(), ldc, scalar, mode=mode, mode_param=mode_param, tile=tile)
        _lib.check(st, "gemm")
    else:
        lib = _lib.load()
        I = C.c_int32
        if dt == np.float64:
            lib.dgemm_(ta.encode(), tb.encode(), C.byref(I(m)), C.byref(I(n)), C.byref(I(k)),
                       C.byref(C.c_double(alpha)), da.data_ptr(), C.byref(I(lda)), db.data_ptr(), C.byref(I(ldb)),
                       C.byref(C.c_double(beta)), dc.data_ptr(), C.byref(I(ldc)))
        else:
            lib.sgemm_(ta.encode(), tb.encode(), C.byref(I(m)), C.byref(I(n)), C.byref(I(k)),
                       C.byref(C.c_float(alpha)), da.data_ptr(), C.byref(I(lda)), db.data_ptr(), C.byref(I(ldb)),
                       C.byref(C.c_float(beta)), dc.data_ptr(), C.byref(I(ldc)))
        assert lib.zynum_gpu_last_status() == 0
    torch.cuda.synchronize()
    return dc.cpu().numpy()


def logical(x, rows, cols, ld):
    return x[: ld * cols].reshape(cols, ld)[:, :rows].T


def normwise_bound(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, u):
    A = logical(a, m, k, lda) if ta == "N" else logical(a, k, m, lda).T
    B = logical(b, k, n, ldb) if tb == "N" else logical(b, n, k, ldb).T
    mag = abs(alpha) * (np.abs(A).astype(np.float64) @ np.abs(B).astype(np.float64))
    if beta != 0:
        mag = mag + abs(beta) * np.abs(logical(c, m, n, ldc)).astype(np.float64)
    return 2.0 * (k + 2) * u * mag


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("shape", [(16, 16, 4), (16, 16, 16), (33, 17, 9), (128, 128, 16), (130, 129, 37)])
def test_mfma_lane_layout_integer_exact(torch_cuda, ta, tb, shape):
    """Small integer GEMMs pin the MFMA operand/result lane layout bitwise."""
    m, n, k = shape
    lda = (m if ta == "N" else k) + 1
    ldb = (k if tb == "N" else n) + 3
    ldc = m + 2
    a, b, c = operands(ta, tb, m, n, k, lda, ldb, ldc, np.float64, True, salts=(2, 6, 10))
    want = int_gemm(ta, tb, m, n, k, 1.0, a, lda, b, ldb, 0.0, c, ldc)
    got = gpu_gemm(torch_cuda, ta, tb, m, n, k, 1.0, a, lda, b, ldb, 0.0, c, ldc)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("case", all_cases())
def test_reference_cases_integer_bitwise(torch_cuda, oracle, dtype, case):
    """The reference's padded/tail/alpha-beta shapes on integer inputs: GPU ==
    oracle == exact, bitwise, padding untouched (fortran_compat_test.zig:353-386)."""
    ta, tb, m, n, k, lda, ldb, ldc, al, be = case
    if al == 0.001:
        al = 0.125
    if dtype == np.float32 and k > 1024:
        pytest.skip("FP32 integer sums exceed 2^24 at this depth")
    a, b, c = operands(ta, tb, m, n, k, lda, ldb, ldc, dtype, True)
    want = c.copy()
    (oracle.dgemm if dtype == np.float64 else oracle.sgemm)(ta, tb, m, n, k, al, a, lda, b, ldb, be, want, ldc)
    got = gpu_gemm(torch_cuda, ta, tb, m, n, k, al, a, lda, b, ldb, be, c, ldc, via="fortran")
    assert np.array_equal(got, want)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("case", all_cases())
def test_reference_cases_sample_values(torch_cuda, oracle, dtype, case):
    """The reference's tolerance (1e-10 f64 / 1e-3 f32, fortran_compat_test.zig:375)
    and the normwise bound against the oracle; padding bitwise unchanged."""
    ta, tb, m, n, k, lda, ldb, ldc, al, be = case
    a, b, c = operands(ta, tb, m, n, k, lda, ldb, ldc, dtype, False)
    want = c.copy()
    (oracle.dgemm if dtype == np.float64 else oracle.sgemm)(ta, tb, m, n, k, al, a, lda, b, ldb, be, want, ldc)
    got = gpu_gemm(torch_cuda, ta, tb, m, n, k, al, a, lda, b, ldb, be, c, ldc, via="fortran")
    tol = 1e-10 if dtype == np.float64 else 1e-3
    G, W = logical(got, m, n, ldc), logical(want, m, n, ldc)
    assert np.max(np.abs(G.astype(np.float64) - W)) <= tol * max(1.0, k / 64)
    u = 2.0**-53 if dtype == np.float64 else 2.0**-24
    bound = normwise_bound(ta, tb, m, n, k, al, a, lda, b, ldb, be, c, ldc, u)
    assert np.all(np.abs(G.astype(np.float64) - W) <= bound + 1e-300)
    mask = np.ones_like(got, dtype=bool)
    logical(mask, m, n, ldc)[:] = False
    assert np.array_equal(got[mask], c[mask])


def test_known_answers_and_nan_isolation(torch_cuda):
    # fortran_compat_test.zig:537-563, 1012-1045
    a = np.array([1, 2, 3, 4, 5, 6.0])
    b = np.array([7, 8, 9, 10, 11, 12.0])
    c = np.zeros(4)
    assert gpu_gemm(torch_cuda, "N", "N", 2, 2, 3, 1.0, a, 2, b, 3, 0.0, c, 2).tolist() == [76, 100, 103, 136]
    nan = float("nan")
    a = np.array([1, 2, 3, 4.0])
    b = np.array([5, 6, 7, 8.0])
    got = gpu_gemm(torch_cuda, "N", "N", 2, 2, 2, 1.0, a, 2, b, 2, 0.0, np.full(4, nan), 2)
    assert got.tolist() == [23, 34, 31, 46]
    got = gpu_gemm(torch_cuda, "N", "N", 2, 2, 2, 0.0, np.full(4, nan), 2, np.full(4, nan), 2, 2.0,
                   np.array([1, 2, 3, 4.0]), 2)
    assert got.tolist() == [2, 4, 6, 8]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_cblas_row_major_pairs(torch_cuda, oracle, dtype, ta, tb):
    """cblas_compat_test.zig:326-364 through cblas_dgemm/cblas_sgemm (row-major)."""
    from zynum_amd import _lib

    lib = _lib.load()
    m, n, k = 35, 37, 67
    a_rows, a_cols = (m, k) if ta == "N" else (k, m)
    b_rows, b_cols = (k, n) if tb == "N" else (n, k)
    lda, ldb, ldc = a_cols + 4, b_cols + 4, n + 5
    rng = XorShift(0x74A3119D)
    a = fill_row_major(rng, a_rows, a_cols, lda, dtype)
    b = fill_row_major(rng, b_rows, b_cols, ldb, dtype)
    c = fill_row_major(rng, m, n, ldc, dtype)
    want = c.copy()
    (oracle.dgemm if dtype == np.float64 else oracle.sgemm)(tb, ta, n, m, k, -0.5, b, ldb, a, lda, 2.0, want, ldc)
    da, db, dc = (torch_cuda.from_numpy(x.copy()).cuda() for x in (a, b, c))
    tr = {"N": _lib.CblasNoTrans, "T": _lib.CblasTrans}
    fn = lib.cblas_dgemm if dtype == np.float64 else lib.cblas_sgemm
    fn(_lib.CblasRowMajor, tr[ta], tr[tb], m, n, k, -0.5, da.data_ptr(), lda, db.data_ptr(), ldb, 2.0,
       dc.data_ptr(), ldc)
    got = dc.cpu().numpy()
    tol = 1e-10 if dtype == np.float64 else 1e-3
    np.testing.assert_allclose(got, want, rtol=0, atol=tol)


def test_host_pointers_are_staged(oracle):
    """Host operands run through device staging with the same kernels."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from zynum_amd import _lib

    lib = _lib.load()
    m, n, k, lda, ldb, ldc = 70, 97, 31, 73, 37, 75
    a, b, c = operands("N", "N", m, n, k, lda, ldb, ldc, np.float64, True)
    want = c.copy()
    oracle.dgemm("N", "N", m, n, k, 0.75, a, lda, b, ldb, -0.25, want, ldc)
    got = c.copy()
    I = C.c_int32
    lib.dgemm_(b"N", b"N", C.byref(I(m)), C.byref(I(n)), C.byref(I(k)), C.byref(C.c_double(0.75)), a.ctypes.data,
               C.byref(I(lda)), b.ctypes.data, C.byref(I(ldb)), C.byref(C.c_double(-0.25)), got.ctypes.data,
               C.byref(I(ldc)))
    assert lib.zynum_gpu_last_status() == 0
    assert np.array_equal(got, want)


def test_config1_cblas_dgemm_1024(torch_cuda, oracle):
    """cblas_dgemm(ColMajor, NN, 1024^3, alpha=1.5, beta=0.5), inputs from seed 0
    (A, B, C in that order): normwise bound against the oracle, and the
    committed double-double golden entries."""
    from zynum_amd import _lib, synth

    n = 1024
    a = synth.matrix(0, 0, n, n, n)
    b = synth.matrix(0, n * n, n, n, n)
    c = synth.matrix(0, 2 * n * n, n, n, n)
    want = c.copy()
    oracle.dgemm("N", "N", n, n, n, 1.5, a, n, b, n, 0.5, want, n)
    da, db, dc = (torch_cuda.from_numpy(x.copy()).cuda() for x in (a, b, c))
    _lib.load().cblas_dgemm(_lib.CblasColMajor, _lib.CblasNoTrans, _lib.CblasNoTrans, n, n, n, 1.5, da.data_ptr(),
                            n, db.data_ptr(), n, 0.5, dc.data_ptr(), n)
    got = dc.cpu().numpy()
    bound = normwise_bound("N", "N", n, n, n, 1.5, a, n, b, n, 0.5, c, n, 2.0**-53)
    assert np.all(np.abs(logical(got, n, n, n) - logical(want, n, n, n)) <= bound)
    g = json.loads((GOLDEN / "config1_dd_samples.json").read_text())
    rows, cols = np.array(g["rows"]), np.array(g["cols"])
    hi, lo = np.array(g["hi"]), np.array(g["lo"])
    err = np.abs((got[rows + cols * n] - hi) - lo)
    assert np.all(err <= bound[rows, cols])
    # the committed oracle values at the same entries
    assert np.all(np.abs(got[rows + cols * n] - np.array(g["oracle_fp64"])) <= bound[rows, cols])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("beta", [0.0, 1.0])
def test_8192_padded_integer_sampled(torch_cuda, dtype, ta, tb, beta):
    """Full-size configuration (8192^3, ld 8200, four transposes, beta 0 and 1)
    on integer inputs in [-8, 8]: every entry is exact, so 2048 sampled entries
    must equal exact int64 dot products bitwise; padding rows untouched."""
    from zynum_amd import _lib, blas

    torch = torch_cuda
    n, ld = 8192, 8200
    g = torch.Generator(device="cuda").manual_seed(1234 + (ta == "T") * 2 + (tb == "T"))
    tdt = torch.float64 if dtype == np.float64 else torch.float32

    def mk():
        x = torch.randint(-8, 9, (n, ld), generator=g, device="cuda").to(tdt)
        x[:, n:] = -999.0
        return x  # row j of the tensor = column j of the column-major matrix

    A, B, Cm = mk(), mk(), mk()
    C0 = Cm.clone()
    st = blas.gemm_request(ta == "T", tb == "T", n, n, n, 1.0, A.data_ptr(), ld, B.data_ptr(), ld, beta,
                           Cm.data_ptr(), ld, _lib.F64 if dtype == np.float64 else _lib.F32)
    _lib.check(st)
    torch.cuda.synchronize()
    rs = np.random.default_rng(5)
    rows = torch.from_numpy(rs.integers(0, n, 2048)).cuda()
    cols = torch.from_numpy(rs.integers(0, n, 2048)).cuda()
    Ai = A[:, :n].to(torch.int64)  # Ai[p, i] = A(i, p) column-major
    Bi = B[:, :n].to(torch.int64)
    arow = Ai[:, rows].T if ta == "N" else Ai[rows, :]        # op(A)(row, :)
    bcol = Bi[cols, :] if tb == "N" else Bi[:, cols].T         # op(B)(:, col)
    exact = (arow * bcol).sum(dim=1).to(torch.float64)
    if beta != 0:
        exact = exact + C0[cols, rows].to(torch.float64)
    got = Cm[cols, rows].to(torch.float64)
    assert torch.equal(got, exact)
    assert torch.all(Cm[:, n:] == -999.0)


PLAIN_TILES = 12


@pytest.mark.parametrize("tile", range(1, PLAIN_TILES + 1))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "T")])
def test_every_tile_config_bitwise(torch_cuda, tile, dtype, ta, tb):
    """Every FP64/FP32 tile configuration gives the exact product on integer
    inputs over a ragged shape spanning several tiles (padded ld, alpha/beta),
    so the tuned tile choice cannot change results."""
    m, n, k = 301, 263, 150
    lda = (m if ta == "N" else k) + 3
    ldb = (k if tb == "N" else n) + 1
    ldc = m + 2
    a, b, c = operands(ta, tb, m, n, k, lda, ldb, ldc, dtype, True, salts=(5, 9, 2))
    want = int_gemm(ta, tb, m, n, k, -0.5, a, lda, b, ldb, 2.0, c, ldc)
    got = gpu_gemm(torch_cuda, ta, tb, m, n, k, -0.5, a, lda, b, ldb, 2.0, c, ldc, tile=tile)
    assert np.array_equal(got, want.astype(dtype))


DIRECT_LOAD_TILE = 12  # 1-based request tile of the 16-wave 128x128 configuration


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("shape,pads", [
    ((128, 128, 8), (2, 2)),        # one tile, one 8-k stage (direct loads)
    ((256, 384, 1024), (2, 6)),     # direct loads, 128 stages
    ((1024, 640, 2048), (2, 2)),    # direct loads, several tiles per XCD
    ((256, 384, 1020), (2, 6)),     # k % 8 = 4: register-staged tile
    ((256, 384, 1024), (1, 6)),     # odd lda: register-staged tile
    ((1000, 384, 512), (2, 2)),     # m not a multiple of 128: register-staged tile
])
def test_f64_direct_load_tile_integer_exact(torch_cuda, ta, tb, shape, pads):
    """The FP64 16-wave tile fed by direct-to-LDS loads (full tiles, k a
    multiple of 8, even leading dimensions; each transpose pair keeps its own
    LDS layout) and the shapes just outside it give the exact product on
    integer inputs with alpha -1/2, beta 2."""
    m, n, k = shape
    lda = (m if ta == "N" else k) + pads[0]
    ldb = (k if tb == "N" else n) + pads[1]
    ldc = m + 1
    a, b, c = operands(ta, tb, m, n, k, lda, ldb, ldc, np.float64, True, salts=(3, 8, 1))
    want = int_gemm(ta, tb, m, n, k, -0.5, a, lda, b, ldb, 2.0, c, ldc)
    got = gpu_gemm(torch_cuda, ta, tb, m, n, k, -0.5, a, lda, b, ldb, 2.0, c, ldc, tile=DIRECT_LOAD_TILE)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "T")])
def test_f64_direct_load_needs_16_byte_aligned_operands(torch_cuda, ta, tb):
    """A and B starting one double into their buffers (8-byte aligned) fall
    back to the register-staged tile, under the automatic choice and with the
    direct-load configuration forced; both exact on integer inputs."""
    from zynum_amd import _lib, blas

    torch = torch_cuda
    m, n, k = 256, 256, 512
    lda = (m if ta == "N" else k) + 2
    ldb = (k if tb == "N" else n) + 2
    ldc = m + 2
    g = torch.Generator().manual_seed(3)
    A = torch.randint(-8, 9, (lda * (k if ta == "N" else m) + 1,), generator=g).to(torch.float64).cuda()
    B = torch.randint(-8, 9, (ldb * (n if tb == "N" else k) + 1,), generator=g).to(torch.float64).cuda()
    a, b = A[1:].cpu().numpy(), B[1:].cpu().numpy()
    c = np.zeros(ldc * n)
    want = int_gemm(ta, tb, m, n, k, 1.0, a, lda, b, ldb, 0.0, c, ldc)
    assert (A.data_ptr() + 8) % 16 == 8
    for tile in (0, DIRECT_LOAD_TILE):
        C_ = torch.zeros(ldc * n, dtype=torch.float64, device="cuda")
        st = blas.gemm_request(ta == "T", tb == "T", m, n, k, 1.0, A.data_ptr() + 8, lda, B.data_ptr() + 8, ldb, 0.0,
                               C_.data_ptr(), ldc, _lib.F64, tile=tile)
        _lib.check(st, "gemm")
        torch.cuda.synchronize()
        assert np.array_equal(C_.cpu().numpy(), want)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_f64_direct_load_tile_equals_register_staged(torch_cuda, ta, tb):
    """On random FP64 inputs the direct-load form gives bitwise the result of
    the register-staged tile (same MFMA order): the same matrices stored with
    odd leading dimensions take the register-staged kernel."""
    m, n, k = 512, 384, 1032
    rs = np.random.default_rng(17)
    A = rs.uniform(-1, 1, (m, k))
    B = rs.uniform(-1, 1, (k, n))
    Cm = rs.uniform(-1, 1, (m, n))
    opa = A if ta == "N" else A.T.copy()  # the stored matrix (column-major)
    opb = B if tb == "N" else B.T.copy()
    outs = []
    for pad in (2, 1):  # even leading dimensions: direct loads; odd: register staging
        lda, ldb, ldc = opa.shape[0] + pad, opb.shape[0] + pad, m + 3
        a = np.zeros(lda * opa.shape[1]); a.reshape(opa.shape[1], lda)[:, : opa.shape[0]] = opa.T
        b = np.zeros(ldb * opb.shape[1]); b.reshape(opb.shape[1], ldb)[:, : opb.shape[0]] = opb.T
        c = np.zeros(ldc * n); c.reshape(n, ldc)[:, :m] = Cm.T
        got = gpu_gemm(torch_cuda, ta, tb, m, n, k, 1.25, a, lda, b, ldb, -0.75, c, ldc, tile=DIRECT_LOAD_TILE)
        outs.append(logical(got, m, n, ldc).copy())
    assert np.array_equal(outs[0], outs[1])
    ref = 1.25 * (A @ B) - 0.75 * Cm
    assert np.max(np.abs(outs[0] - ref)) < 2 * (k + 2) * 2.0**-53 * np.max(1.25 * np.abs(A) @ np.abs(B) + 0.75)


def test_plain_tile_past_last_declined(torch_cuda):
    from zynum_amd import _lib, blas

    a = torch_cuda.ones(64 * 64, dtype=torch_cuda.float64, device="cuda")
    st = blas.gemm_request(False, False, 64, 64, 64, 1.0, a.data_ptr(), 64, a.data_ptr(), 64, 0.0, a.data_ptr(), 64,
                           _lib.F64, tile=PLAIN_TILES + 1)
    assert st == _lib.DECLINE_UNSUPPORTED
