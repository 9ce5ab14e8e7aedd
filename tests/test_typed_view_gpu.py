"""The typed ``matrixMultiply`` view (operations.zig:501-559) on the GPU tier.

Bars:
* the reference's own 1x1 contract cases (public_surface_contract_test.zig:
  1360-1420) for f32, f64, complex64 and complex128: defaults give 63, a
  NaN-poisoned result with the default result scale gives exact 0, explicit
  scales give 6, and the workspace form the same with a poisoned workspace;
* transposed() / adjoint() views on integer-valued inputs equal exact integer
  arithmetic bitwise, and equal the oracle's GEMM on the same storage;
* random inputs within the normwise bound of the oracle;
* the workspace form with the result aliasing an input equals the oracle run
  on copies (operations.zig:529-559 copies C out, multiplies, copies back).
Device tensors and host numpy storage (staged by the library) both run.
"""
import numpy as np
import pytest

from conftest import cop, int_cgemm, int_gemm

pytestmark = pytest.mark.gpu

DTYPES = [np.float32, np.float64, np.complex64, np.complex128]


def storage(torch, values, where):
    arr = np.ascontiguousarray(values)
    return torch.from_numpy(arr.copy()).cuda() if where == "device" else arr.copy()


def host(x):
    return x.cpu().numpy() if hasattr(x, "cpu") else x


@pytest.mark.parametrize("where", ["device", "host"])
@pytest.mark.parametrize("dtype", DTYPES)
def test_contract_one_by_one(torch_cuda, dtype, where):
    """public_surface_contract_test.zig:1360-1420 (left 9, right 7)."""
    from zynum_amd import blas

    poison = np.array([np.nan], dtype=dtype) if np.dtype(dtype).kind == "f" else np.array(
        [complex(np.nan, np.nan)], dtype=dtype)
    left = blas.ConstMatrix(storage(torch_cuda, np.array([9], dtype=dtype), where), 1, 1, 1)
    right = blas.ConstMatrix(storage(torch_cuda, np.array([7], dtype=dtype), where), 1, 1, 1)

    out = storage(torch_cuda, poison, where)
    result = blas.Matrix(out, 1, 1, 1)
    assert blas.matrix_multiply_workspace_length(result) == 1
    blas.matrix_multiply(left, right, result)
    torch_cuda.cuda.synchronize()
    assert host(out)[0] == 63, "matrixMultiply defaults"

    out = storage(torch_cuda, poison, where)
    blas.matrix_multiply(left, right, blas.Matrix(out, 1, 1, 1), product_scale=0)
    torch_cuda.cuda.synchronize()
    got = host(out)[0]
    assert got == 0 and not np.isnan(got), "matrixMultiply default result scale"

    out = storage(torch_cuda, np.array([3], dtype=dtype), where)
    blas.matrix_multiply(left, right, blas.Matrix(out, 1, 1, 1), product_scale=0, result_scale=2)
    torch_cuda.cuda.synchronize()
    assert host(out)[0] == 6, "matrixMultiply explicit scales"

    for kwargs, start, want in (({}, poison, 63), ({"product_scale": 0}, poison, 0),
                                ({"product_scale": 0, "result_scale": 2}, np.array([3], dtype=dtype), 6)):
        out = storage(torch_cuda, start, where)
        work = storage(torch_cuda, poison, where)
        blas.matrix_multiply_with_workspace(left, right, blas.Matrix(out, 1, 1, 1), work, **kwargs)
        torch_cuda.cuda.synchronize()
        got = host(out)[0]
        assert got == want and not np.isnan(got), f"matrixMultiplyWithWorkspace {kwargs}"


def view_storage(rng, rows, cols, ld, dtype, integer):
    """Column-major rows x cols operand in an ld-padded buffer (sentinel -999 padding)."""
    buf = np.full(ld * cols, -999, dtype=dtype)
    if integer:
        vals = rng.integers(-8, 9, (rows, cols)).astype(dtype)
        if np.dtype(dtype).kind == "c":
            vals = vals + 1j * rng.integers(-8, 9, (rows, cols))
    else:
        vals = rng.uniform(-1, 1, (rows, cols)).astype(dtype)
        if np.dtype(dtype).kind == "c":
            vals = vals + 1j * rng.uniform(-1, 1, (rows, cols))
    buf.reshape(cols, ld)[:, :rows] = vals.T.astype(dtype)
    return buf


OPS = {"normal": "N", "transposed": "T", "adjoint": "C"}


def make_view(blas, buf, rows, cols, ld, op):
    v = blas.ConstMatrix(buf, rows, cols, ld)
    return v if op == "normal" else (v.transposed() if op == "transposed" else v.adjoint())


def oracle_product(oracle, dtype, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    want = c.copy()
    if dtype == np.float64:
        oracle.dgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, want, ldc)
    elif dtype == np.float32:
        oracle.sgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, want, ldc)
    elif dtype == np.complex128:
        oracle.zgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, want, ldc)
    else:
        oracle.cgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, want, ldc)
    return want


@pytest.mark.parametrize("where", ["device", "host"])
@pytest.mark.parametrize("lop,rop", [("normal", "normal"), ("transposed", "normal"), ("normal", "transposed"),
                                     ("adjoint", "transposed"), ("transposed", "adjoint")])
@pytest.mark.parametrize("dtype", DTYPES)
def test_views_integer_bitwise(torch_cuda, oracle, dtype, lop, rop, where):
    """op(left) from transposed()/adjoint() views: exact integer arithmetic and the oracle agree bitwise."""
    from zynum_amd import blas

    real = np.dtype(dtype).kind == "f"
    rng = np.random.default_rng(sum(ord(ch) * 31**i for i, ch in enumerate(str(np.dtype(dtype)) + lop + rop)))
    m, n, k = 67, 45, 83
    ta, tb = OPS[lop], OPS[rop]
    if real:
        ta, tb = ("T" if ta == "C" else ta), ("T" if tb == "C" else tb)
    lr, lc = (m, k) if ta == "N" else (k, m)
    rr, rc = (k, n) if tb == "N" else (n, k)
    lda, ldb, ldc = lr + 3, rr + 1, m + 2
    a = view_storage(rng, lr, lc, lda, dtype, True)
    b = view_storage(rng, rr, rc, ldb, dtype, True)
    c = view_storage(rng, m, n, ldc, dtype, True)
    alpha, beta = (2, -1) if real else (complex(2, -1), complex(1, 1))
    exact = (int_gemm if real else int_cgemm)(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
    want = oracle_product(oracle, dtype, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
    assert np.array_equal(want, exact)

    A, B, Cb = (storage(torch_cuda, x, where) for x in (a, b, c))
    # a real adjoint() view is its transpose (views.zig:165-176)
    blas.matrix_multiply(make_view(blas, A, lr, lc, lda, lop), make_view(blas, B, rr, rc, ldb, rop),
                         blas.Matrix(Cb, m, n, ldc), product_scale=alpha, result_scale=beta)
    torch_cuda.cuda.synchronize()
    assert np.array_equal(host(Cb), want)


@pytest.mark.parametrize("dtype", DTYPES)
def test_views_random_within_bound(torch_cuda, oracle, dtype):
    """Random inputs: |C - C_oracle| <= c (K + 2) u (|s| |op(L)||op(R)| + |r| |C0|) elementwise,
    c = 2 (real) or 4 (complex, two real products and a sum per term)."""
    from zynum_amd import blas

    real = np.dtype(dtype).kind == "f"
    rng = np.random.default_rng(11)
    m, n, k = 150, 97, 211
    lop, rop = ("transposed", "normal") if real else ("adjoint", "normal")
    ta = "T" if real else "C"
    lda, ldb, ldc = k + 1, k + 2, m + 5
    a = view_storage(rng, k, m, lda, dtype, False)
    b = view_storage(rng, k, n, ldb, dtype, False)
    c = view_storage(rng, m, n, ldc, dtype, False)
    alpha, beta = (1.5, 0.5) if real else (complex(1.5, -0.25), complex(0.5, 0.5))
    want = oracle_product(oracle, dtype, ta, "N", m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
    A, B, Cb = (storage(torch_cuda, x, "device") for x in (a, b, c))
    blas.matrix_multiply(make_view(blas, A, k, m, lda, lop), make_view(blas, B, k, n, ldb, rop),
                         blas.Matrix(Cb, m, n, ldc), product_scale=alpha, result_scale=beta)
    torch_cuda.cuda.synchronize()
    got = host(Cb)

    def mat(x, rows, cols, ld):
        return x[: ld * cols].reshape(cols, ld)[:, :rows].T

    u = 2.0**-24 if dtype in (np.float32, np.complex64) else 2.0**-53
    mag = abs(alpha) * (np.abs(cop(mat(a, k, m, lda), ta)).astype(np.float64) @ np.abs(mat(b, k, n, ldb)))
    mag = mag + abs(beta) * np.abs(mat(c, m, n, ldc))
    cfac = 2.0 if real else 4.0
    err = np.abs(mat(got, m, n, ldc).astype(np.complex128) - mat(want, m, n, ldc))
    assert np.all(err <= cfac * (k + 2) * u * mag)
    # padding outside the logical result untouched
    pad = got.reshape(n, ldc)[:, m:]
    assert np.all(pad == -999)


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_workspace_form_with_aliased_result(torch_cuda, oracle, dtype):
    """matrixMultiplyWithWorkspace: the result may alias an input (operations.zig:524-559):
    C := s * A * C + r * C with C both the right operand and the result."""
    from zynum_amd import blas

    real = np.dtype(dtype).kind == "f"
    rng = np.random.default_rng(5)
    m = 96
    a = view_storage(rng, m, m, m, dtype, True)
    c = view_storage(rng, m, m, m, dtype, True)
    alpha, beta = (3, -2) if real else (complex(1, 2), complex(-1, 1))
    want = oracle_product(oracle, dtype, "N", "N", m, m, m, alpha, a, m, c.copy(), m, beta, c, m)
    A, Cb = storage(torch_cuda, a, "device"), storage(torch_cuda, c, "device")
    with pytest.raises(blas.AliasingNotAllowed):
        blas.matrix_multiply(blas.ConstMatrix(A, m, m, m), blas.ConstMatrix(Cb, m, m, m), blas.Matrix(Cb, m, m, m))
    work = torch_cuda.full((m * m,), float("nan"), dtype=Cb.dtype, device="cuda")
    blas.matrix_multiply_with_workspace(blas.ConstMatrix(A, m, m, m), blas.ConstMatrix(Cb, m, m, m),
                                        blas.Matrix(Cb, m, m, m), work, product_scale=alpha, result_scale=beta)
    torch_cuda.cuda.synchronize()
    assert np.array_equal(host(Cb), want)
