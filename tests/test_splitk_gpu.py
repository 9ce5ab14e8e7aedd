"""Split-K form of the plain GEMM (few output tiles, long K): K ranges are
multiplied into a partial buffer and summed in increasing order before the
alpha/beta epilogue.  Bars: bitwise on integer-valued inputs (every order is
exact), the componentwise bound 2 (K + 2) u (|alpha| |A| |B| + |beta| |C|)
against a float64 product on random inputs, NaN isolation for beta = 0, and
padding rows untouched."""
import numpy as np
import pytest

from conftest import int_gemm
from test_gemm_gpu import gpu_gemm, logical, normwise_bound
from test_oracle import operands

pytestmark = pytest.mark.gpu

# (m, n, k): every one has fewer than 128 64x64 tiles and K >= 1024
SPLIT_SHAPES = [(64, 64, 4096), (17, 19, 2049), (128, 128, 4096), (200, 300, 1100), (1, 4096, 2048),
                (4096, 1, 2048), (63, 65, 1031)]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("shape", SPLIT_SHAPES)
def test_split_shapes_integer_bitwise(torch_cuda, dtype, ta, tb, shape):
    m, n, k = shape
    lda = (m if ta == "N" else k) + 3
    ldb = (k if tb == "N" else n) + 1
    ldc = m + 2
    a, b, c = operands(ta, tb, m, n, k, lda, ldb, ldc, dtype, True, salts=(3, 8, 4))
    want = int_gemm(ta, tb, m, n, k, -0.5, a, lda, b, ldb, 2.0, c, ldc)
    for via in ("request", "fortran"):
        got = gpu_gemm(torch_cuda, ta, tb, m, n, k, -0.5, a, lda, b, ldb, 2.0, c, ldc, via=via)
        assert np.array_equal(got, want.astype(dtype)), via


@pytest.mark.parametrize("dtype,u", [(np.float64, 2.0 ** -53), (np.float32, 2.0 ** -24)])
@pytest.mark.parametrize("shape", [(64, 64, 8192), (96, 160, 3000), (128, 128, 4096)])
def test_split_shapes_random_within_bound(torch_cuda, dtype, u, shape):
    m, n, k = shape
    rng = np.random.default_rng(m + n + k)
    lda, ldb, ldc = m + 5, k + 7, m + 1
    a = rng.uniform(-1, 1, lda * k).astype(dtype)
    b = rng.uniform(-1, 1, ldb * n).astype(dtype)
    c = rng.uniform(-1, 1, ldc * n).astype(dtype)
    alpha, beta = 1.5, 0.75
    got = gpu_gemm(torch_cuda, "N", "N", m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
    A, B, C0 = logical(a, m, k, lda), logical(b, k, n, ldb), logical(c, m, n, ldc)
    exact = alpha * (A.astype(np.float64) @ B.astype(np.float64)) + beta * C0.astype(np.float64)
    err = np.abs(logical(got, m, n, ldc).astype(np.float64) - exact)
    bound = normwise_bound("N", "N", m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, u)
    assert np.all(err <= bound)
    # padding rows keep their values
    assert np.array_equal(got.reshape(n, ldc)[:, m:], c.reshape(n, ldc)[:, m:])


def test_split_beta_zero_isolates_nan_in_c(torch_cuda):
    m, n, k = 64, 64, 4096
    rng = np.random.default_rng(11)
    a = rng.integers(-3, 4, m * k).astype(np.float64)
    b = rng.integers(-3, 4, k * n).astype(np.float64)
    c = np.full(m * n, np.nan)
    got = gpu_gemm(torch_cuda, "N", "N", m, n, k, 1.0, a, m, b, k, 0.0, c, m)
    want = (a.reshape(k, m).T @ b.reshape(n, k).T)
    assert np.array_equal(got.reshape(n, m).T, want)


def test_split_matches_single_tile_walk_on_integers(torch_cuda):
    """The same call with a forced tile (which never splits) and with the tuned
    choice (which splits here) agree on integer inputs."""
    m, n, k = 128, 128, 4096
    a, b, c = operands("N", "T", m, n, k, m, n, m, np.float64, True, salts=(2, 6, 1))
    split = gpu_gemm(torch_cuda, "N", "T", m, n, k, 0.25, a, m, b, n, -1.0, c, m)
    whole = gpu_gemm(torch_cuda, "N", "T", m, n, k, 0.25, a, m, b, n, -1.0, c, m, tile=1)
    assert np.array_equal(split, whole)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("C", "N"), ("N", "T"), ("T", "C")])
@pytest.mark.parametrize("shape", [(64, 64, 2048), (17, 40, 1500)])
def test_complex_split_shapes_integer_bitwise(torch_cuda, dtype, ta, tb, shape):
    """Complex GEMM on the real core splits K the same way (the expanded
    problem is 2m x n x 2k); integer inputs give the exact result."""
    from conftest import fill_padded_complex, int_cgemm
    from test_complex_gpu import _dims, gpu_cgemm

    m, n, k = shape
    ar, ac, br, bc = _dims(ta, tb, m, n, k)
    lda, ldb, ldc = ar + 4, br + 5, m + 6
    a = fill_padded_complex(ar, ac, lda, 7, dtype, integer=True)
    b = fill_padded_complex(br, bc, ldb, 9, dtype, integer=True)
    c = fill_padded_complex(m, n, ldc, 13, dtype, integer=True)
    alpha, beta = 1 - 1j, -1 + 2j
    want = int_cgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
    got = gpu_cgemm(torch_cuda, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
    assert np.array_equal(got, want)
