"""zynum_gpu_gemm_noalloc against zynum_gpu_gemm through the C ABI.

Bar: bitwise equal results for the same request (same kernels, same split-K
plan; the partials live in the fixed device array instead of allocated
memory), on two calls in a row, the second with a smaller input, for shapes
that split K (the array is used and keeps the first call's contents) and
shapes that do not; the padding rows of C untouched; host operands, other
modes and complex scalars decline before writing.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def request(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, scalar, stream=None, mode=0):
    from zynum_amd import _lib, blas

    r = _lib.GemmRequest()
    r.transa = blas._trans_code(ta)
    r.transb = blas._trans_code(tb)
    r.scalar = scalar
    r.m, r.n, r.k, r.lda, r.ldb, r.ldc = m, n, k, lda, ldb, ldc
    r.alpha_re, r.beta_re = blas._bits(alpha, scalar), blas._bits(beta, scalar)
    r.a, r.b, r.c = a, b, c
    r.stream = stream
    r.mode = mode
    return r


def run_both(torch, dtype, ta, tb, m, n, k, seed, stream=None):
    from zynum_amd import _lib

    lib = _lib.load()
    rng = np.random.default_rng(seed)
    lda = (m if ta == "N" else k) + 3
    ldb = (k if tb == "N" else n) + 1
    ldc = m + 2
    a = torch.from_numpy(rng.uniform(-1, 1, lda * (k if ta == "N" else m)).astype(dtype)).cuda()
    b = torch.from_numpy(rng.uniform(-1, 1, ldb * (n if tb == "N" else k)).astype(dtype)).cuda()
    c0 = np.full(ldc * n, -777.0, dtype=dtype)
    c0.reshape(n, ldc)[:, :m] = rng.uniform(-1, 1, (n, m))
    c_ref = torch.from_numpy(c0.copy()).cuda()
    c_na = torch.from_numpy(c0.copy()).cuda()
    scalar = _lib.F64 if dtype == np.float64 else _lib.F32
    r = request(ta, tb, m, n, k, 1.25, a.data_ptr(), lda, b.data_ptr(), ldb, -0.5, c_ref.data_ptr(), ldc, scalar,
                stream)
    _lib.check(lib.zynum_gpu_gemm(C.byref(r)), "gemm")
    r.c = c_na.data_ptr()
    _lib.check(lib.zynum_gpu_gemm_noalloc(C.byref(r)), "gemm_noalloc")
    torch.cuda.synchronize()
    assert torch.equal(c_na, c_ref), f"{dtype.__name__} {ta}{tb} {m}x{n}x{k}"
    assert bool((c_na.view(n, ldc)[:, m:] == -777.0).all())
    return c_na


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_noalloc_matches_gemm_split_k_then_smaller(torch_cuda, dtype, ta, tb):
    # first call: 127 output tiles of 64 x 64, K 4096 -> 5 K ranges (the largest partial buffer a plan uses)
    run_both(torch_cuda, dtype, ta, tb, 127 * 64, 64, 4096, 1)
    # second call, smaller: 4 tiles, K 2048 -> split again, over fewer partials
    run_both(torch_cuda, dtype, ta, tb, 128, 100, 2048, 2)
    # and a shape that does not split
    run_both(torch_cuda, dtype, ta, tb, 300, 200, 100, 3)


def test_noalloc_on_a_caller_stream(torch_cuda):
    s = torch_cuda.cuda.Stream()
    run_both(torch_cuda, np.float64, "N", "N", 256, 256, 8192, 4, stream=s.cuda_stream)
    run_both(torch_cuda, np.float64, "N", "T", 64, 64, 1024, 5, stream=s.cuda_stream)


def test_noalloc_declines_before_writing(torch_cuda):
    from zynum_amd import _lib

    lib = _lib.load()
    m = n = k = 64
    a = torch_cuda.ones(m * k, dtype=torch_cuda.float64, device="cuda")
    c = torch_cuda.full((m * n,), 7.0, dtype=torch_cuda.float64, device="cuda")
    host_b = np.ones(k * n)
    r = request("N", "N", m, n, k, 1.0, a.data_ptr(), m, host_b.ctypes.data, k, 0.0, c.data_ptr(), m, _lib.F64)
    assert lib.zynum_gpu_gemm_noalloc(C.byref(r)) == _lib.DECLINE_UNSUPPORTED  # host operand
    r = request("N", "N", m, n, k, 1.0, a.data_ptr(), m, a.data_ptr(), k, 0.0, c.data_ptr(), m, _lib.F64, mode=3)
    assert lib.zynum_gpu_gemm_noalloc(C.byref(r)) == _lib.DECLINE_UNSUPPORTED  # split precision needs a workspace
    r = request("N", "N", m, n, k, 1.0, a.data_ptr(), m, a.data_ptr(), k, 0.0, c.data_ptr(), m, _lib.C64)
    assert lib.zynum_gpu_gemm_noalloc(C.byref(r)) == _lib.DECLINE_UNSUPPORTED  # complex expansion needs a workspace
    r = request("N", "N", m, n, k, 1.0, a.data_ptr(), m - 1, a.data_ptr(), k, 0.0, c.data_ptr(), m, _lib.F64)
    assert lib.zynum_gpu_gemm_noalloc(C.byref(r)) == _lib.DECLINE_ARGUMENT  # lda < m
    torch_cuda.cuda.synchronize()
    assert bool((c == 7.0).all())
