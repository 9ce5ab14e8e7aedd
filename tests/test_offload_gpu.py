"""Host-pointer offload on the GPU: the staged call is the same computation as
the device-operand call (the model's wall-time predictions are checked
outside this parity suite, in test_offload_timing.py under the perf marker,
and reported by the benchmark's offload leg)."""
import pytest

from zynum_amd import _lib

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch(torch_cuda):
    return torch_cuda


@pytest.fixture(scope="module")
def lib(torch):
    return _lib.load()


@pytest.mark.parametrize("padded", [False, True])
def test_large_host_operands_match_device_operands_bitwise(torch, lib, padded):
    """Operands of several MiB go through the chunked, multi-threaded staging
    (several chunks per matrix); the result equals the device-operand call bit
    for bit and the padding rows of C keep their sentinel."""
    import ctypes as C

    import numpy as np

    m, n, k = 1500, 1300, 700
    lda, ldb, ldc = (m + 31, k + 17, m + 17) if padded else (m, k, m)
    rng = np.random.default_rng(5)
    a = rng.uniform(-1, 1, lda * k)
    b = rng.uniform(-1, 1, ldb * n)
    c0 = np.full(ldc * n, -777.0)
    c0.reshape(n, ldc)[:, :m] = rng.uniform(-1, 1, (n, m))
    alpha, beta = np.array([1.25]), np.array([0.5])
    I = C.c_int32
    dp = C.POINTER(C.c_double)

    def call(pa, pb, pc):
        lib.dgemm_(b"N", b"N", C.byref(I(m)), C.byref(I(n)), C.byref(I(k)), alpha.ctypes.data_as(dp), pa,
                   C.byref(I(lda)), pb, C.byref(I(ldb)), beta.ctypes.data_as(dp), pc, C.byref(I(ldc)))
        assert lib.zynum_gpu_last_status() == 0

    c_host = c0.copy()
    call(a.ctypes.data, b.ctypes.data, c_host.ctypes.data)
    da, db, dc = (torch.from_numpy(x.copy()).cuda() for x in (a, b, c0))
    torch.cuda.synchronize()
    call(da.data_ptr(), db.data_ptr(), dc.data_ptr())
    c_dev = dc.cpu().numpy()
    assert np.array_equal(c_host, c_dev)
    if padded:
        assert np.all(c_host.reshape(n, ldc)[:, m:] == -777.0)
    # and a sanity check of the product itself on a corner
    A = a.reshape(k, lda)[:, :m].T
    B = b.reshape(n, ldb)[:, :k].T
    want = 1.25 * (A[:64] @ B[:, :64]) + 0.5 * c0.reshape(n, ldc)[:64, :64].T
    assert np.allclose(c_host.reshape(n, ldc)[:64, :64].T, want, rtol=0, atol=1e-11)


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_copy_thread_count_does_not_change_results(torch, lib, threads):
    """zynum_gpu_set_copy_threads changes how host operands are copied, never what is computed."""
    import ctypes as C

    import numpy as np

    m, n, k = 900, 700, 1100
    rng = np.random.default_rng(threads)
    a, b = rng.uniform(-1, 1, m * k), rng.uniform(-1, 1, k * n)
    c0 = rng.uniform(-1, 1, m * n)
    I = C.c_int32
    dp = C.POINTER(C.c_double)
    alpha, beta = np.array([0.5]), np.array([2.0])

    def call(pa, pb, pc):
        lib.dgemm_(b"N", b"N", C.byref(I(m)), C.byref(I(n)), C.byref(I(k)), alpha.ctypes.data_as(dp), pa,
                   C.byref(I(m)), pb, C.byref(I(k)), beta.ctypes.data_as(dp), pc, C.byref(I(m)))
        assert lib.zynum_gpu_last_status() == 0

    assert lib.zynum_gpu_set_copy_threads(threads) == _lib.OK
    try:
        c_host = c0.copy()
        call(a.ctypes.data, b.ctypes.data, c_host.ctypes.data)
    finally:
        assert lib.zynum_gpu_set_copy_threads(0) == _lib.OK
    da, db, dc = (torch.from_numpy(x.copy()).cuda() for x in (a, b, c0))
    torch.cuda.synchronize()
    call(da.data_ptr(), db.data_ptr(), dc.data_ptr())
    assert np.array_equal(c_host, dc.cpu().numpy())
