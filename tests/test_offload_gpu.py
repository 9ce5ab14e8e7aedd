"""Host-pointer offload on the GPU: the staged call is the same computation as
the device-operand call, and the offload model predicts its wall time."""
import pytest

from zynum_amd import _lib, autotune

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch(torch_cuda):
    return torch_cuda


@pytest.fixture(scope="module")
def lib(torch):
    return _lib.load()


@pytest.mark.parametrize("scalar,edge", [(_lib.F64, 256), (_lib.F64, 1024), (_lib.F32, 512), (_lib.C64, 512)])
def test_model_predicts_host_and_device_call_times(torch, lib, scalar, edge):
    """Predictions within a factor of 2.5 of the best of 7 measured calls (the
    table comes from the same kind of machine; this bounds drift, not noise)."""
    for host, mask in ((False, 0), (True, 7)):
        got = autotune.blas_call_ns(torch, lib, scalar, edge, edge, edge, host, 0.0, 7)
        want = lib.zynum_gpu_offload_predict_ns(scalar, edge, edge, edge, mask, 0)
        print(f"scalar={scalar} edge={edge} host={host}: measured {got / 1e3:.1f} us, model {want / 1e3:.1f} us",
              flush=True)
        assert want / 2.5 <= got <= want * 2.5


def test_skinny_host_shapes_follow_the_staging_fit(torch, lib):
    """Shapes dominated by input bytes (large k) or by output bytes (k = 16)."""
    for m, n, k in ((64, 64, 65536), (2048, 2048, 16)):
        got = autotune.blas_call_ns(torch, lib, _lib.F64, m, n, k, True, 0.0, 5)
        want = lib.zynum_gpu_offload_predict_ns(_lib.F64, m, n, k, 7, 0)
        print(f"dgemm {m}x{n}x{k} host: measured {got / 1e3:.1f} us, model {want / 1e3:.1f} us", flush=True)
        assert want / 2.5 <= got <= want * 2.5


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
