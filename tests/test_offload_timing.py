"""Wall-clock checks of the host-pointer offload model (perf marker, not part
of the gpu parity suite: timings depend on load and clocks).

Run on a GPU with ``pytest -m perf tests/test_offload_timing.py``."""
import pytest

from zynum_amd import _lib, autotune

pytestmark = pytest.mark.perf


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
