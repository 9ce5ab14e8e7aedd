"""Grouped GEMM of uneven problems: one persistent launch (plain or replayed
from a HIP graph) against single plain GEMM calls and the oracle.

Bars: every problem's result is bitwise the plain single-GEMM result (same
MFMA tile routine and per-element order), whatever tile configuration the
plan uses; the oracle agrees within the normwise bound; graph replays with
eager work in between reproduce the eager result every time.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def batch(torch, count, seed, dtype, transa=False, transb=False, alpha=1.0, beta=0.0, pad=0):
    from zynum_amd import grouped

    shapes = grouped.draw_shapes(count, seed, 16, 512)
    return grouped.Batch(torch, shapes, dtype, seed=seed, transa=transa, transb=transb, alpha=alpha, beta=beta,
                         pad=pad)


@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_grouped_bitwise_equals_single_calls(torch_cuda, dtype_name, ta, tb):
    torch = torch_cuda
    dt = getattr(torch, dtype_name)
    b = batch(torch, 96, 11, dt, ta, tb, alpha=0.75, beta=-0.25, pad=3)
    b.reset_c()
    b.loop()
    torch.cuda.synchronize()
    want = b.C.clone()
    for tile in (0, 1, 2, 3):
        b.reset_c()
        p = b.plan(tile)
        p.run()
        torch.cuda.synchronize()
        assert torch.equal(b.C, want), f"tile {tile}"
        p.close()


def test_grouped_matches_oracle(torch_cuda, oracle):
    torch = torch_cuda
    b = batch(torch, 40, 5, torch.float64, alpha=1.5, beta=0.5, pad=1)
    b.reset_c()
    p = b.plan()
    p.run()
    torch.cuda.synchronize()
    A, B, C0 = b.A.cpu().numpy(), b.B.cpu().numpy(), b.C0.cpu().numpy()
    for i, (m, n, k) in enumerate(b.shapes):
        lda, ldb, ldc = b.lds[i]
        oa, ob, oc = b.offs[0][i], b.offs[1][i], b.offs[2][i]
        c = C0[oc: oc + ldc * n].copy()
        oracle.dgemm("N", "N", int(m), int(n), int(k), 1.5, A[oa:], lda, B[ob:], ldb, 0.5, c, ldc)
        want = c.reshape(n, ldc)[:, :m].T
        Am = A[oa: oa + lda * k].reshape(k, lda)[:, :m].T
        Bm = B[ob: ob + ldb * n].reshape(n, ldb)[:, :k].T
        bound = 2 * (k + 2) * 2.0**-53 * (1.5 * np.abs(Am) @ np.abs(Bm) + 0.5 * np.abs(C0[oc: oc + ldc * n].reshape(n, ldc)[:, :m].T))
        assert np.all(np.abs(b.result(i) - want) <= bound)


def test_graph_replays_match_eager(torch_cuda):
    """Replays of the captured (counter reset + persistent kernel) graph, with
    eager work and synchronizations between them, each equal the eager run."""
    torch = torch_cuda
    b = batch(torch, 64, 3, torch.float64, beta=0.0)
    p = b.plan()
    p.run()
    torch.cuda.synchronize()
    want = b.C.clone()
    s = torch.cuda.Stream()
    for rep in range(4):
        b.C.fill_(float("nan"))
        _ = (b.A[:1000] * 2).sum()  # eager work between replays
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            p.graph_run(s.cuda_stream)
        s.synchronize()
        assert torch.equal(b.C, want), f"replay {rep}"


def test_scale_only_problems_and_invalid_batch(torch_cuda):
    """k == 0 and alpha == 0 problems scale C (exact 0 when beta == 0); an
    invalid problem declines the whole plan before anything runs."""
    import ctypes as C

    from zynum_amd import _lib, grouped

    torch = torch_cuda
    b = grouped.Batch(torch, [(37, 29, 0), (40, 41, 16)], torch.float64, seed=1, alpha=0.0, beta=2.0)
    b.reset_c()
    p = b.plan()
    p.run()
    torch.cuda.synchronize()
    assert torch.equal(b.C, 2.0 * b.C0)
    bad = grouped.Batch(torch, [(8, 8, 8)], torch.float64, seed=1)
    bad.problems[0].lda = 4
    h = C.c_void_p()
    st = _lib.load().zynum_gpu_grouped_plan_create(_lib.F64, bad.problems, 1, 0, C.byref(h))
    assert st == _lib.DECLINE_ARGUMENT and not h.value


def test_benchmark_batch_runner(torch_cuda, capsys):
    """`python -m zynum_amd.grouped` (the command the profile's counters come
    from) plans the benchmark batch and runs it without graph capture."""
    from zynum_amd import grouped

    assert grouped.main(["--dtype", "f32", "--reps", "1"]) == 0
    out = capsys.readouterr().out
    assert "grouped f32:" in out and "'tiles': " in out and "last run" in out
