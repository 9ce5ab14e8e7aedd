"""Host-offload crossover (zynum_gpu_offload_*): the model is host arithmetic
over the measured table in offload_tune.inc, so these run without a GPU."""
import re
from pathlib import Path

import pytest

from zynum_amd import _lib

CSRC = Path(__file__).resolve().parents[1] / "zynum_amd" / "csrc"


@pytest.fixture(scope="module")
def lib():
    return _lib.load()


def _table():
    text = (CSRC / "offload_tune.inc").read_text()
    pts = [(int(s), int(w), int(t)) for s, w, t in re.findall(r"\{(\d), (\d+), (\d+)\}", text)]
    call, a, b = (float(x) for x in re.search(r"kOffloadStage = \{([-\d.e]+), ([-\d.e]+), ([-\d.e]+)\}",
                                              text).groups())
    return pts, (call, a, b)


def test_invalid_arguments(lib):
    assert lib.zynum_gpu_offload_predict_ns(4, 8, 8, 8, 0, 0) < 0
    assert lib.zynum_gpu_offload_predict_ns(1, -1, 8, 8, 0, 0) < 0
    assert lib.zynum_gpu_offload_predict_ns(1, 8, 8, 8, 8, 0) < 0
    assert lib.zynum_gpu_offload_choose(1, 8, 8, 8, 0, 0, 0.0) == -_lib.DECLINE_ARGUMENT
    assert lib.zynum_gpu_offload_choose(1, 8, 8, 8, 0, 0, float("nan")) == -_lib.DECLINE_ARGUMENT
    assert lib.zynum_gpu_offload_choose(-1, 8, 8, 8, 0, 0, 100.0) == -_lib.DECLINE_ARGUMENT
    assert lib.zynum_gpu_offload_crossover(1, 9, 0, 100.0) == -_lib.DECLINE_ARGUMENT
    assert lib.zynum_gpu_offload_crossover(1, 7, 0, float("inf")) == -_lib.DECLINE_ARGUMENT


def test_quick_returns_stay_on_the_cpu(lib):
    assert lib.zynum_gpu_offload_predict_ns(1, 0, 100, 100, 7, 1) == 0.0
    for m, n, k in ((0, 5, 5), (5, 0, 5), (5, 5, 0)):
        assert lib.zynum_gpu_offload_choose(1, m, n, k, 7, 1, 1e-6) == 0


def test_device_times_at_and_between_measured_points(lib):
    pts, _ = _table()
    for s in range(4):
        mine = [p for p in pts if p[0] == s]
        for _, w, t in mine:
            e = round(w ** (1 / 3))
            assert e ** 3 == w
            assert lib.zynum_gpu_offload_predict_ns(s, e, e, e, 0, 0) == pytest.approx(t, rel=1e-9)
        for (_, w1, t1), (_, w2, t2) in zip(mine, mine[1:]):
            e1, e2 = round(w1 ** (1 / 3)), round(w2 ** (1 / 3))
            got = lib.zynum_gpu_offload_predict_ns(s, e1, e1, e2, 0, 0)  # work strictly between the two
            assert min(t1, t2) * (1 - 1e-12) <= got <= max(t1, t2) * (1 + 1e-12)
        # below the first point: the first point's time; beyond the last: linear in work
        _, w0, t0 = mine[0]
        assert lib.zynum_gpu_offload_predict_ns(s, 1, 1, 1, 0, 0) == pytest.approx(t0)
        _, wl, tl = mine[-1]
        el = round(wl ** (1 / 3))
        assert lib.zynum_gpu_offload_predict_ns(s, el, el, 2 * el, 0, 0) == pytest.approx(2 * tl, rel=1e-9)


def test_staging_cost_per_operand(lib):
    """A host operand adds the call cost once plus its bytes at the fitted
    rate; C is staged in only when beta is nonzero."""
    _, (call, a, b) = _table()
    m, n, k = 300, 200, 100
    base = lib.zynum_gpu_offload_predict_ns(1, m, n, k, 0, 0)
    es = 8
    assert lib.zynum_gpu_offload_predict_ns(1, m, n, k, 1, 0) - base == pytest.approx(call + m * k * es * a)
    assert lib.zynum_gpu_offload_predict_ns(1, m, n, k, 2, 0) - base == pytest.approx(call + k * n * es * a)
    assert lib.zynum_gpu_offload_predict_ns(1, m, n, k, 4, 0) - base == pytest.approx(call + m * n * es * b)
    assert lib.zynum_gpu_offload_predict_ns(1, m, n, k, 4, 1) - base == pytest.approx(
        call + m * n * es * (a + b))
    assert lib.zynum_gpu_offload_predict_ns(1, m, n, k, 7, 1) - base == pytest.approx(
        call + (m * k + k * n + m * n) * es * a + m * n * es * b)
    # complex double stages 16 bytes per element
    base16 = lib.zynum_gpu_offload_predict_ns(3, m, n, k, 0, 0)
    assert lib.zynum_gpu_offload_predict_ns(3, m, n, k, 1, 0) - base16 == pytest.approx(call + m * k * 16 * a)


def test_predictions_grow_with_every_dimension(lib):
    for s in range(4):
        for mask in (0, 7):
            prev = None
            for e in (1, 7, 33, 100, 257, 1000, 3000, 9000):
                t = lib.zynum_gpu_offload_predict_ns(s, e, e, e, mask, 1)
                assert prev is None or t >= prev
                prev = t
            t0 = lib.zynum_gpu_offload_predict_ns(s, 500, 400, 300, mask, 1)
            assert lib.zynum_gpu_offload_predict_ns(s, 1000, 400, 300, mask, 1) >= t0
            assert lib.zynum_gpu_offload_predict_ns(s, 500, 800, 300, mask, 1) >= t0
            assert lib.zynum_gpu_offload_predict_ns(s, 500, 400, 600, mask, 1) >= t0


@pytest.mark.parametrize("scalar", [0, 1, 2, 3])
@pytest.mark.parametrize("mask", [0, 7])
@pytest.mark.parametrize("gflops", [5.0, 100.0, 420.0, 2000.0])
def test_crossover_is_where_the_choice_flips(lib, scalar, mask, gflops):
    e = lib.zynum_gpu_offload_crossover(scalar, mask, 1, gflops)
    assert e >= 1 or e == -1
    if e > 1:
        assert lib.zynum_gpu_offload_choose(scalar, e - 1, e - 1, e - 1, mask, 1, gflops) == 0
    if e >= 1:
        assert lib.zynum_gpu_offload_choose(scalar, e, e, e, mask, 1, gflops) == 1
        assert lib.zynum_gpu_offload_choose(scalar, 2 * e, 2 * e, 2 * e, mask, 1, gflops) == 1
        # the CPU-vs-GPU comparison itself: CPU time at the edge exceeds the prediction
        flops = (8 if scalar >= 2 else 2) * e ** 3
        assert flops / gflops > lib.zynum_gpu_offload_predict_ns(scalar, e, e, e, mask, 1)


def test_staging_and_a_faster_cpu_move_the_crossover_up(lib):
    for s in range(4):
        dev = lib.zynum_gpu_offload_crossover(s, 0, 0, 100.0)
        host = lib.zynum_gpu_offload_crossover(s, 7, 1, 100.0)
        fast = lib.zynum_gpu_offload_crossover(s, 7, 1, 1000.0)
        assert 1 <= dev <= host <= fast
        # a large DGEMM is always worth offloading against the 16-thread CPU path's ~0.4 TFLOP/s
        assert lib.zynum_gpu_offload_choose(s, 8192, 8192, 8192, 7, 1, 420.0) == 1
        # a 1x1x1 product never is
        assert lib.zynum_gpu_offload_choose(s, 1, 1, 1, 7, 1, 1.0) == 0
