"""The autotuner's table writers: the fastest configuration per class is the
one written, rows keep the layout the C++ tables are compiled from, and the
committed tables parse with that layout."""
import re
from pathlib import Path

from zynum_amd import autotune

CSRC = Path(__file__).resolve().parents[1] / "zynum_amd" / "csrc"


def _plain_rows(text):
    return [(int(e), int(m, 16), int(t), int(c))
            for e, m, t, c in re.findall(r"\{(\d+), 0x([0-9a-f]+), (\d+), (\d+)\},", text)]


def _pick(rows, m, n, eb, ta, tb):
    """The selection rule of pick_tile (tune.cpp)."""
    tiles = ((m + 127) // 128) * ((n + 127) // 128)
    pair = 1 << ((2 if ta else 0) + (1 if tb else 0))
    return next(c for e, mask, t, c in rows if e == eb and mask & pair and tiles >= t)


def test_plain_table_picks_fastest_per_class():
    scores = {(eb, n, p, c): (10.0 if c == (3 if n == 8192 else 5) else 1.0) for eb in (8, 4)
              for n in autotune.PLAIN_SIZES for p in autotune.PAIRS for c in range(autotune.PLAIN_TILES)}
    rows = _plain_rows(autotune.plain_table(scores))
    assert (8, 0xF, 512, 3) in rows and (8, 0xF, 128, 5) in rows and (4, 0xF, 0, 5) in rows
    assert len(rows) == 2 * len(autotune.PLAIN_SIZES)


def test_plain_table_splits_pairs_with_different_winners():
    """A pair whose fastest configuration differs gets a row of its own, and
    every pair of every class is covered exactly once."""
    scores = {(eb, n, p, c): (10.0 if c == (10 if p == "TN" else 11) else 1.0) for eb in (8, 4)
              for n in autotune.PLAIN_SIZES for p in autotune.PAIRS for c in range(autotune.PLAIN_TILES)}
    rows = _plain_rows(autotune.plain_table(scores))
    assert (8, 0xB, 512, 11) in rows and (8, 0x4, 512, 10) in rows
    for eb in (8, 4):
        for n in autotune.PLAIN_SIZES:
            masks = [mask for e, mask, t, _ in rows if e == eb and t == autotune.PLAIN_CLASS_MIN_TILES[n]]
            assert sum(masks) == 0xF and all(a & b == 0 for i, a in enumerate(masks) for b in masks[i + 1:])
            for ta in (False, True):
                for tb in (False, True):
                    want = 10 if (ta and not tb) else 11
                    assert _pick(rows, n, n, eb, ta, tb) == want


def test_each_measured_size_selects_its_own_row():
    """The first row whose threshold a size's 128x128 tile count reaches must
    be the row measured at that size (a 1024^3 GEMM must not pick the 2048 row)."""
    sizes = sorted(autotune.PLAIN_SIZES, reverse=True)
    for n in sizes:
        tiles = ((n + 127) // 128) ** 2
        first = next(s for s in sizes if tiles >= autotune.PLAIN_CLASS_MIN_TILES[s])
        assert first == n, (n, first)
    i8_sizes = sorted(autotune.I8_SIZES, reverse=True)
    for n in i8_sizes:
        tiles = ((n + 255) // 256) ** 2
        assert next(s for s in i8_sizes if tiles >= autotune.I8_CLASS_MIN_TILES[s]) == n


def test_grouped_and_int8_tables():
    g = {(eb, hi, c): float(c == (0 if hi <= 48 else 1)) for eb in (8, 4)
         for _, hi in autotune.GROUPED_BUCKETS for c in range(autotune.GROUPED_TILES)}
    text = autotune.grouped_table(g)
    assert "{8, 48, 0}" in text and "{8, 1 << 30, 1}" in text
    i8 = {(n, c): float(c == 10) for n in autotune.I8_SIZES for c in range(autotune.I8_TILES)}
    assert "{128, 10}" in autotune.i8_table(i8)


def test_committed_tables_have_the_compiled_layout():
    plain = (CSRC / "tune_table.inc").read_text()
    assert "kPlainTuneTable" in plain
    rows = _plain_rows(plain)
    assert rows and all(0 <= c < autotune.PLAIN_TILES for *_, c in rows)
    for eb in (8, 4):  # every pair of every element size resolves, down to one 128x128 tile
        for ta in (False, True):
            for tb in (False, True):
                _pick(rows, 1, 1, eb, ta, tb)
    grouped = (CSRC / "grouped_tune.inc").read_text()
    assert "kGroupedTuneTable" in grouped and "1 << 30" in grouped
    i8 = (CSRC / "i8_tune.inc").read_text()
    rows = re.findall(r"\{(\d+), (\d+)\},", i8)
    assert rows and all(0 <= int(c) < autotune.I8_TILES for _, c in rows) and rows[-1][0] == "0"


def test_offload_fit_recovers_the_staging_model():
    """The staging fit returns the call cost and per-byte rates that generated
    exact rows, including shapes dominated by one direction."""
    call, a, b = 12000.0, 0.05, 0.08
    rows = [(i, o, call + a * i + b * o) for i, o in
            ((98304, 32768), (1 << 20, 1 << 19), (24 << 20, 8 << 20), (4 << 27, 32768), (1 << 19, 1 << 27),
             (3 << 27, 1 << 27))]
    got = autotune.fit_stage(rows)
    assert abs(got[0] - call) < 1e-3 * call and abs(got[1] - a) < 1e-3 * a and abs(got[2] - b) < 1e-3 * b


def test_offload_table_layout_and_monotone_device_times():
    dev = {(s, e): 20000.0 + (e ** 3) * (0.01 if s < 2 else 0.04) for s in range(4) for e in autotune.OFFLOAD_EDGES}
    dev[(1, 64)] = 1.0  # a noisy dip must not make the table non-monotone
    rows = [(1 << 20, 1 << 19, 9000.0 + 0.04 * (1 << 20) + 0.05 * (1 << 19)),
            (1 << 24, 1 << 10, 9000.0 + 0.04 * (1 << 24) + 0.05 * (1 << 10)),
            (1 << 10, 1 << 24, 9000.0 + 0.04 * (1 << 10) + 0.05 * (1 << 24))]
    text = autotune.offload_table(dev, rows)
    pts = [(int(s), float(w), float(t)) for s, w, t in re.findall(r"\{(\d), (\d+), (\d+)\}", text)]
    assert len(pts) == 4 * len(autotune.OFFLOAD_EDGES)
    for s in range(4):
        mine = [p for p in pts if p[0] == s]
        assert all(x[1] < y[1] and x[2] <= y[2] for x, y in zip(mine, mine[1:]))
    stage = re.search(r"kOffloadStage = \{([-\d.e]+), ([-\d.e]+), ([-\d.e]+)\}", text)
    assert stage and abs(float(stage.group(1)) - 9000.0) < 1.0


def test_committed_offload_table_parses():
    text = (CSRC / "offload_tune.inc").read_text()
    pts = [(int(s), float(w), float(t)) for s, w, t in re.findall(r"\{(\d), (\d+), (\d+)\}", text)]
    assert {p[0] for p in pts} == {0, 1, 2, 3}
    for s in range(4):
        mine = [p for p in pts if p[0] == s]
        assert all(x[1] < y[1] and x[2] <= y[2] for x, y in zip(mine, mine[1:]))
    assert re.search(r"kOffloadStage = \{[-\d.e]+, [-\d.e]+, [-\d.e]+\};", text)
