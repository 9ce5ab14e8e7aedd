"""The autotuner's table writers: the fastest configuration per class is the
one written, rows keep the layout the C++ tables are compiled from, and the
committed tables parse with that layout."""
import re
from pathlib import Path

from zynum_amd import autotune

CSRC = Path(__file__).resolve().parents[1] / "zynum_amd" / "csrc"


def test_plain_table_picks_fastest_per_class():
    scores = {(eb, n, c): (10.0 if c == (3 if n == 8192 else 5) else 1.0) for eb in (8, 4)
              for n in autotune.PLAIN_SIZES for c in range(autotune.PLAIN_TILES)}
    rows = re.findall(r"\{(\d+), (\d+), (\d+)\},", autotune.plain_table(scores))
    assert ("8", "512", "3") in rows and ("8", "128", "5") in rows and ("4", "0", "5") in rows
    assert len(rows) == 2 * len(autotune.PLAIN_SIZES)


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
    assert "kPlainTuneTable" in plain and re.search(r"\{8, \d+, \d+\}", plain)
    grouped = (CSRC / "grouped_tune.inc").read_text()
    assert "kGroupedTuneTable" in grouped and "1 << 30" in grouped
    i8 = (CSRC / "i8_tune.inc").read_text()
    rows = re.findall(r"\{(\d+), (\d+)\},", i8)
    assert rows and all(0 <= int(c) < autotune.I8_TILES for _, c in rows) and rows[-1][0] == "0"
