"""Autotuner of the GEMM core: times every tile configuration on the GPU it
runs on and writes the selection tables that are compiled into the library.

This replaces the CPU planner's blocking heuristics (planner.zig:88-211,
tuning.zig:68-155): instead of rules on thread counts and panel widths, the
tile shape per size class comes from measurements on the target MI355X.

    python -m zynum_amd.autotune            # measure and print the tables
    python -m zynum_amd.autotune --write    # also rewrite csrc/{tune_table,grouped_tune,i8_tune}.inc
                                            # (rebuild the library afterwards)
    python -m zynum_amd.autotune --out DIR  # write the tables into DIR instead

Tables:
  tune_table.inc    plain FP64/FP32 MFMA GEMM: first row whose element size
                    matches and whose minimum count of 128x128 output tiles the
                    shape reaches gives the configuration.
  grouped_tune.inc  grouped GEMM: configuration per bucket of max(m, n).
  i8_tune.inc       INT8 GEMM of the split-precision forms: first row whose
                    minimum count of 256x256 output tiles the shape reaches.
"""
from __future__ import annotations

import argparse
import math
import sys
from pathlib import Path

import numpy as np

from . import _lib, blas, grouped, synth

CSRC = Path(__file__).resolve().parent / "csrc"
PLAIN_TILES = 11         # kTileConfigs
GROUPED_TILES = 3        # kGroupedCfgs
I8_TILES = 14            # kI8Tiles
PLAIN_SIZES = (8192, 2048, 1024)
PLAIN_CLASS_MIN_TILES = {8192: 512, 2048: 128, 1024: 0}  # 128x128 tiles: 4096, 256, 64 at these sizes
GROUPED_BUCKETS = ((16, 48), (48, 96), (96, 192), (192, 512))
I8_SIZES = (8192, 2048)
I8_CLASS_MIN_TILES = {8192: 128, 2048: 0}               # 256x256 tiles: 1024, 64


def _time(torch, stream, fn, reps: int) -> float:
    fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record(stream)
    for _ in range(reps):
        fn()
    e.record(stream)
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


def plain_sweep(torch, stream, log=print) -> dict:
    """TFLOP/s of every plain tile configuration, FP64 and FP32, NN, beta 0."""
    out = {}
    sh = stream.cuda_stream
    for scalar, dt, eb in ((_lib.F64, torch.float64, 8), (_lib.F32, torch.float32, 4)):
        for n in PLAIN_SIZES:
            ld = n + 8
            A = torch.rand(ld * n, dtype=dt, device="cuda") * 2 - 1
            B = torch.rand(ld * n, dtype=dt, device="cuda") * 2 - 1
            C = torch.zeros(ld * n, dtype=dt, device="cuda")
            reps = 5 if n >= 8192 else 20
            for tile in range(1, PLAIN_TILES + 1):
                def go(tile=tile):
                    _lib.check(blas.gemm_request(False, False, n, n, n, 1.0, A.data_ptr(), ld, B.data_ptr(), ld, 0.0,
                                                 C.data_ptr(), ld, scalar, stream=sh, flags=_lib.FLAG_ASYNC,
                                                 tile=tile), "plain sweep")
                ms = _time(torch, stream, go, reps)
                tf = 2.0 * n ** 3 / (ms * 1e-3) / 1e12
                out[(eb, n, tile - 1)] = tf
                log(f"plain elem={eb} n={n} cfg={tile - 1}: {ms:.3f} ms {tf:.1f} TFLOP/s")
            del A, B, C
    return out


def grouped_sweep(torch, stream, log=print) -> dict:
    """TFLOP/s of each grouped configuration per size bucket (1024 problems,
    max(m, n) in the bucket, k log-uniform in [16, 512])."""
    out = {}
    for dt, eb in ((torch.float64, 8), (torch.float32, 4)):
        for lo, hi in GROUPED_BUCKETS:
            base = grouped.draw_shapes(1024, 17 + lo, 16, 512)
            mn = grouped.draw_shapes(1024, 29 + lo, lo, hi)
            shapes = np.stack([mn[:, 0], mn[:, 1], base[:, 2]], axis=1)
            b = grouped.Batch(torch, shapes, dt, seed=lo)
            for tile in range(1, GROUPED_TILES + 1):
                plan = b.plan(tile)
                ms = _time(torch, stream, lambda: plan.graph_run(stream.cuda_stream), 10)
                tf = b.flops / (ms * 1e-3) / 1e12
                out[(eb, hi, tile - 1)] = tf
                log(f"grouped elem={eb} bucket=({lo},{hi}] cfg={tile - 1}: {ms:.3f} ms {tf:.2f} TFLOP/s")
                plan.close()
            del b
    return out


def i8_sweep(torch, stream, log=print) -> dict:
    """FP64-equivalent TFLOP/s of the 14-modulus CRT DGEMM per INT8 configuration."""
    out = {}
    sh = stream.cuda_stream
    for n in I8_SIZES:
        ld = n + 8
        A = torch.from_numpy(synth.wide_range(7, 0, n, n, ld)).cuda()
        B = torch.from_numpy(synth.wide_range(7, n * n, n, n, ld)).cuda()
        C = torch.zeros(ld * n, dtype=torch.float64, device="cuda")
        for tile in range(1, I8_TILES + 1):
            def go(tile=tile):
                _lib.check(blas.gemm_request(False, False, n, n, n, 1.0, A.data_ptr(), ld, B.data_ptr(), ld, 0.0,
                                             C.data_ptr(), ld, _lib.F64, stream=sh, mode=_lib.MODE_OZAKI_CRT,
                                             mode_param=14, flags=_lib.FLAG_ASYNC, tile=tile), "int8 sweep")
            ms = _time(torch, stream, go, 3 if n >= 8192 else 10)
            tf = 2.0 * n ** 3 / (ms * 1e-3) / 1e12
            out[(n, tile - 1)] = tf
            log(f"int8 crt14 n={n} cfg={tile - 1}: {ms:.3f} ms {tf:.1f} TFLOP/s-eq")
        del A, B, C
    return out


def _best(scores: dict, key_prefix: tuple, count: int) -> int:
    vals = [scores.get(key_prefix + (c,), -math.inf) for c in range(count)]
    return int(np.argmax(vals))


def plain_table(scores: dict) -> str:
    rows, notes = [], []
    for eb in (8, 4):
        for n in PLAIN_SIZES:
            cfg = _best(scores, (eb, n), PLAIN_TILES)
            rows.append(f"    {{{eb}, {PLAIN_CLASS_MIN_TILES[n]}, {cfg}}},")
            notes.append(f"// measured: {n}^3 NN elem {eb}: " +
                         "  ".join(f"cfg{c} {scores[(eb, n, c)]:.1f}" for c in range(PLAIN_TILES)) + " TFLOP/s")
    return ("// Tile-selection table for the plain MFMA GEMM.  Written by\n"
            "// zynum_amd/autotune.py from timings on MI355X; edit by re-running it.\n"
            "// Columns: element bytes, minimum number of 128x128 output tiles, tile config.\n"
            "// The first row whose element size matches and whose threshold the shape\n"
            "// reaches is used.\n" + "\n".join(notes) + "\n"
            "static const TuneRow kPlainTuneTable[] = {\n" + "\n".join(rows) + "\n};\n")


def grouped_table(scores: dict) -> str:
    rows, notes = [], []
    for eb in (8, 4):
        line = []
        for i, (lo, hi) in enumerate(GROUPED_BUCKETS):
            cfg = _best(scores, (eb, hi), GROUPED_TILES)
            bound = "1 << 30" if i == len(GROUPED_BUCKETS) - 1 else str(hi)
            line.append(f"{{{eb}, {bound}, {cfg}}}")
            notes.append(f"// measured: elem {eb} max(m,n) in ({lo},{hi}]: " +
                         "  ".join(f"cfg{c} {scores[(eb, hi, c)]:.2f}" for c in range(GROUPED_TILES)) + " TFLOP/s")
        rows.append("    " + ", ".join(line) + ",")
    return ("// Tile configuration per size bucket for the grouped GEMM.  Written by\n"
            "// zynum_amd/autotune.py from timings on MI355X; edit by re-running it.\n"
            "// Columns: element bytes, largest max(m, n) of the bucket, configuration\n"
            "// (0: 32x64 / 64x32, 1: 64x64, 2: 128x128).  First matching row wins.\n" + "\n".join(notes) + "\n"
            "static const GroupedTuneRow kGroupedTuneTable[] = {\n" + "\n".join(rows) + "\n};\n")


def i8_table(scores: dict) -> str:
    rows, notes = [], []
    for n in I8_SIZES:
        cfg = _best(scores, (n,), I8_TILES)
        rows.append(f"    {{{I8_CLASS_MIN_TILES[n]}, {cfg}}},")
        notes.append(f"// measured: CRT14 {n}^3: " +
                     "  ".join(f"cfg{c} {scores[(n, c)]:.1f}" for c in range(I8_TILES)) + " TFLOP/s-eq")
    return ("// INT8 GEMM configuration of the split-precision forms.  Written by\n"
            "// zynum_amd/autotune.py from timings on MI355X; edit by re-running it.\n"
            "// Columns: minimum number of 256x256 output tiles, configuration.\n"
            "// The first row whose threshold the shape reaches is used.\n" + "\n".join(notes) + "\n"
            "static const I8TuneRow kI8TuneTable[] = {\n" + "\n".join(rows) + "\n};\n")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--write", action="store_true", help="rewrite the .inc tables under zynum_amd/csrc")
    ap.add_argument("--out", type=Path, help="also write the tables into this directory")
    ap.add_argument("--only", choices=("plain", "grouped", "int8"), action="append",
                    help="restrict to some tables (repeatable)")
    args = ap.parse_args(argv)
    import torch

    if not torch.cuda.is_available():
        print("autotune needs a GPU", file=sys.stderr)
        return 2
    _lib.load()
    stream = torch.cuda.Stream()
    which = set(args.only or ("plain", "grouped", "int8"))
    tables = {}
    with torch.cuda.stream(stream):
        if "plain" in which:
            tables["tune_table.inc"] = plain_table(plain_sweep(torch, stream))
        if "grouped" in which:
            tables["grouped_tune.inc"] = grouped_table(grouped_sweep(torch, stream))
        if "int8" in which:
            tables["i8_tune.inc"] = i8_table(i8_sweep(torch, stream))
    for name, text in tables.items():
        print(f"==== {name}\n{text}")
        if args.write:
            (CSRC / name).write_text(text)
        if args.out:
            args.out.mkdir(parents=True, exist_ok=True)
            (args.out / name).write_text(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
