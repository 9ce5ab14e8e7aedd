"""Autotuner of the GEMM core: times every tile configuration on the GPU it
runs on and writes the selection tables that are compiled into the library.

This replaces the CPU planner's blocking heuristics (planner.zig:88-211,
tuning.zig:68-155): instead of rules on thread counts and panel widths, the
tile shape per size class comes from measurements on the target MI355X.

    python -m zynum_amd.autotune            # measure and print the tables
    python -m zynum_amd.autotune --write    # also rewrite csrc/{tune_table,grouped_tune,i8_tune,offload_tune}.inc
                                            # (rebuild the library afterwards)
    python -m zynum_amd.autotune --out DIR  # write the tables into DIR instead

Tables:
  tune_table.inc    plain FP64/FP32 MFMA GEMM: first row whose element size
                    and transpose pair match and whose minimum count of
                    128x128 output tiles the shape reaches gives the
                    configuration.
  grouped_tune.inc  grouped GEMM: configuration per bucket of max(m, n).
  i8_tune.inc       INT8 GEMM of the split-precision forms: first row whose
                    minimum count of 256x256 output tiles the shape reaches.
  offload_tune.inc  host-offload model: device-operand call times of square
                    problems per scalar type and the fitted cost of staging
                    host operands (zynum_gpu_offload_predict_ns and friends).
"""
from __future__ import annotations

import argparse
import math
import sys
import time
from pathlib import Path

import numpy as np

from . import _lib, blas, grouped, synth

CSRC = Path(__file__).resolve().parent / "csrc"
PLAIN_TILES = 12         # kTileConfigs
GROUPED_TILES = 3        # kGroupedCfgs
I8_TILES = 26            # kI8Tiles
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


PAIRS = ("NN", "NT", "TN", "TT")  # bit i of a table row's pair mask


def plain_sweep(torch, stream, log=print) -> dict:
    """TFLOP/s of every plain tile configuration, FP64 and FP32, each transpose
    pair, beta 0."""
    out = {}
    sh = stream.cuda_stream
    for scalar, dt, eb in ((_lib.F64, torch.float64, 8), (_lib.F32, torch.float32, 4)):
        for n in PLAIN_SIZES:
            ld = n + 8
            A = torch.rand(ld * n, dtype=dt, device="cuda") * 2 - 1
            B = torch.rand(ld * n, dtype=dt, device="cuda") * 2 - 1
            C = torch.zeros(ld * n, dtype=dt, device="cuda")
            reps = 5 if n >= 8192 else 20
            for pair in PAIRS:
                ta, tb = pair[0] == "T", pair[1] == "T"
                for tile in range(1, PLAIN_TILES + 1):
                    def go(tile=tile):
                        _lib.check(blas.gemm_request(ta, tb, n, n, n, 1.0, A.data_ptr(), ld, B.data_ptr(), ld, 0.0,
                                                     C.data_ptr(), ld, scalar, stream=sh, flags=_lib.FLAG_ASYNC,
                                                     tile=tile), "plain sweep")
                    ms = _time(torch, stream, go, reps)
                    tf = 2.0 * n ** 3 / (ms * 1e-3) / 1e12
                    out[(eb, n, pair, tile - 1)] = tf
                    log(f"plain elem={eb} n={n} {pair} cfg={tile - 1}: {ms:.3f} ms {tf:.1f} TFLOP/s")
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


OFFLOAD_EDGES = (8, 16, 32, 64, 96, 128, 192, 256, 384, 512, 768, 1024, 1536, 2048, 3072, 4096)
OFFLOAD_SYMBOLS = {_lib.F32: ("sgemm_", np.float32), _lib.F64: ("dgemm_", np.float64),
                   _lib.C32: ("cgemm_", np.complex64), _lib.C64: ("zgemm_", np.complex128)}
# (m, n, k, beta): shapes whose input and output bytes differ, so the staging
# fit can tell the two directions apart
OFFLOAD_STAGE_SHAPES = ((64, 64, 64, 0), (256, 256, 256, 0), (1024, 1024, 1024, 0), (2048, 2048, 2048, 0),
                        (64, 64, 65536, 0), (64, 64, 262144, 0), (2048, 2048, 16, 0), (4096, 4096, 16, 0),
                        (1024, 1024, 1024, 1), (4096, 4096, 16, 1))


def blas_call_ns(torch, lib, scalar: int, m: int, n: int, k: int, host: bool, beta: float, reps: int) -> float:
    """Best wall time of one synchronous Fortran-symbol call (column-major, NN)."""
    import ctypes as C

    sym, dt = OFFLOAD_SYMBOLS[scalar]
    rng = np.random.default_rng(m * 131 + n * 7 + k)
    a = rng.uniform(-1, 1, m * k).astype(dt)
    b = rng.uniform(-1, 1, k * n).astype(dt)
    c = rng.uniform(-1, 1, m * n).astype(dt)
    keep = (a, b, c)
    if not host:
        keep = tuple(torch.from_numpy(x).cuda() for x in (a, b, c))
        torch.cuda.synchronize()  # the Fortran symbols run on the library stream
    pa, pb, pc = (x.ctypes.data if host else x.data_ptr() for x in keep)
    one, bet = np.ones(1, dtype=dt), np.full(1, beta, dtype=dt)
    if scalar in (_lib.F32, _lib.F64):  # real scalars by typed pointer, complex ones as addresses
        ptr = C.POINTER(C.c_float if scalar == _lib.F32 else C.c_double)
        alpha, beta_p = one.ctypes.data_as(ptr), bet.ctypes.data_as(ptr)
    else:
        alpha, beta_p = one.ctypes.data, bet.ctypes.data
    I = C.c_int32
    args = (b"N", b"N", C.byref(I(m)), C.byref(I(n)), C.byref(I(k)), alpha, pa, C.byref(I(m)), pb,
            C.byref(I(k)), beta_p, pc, C.byref(I(m)))
    fn = getattr(lib, sym)
    fn(*args)
    _lib.check(lib.zynum_gpu_last_status(), f"offload sweep {sym}")
    best = math.inf
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter_ns()
        fn(*args)
        best = min(best, time.perf_counter_ns() - t0)
    _lib.check(lib.zynum_gpu_last_status(), f"offload sweep {sym}")
    return float(best)


def offload_sweep(torch, stream, log=print):
    """Device-operand call times per scalar and edge, and host-minus-device
    times for the staging fit."""
    lib = _lib.load()
    dev = {}
    for scalar in OFFLOAD_SYMBOLS:
        for e in OFFLOAD_EDGES:
            ns = blas_call_ns(torch, lib, scalar, e, e, e, False, 0.0, 9 if e <= 1024 else 3)
            dev[(scalar, e)] = ns
            log(f"offload device scalar={scalar} edge={e}: {ns / 1e3:.1f} us")
    stage = []
    for scalar in (_lib.F64, _lib.C64):
        es = np.dtype(OFFLOAD_SYMBOLS[scalar][1]).itemsize
        for m, n, k, beta in OFFLOAD_STAGE_SHAPES:
            if scalar == _lib.C64 and m * n * k > 2048 ** 3:
                continue
            reps = 5
            t_dev = blas_call_ns(torch, lib, scalar, m, n, k, False, float(beta), reps)
            t_host = blas_call_ns(torch, lib, scalar, m, n, k, True, float(beta), reps)
            in_b = (m * k + k * n + (m * n if beta else 0)) * es
            out_b = m * n * es
            stage.append((in_b, out_b, t_host - t_dev))
            log(f"offload staging scalar={scalar} m={m} n={n} k={k} beta={beta}: device {t_dev / 1e3:.1f} us, "
                f"host {t_host / 1e3:.1f} us, in {in_b / 2**20:.1f} MiB out {out_b / 2**20:.1f} MiB")
    return dev, stage


def fit_stage(rows) -> tuple:
    """Non-negative least squares of extra = call + in_bytes * a + out_bytes * b
    over (in_bytes, out_bytes, extra_ns) rows, each weighted by 1/extra so
    that small and large shapes count alike."""
    from scipy.optimize import nnls

    x = np.array([[1.0, i, o] for i, o, _ in rows])
    y = np.array([max(t, 1.0) for *_, t in rows])
    w = 1.0 / y
    coef, _ = nnls(x * w[:, None], y * w)
    return float(coef[0]), float(coef[1]), float(coef[2])


def offload_table(dev: dict, stage_rows) -> str:
    call, a, b = fit_stage(stage_rows)
    rows = []
    for scalar in sorted({s for s, _ in dev}):
        pts, last = [], 0.0
        for e in sorted(e for s, e in dev if s == scalar):
            ns = max(dev[(scalar, e)], last)  # keep the table monotone in work
            last = ns
            pts.append(f"{{{scalar}, {float(e) ** 3:.0f}, {ns:.0f}}}")
        rows.append("    " + ", ".join(pts) + ",")
    notes = [f"// staging rows (in bytes, out bytes, host - device ns): " +
             "; ".join(f"{i} {o} {t:.0f}" for i, o, t in stage_rows)]
    return ("// Host-offload model of the GPU tier.  Written by zynum_amd/autotune.py from\n"
            "// timings on MI355X; edit by re-running it (--only offload).\n"
            "// kOffloadDevice: scalar (0 f32, 1 f64, 2 c64, 3 c128), m*n*k of a square\n"
            "// problem, best wall ns of one synchronous call on device operands; rows of\n"
            "// one scalar in increasing work.\n"
            "// kOffloadStage: fixed ns of a call with host operands, ns per byte staged in,\n"
            "// ns per byte staged out (non-negative least squares over the staging rows).\n" +
            "\n".join(notes) + "\n"
            "static const OffloadPoint kOffloadDevice[] = {\n" + "\n".join(rows) + "\n};\n"
            f"static const OffloadStage kOffloadStage = {{{call:.1f}, {a:.6g}, {b:.6g}}};\n")


def _best(scores: dict, key_prefix: tuple, count: int) -> int:
    vals = [scores.get(key_prefix + (c,), -math.inf) for c in range(count)]
    return int(np.argmax(vals))


def plain_table(scores: dict) -> str:
    rows, notes = [], []
    for eb in (8, 4):
        for n in PLAIN_SIZES:
            best = {}
            for i, pair in enumerate(PAIRS):
                best.setdefault(_best(scores, (eb, n, pair), PLAIN_TILES), []).append(i)
                notes.append(f"// measured: {n}^3 {pair} elem {eb}: " +
                             "  ".join(f"cfg{c} {scores[(eb, n, pair, c)]:.1f}" for c in range(PLAIN_TILES)) +
                             " TFLOP/s")
            for cfg, bits in sorted(best.items(), key=lambda kv: min(kv[1])):
                mask = sum(1 << b for b in bits)
                rows.append(f"    {{{eb}, 0x{mask:x}, {PLAIN_CLASS_MIN_TILES[n]}, {cfg}}},")
    return ("// Tile-selection table for the plain MFMA GEMM.  Written by\n"
            "// zynum_amd/autotune.py from timings on MI355X; edit by re-running it.\n"
            "// Columns: element bytes, transpose pairs covered (bit 0 NN, 1 NT, 2 TN,\n"
            "// 3 TT), minimum number of 128x128 output tiles, tile config.  The first\n"
            "// row whose element size and pair match and whose threshold the shape\n"
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
    ap.add_argument("--only", choices=("plain", "grouped", "int8", "offload"), action="append",
                    help="restrict to some tables (repeatable)")
    args = ap.parse_args(argv)
    import torch

    if not torch.cuda.is_available():
        print("autotune needs a GPU", file=sys.stderr)
        return 2
    _lib.load()
    stream = torch.cuda.Stream()
    which = set(args.only or ("plain", "grouped", "int8", "offload"))
    tables = {}
    with torch.cuda.stream(stream):
        if "plain" in which:
            tables["tune_table.inc"] = plain_table(plain_sweep(torch, stream))
        if "grouped" in which:
            tables["grouped_tune.inc"] = grouped_table(grouped_sweep(torch, stream))
        if "int8" in which:
            tables["i8_tune.inc"] = i8_table(i8_sweep(torch, stream))
        if "offload" in which:
            tables["offload_tune.inc"] = offload_table(*offload_sweep(torch, stream))
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
