"""GEMM shape sweep in Zynum's benchmark CSV format.

Restates the reference's `gemm-sweep` harness (bench/gemm_sweep.zig) for the
GPU tier, so its rows can be set beside the reference's own sweep CSV and
compared by the same report tooling (bench/tools/check_gemm_sweep.py):

* the default shape list (gemm_sweep.zig:58-101) and transpose pair (NN);
* inputs from the harness's LCG (gemm_sweep.zig:156-174: every buffer restarts
  at seed 0x3141592653589793, values (bits % 1000) / 1000 - 0.5), C zeroed
  before every timed call (:349-355);
* one warm call, then `reps` timed calls of the Fortran symbol
  (sgemm_/dgemm_/cgemm_/zgemm_) through the C ABI; best/median/p95/max
  wall-clock nanoseconds with the reference's p95 index (:356-364);
* the CSV header and row format of gemm_sweep.zig:505 and :380-398; GFLOP/s =
  2MNK (8MNK complex) / best;
* `--check` compares sampled entries (all of them for M N <= 4096, otherwise a
  5 x 5 grid, :247-286) with a float64 sum within 1e-9 K (1e-3 K for single
  precision).

Operands are device-resident by default (the GPU tier's in-place route);
`--host` passes host arrays, so each call includes the staging copies (the
host-pointer offload route).  Every call returns with C final, as BLAS
requires, so the times include launch and synchronisation.

    python -m zynum_amd.sweep --csv sweep.csv [--kind dgemm] [--trans NN] \
        [--shape label:m:n:k] [--reps 5] [--check] [--host]
"""
from __future__ import annotations

import argparse
import ctypes as C
import sys
import time
from pathlib import Path

import numpy as np

from . import _lib

DEFAULT_SHAPES = [
    ("m1_n1_k1", 1, 1, 1), ("m8_n8_k8", 8, 8, 8), ("m31_n31_k31", 31, 31, 31), ("m33_n33_k33", 33, 33, 33),
    ("sq64", 64, 64, 64), ("sq96", 96, 96, 96), ("sq128", 128, 128, 128), ("sq192", 192, 192, 192),
    ("sq256", 256, 256, 256), ("sq384", 384, 384, 384), ("sq512", 512, 512, 512), ("sq768", 768, 768, 768),
    ("sq1024", 1024, 1024, 1024), ("m63_n65_k17", 63, 65, 17), ("m65_n63_k33", 65, 63, 33),
    ("m127_n129_k31", 127, 129, 31), ("m129_n127_k33", 129, 127, 33), ("m1_n4096_k256", 1, 4096, 256),
    ("m4096_n1_k256", 4096, 1, 256), ("m1024_n64_k1024", 1024, 64, 1024), ("m2048_n64_k512", 2048, 64, 512),
    ("m4096_n32_k256", 4096, 32, 256), ("m2048_n17_k257", 2048, 17, 257), ("m512_n64_k2048", 512, 64, 2048),
    ("m64_n1024_k1024", 64, 1024, 1024), ("m64_n2048_k512", 64, 2048, 512), ("m32_n4096_k256", 32, 4096, 256),
    ("m17_n2048_k257", 17, 2048, 257), ("m64_n512_k2048", 64, 512, 2048), ("m1024_n1024_k64", 1024, 1024, 64),
    ("m1024_n1024_k128", 1024, 1024, 128), ("m1024_n1024_k256", 1024, 1024, 256),
    ("m256_n256_k2048", 256, 256, 2048), ("m128_n128_k4096", 128, 128, 4096), ("m1536_n256_k256", 1536, 256, 256),
    ("m256_n1536_k256", 256, 1536, 256), ("m512_n256_k768", 512, 256, 768), ("m256_n512_k768", 256, 512, 768),
    ("m768_n512_k256", 768, 512, 256), ("m512_n768_k256", 512, 768, 256), ("m384_n640_k96", 384, 640, 96),
    ("m640_n384_k96", 640, 384, 96),
]
KINDS = {  # kind -> (numpy dtype, Fortran symbol, flop factor)
    "sgemm": (np.float32, "sgemm_", 2.0),
    "dgemm": (np.float64, "dgemm_", 2.0),
    "cgemm": (np.complex64, "cgemm_", 8.0),
    "zgemm": (np.complex128, "zgemm_", 8.0),
}
HEADER = "kind,transa,transb,shape_index,label,m,n,k,library,gflops,best_ns,median_ns,p95_ns,max_ns,reps,check\n"


def lcg_values(count: int) -> np.ndarray:
    """The harness's fill sequence: state <- state * a + c (mod 2^64) from the
    fixed seed, value = ((state >> 32) % 1000) / 1000 - 0.5, computed in closed
    form (state_i = a^i s + c (1 + a + ... + a^(i-1)))."""
    a = np.uint64(6364136223846793005)
    c = np.uint64(1442695040888963407)
    seed = np.uint64(0x3141592653589793)
    if count <= 0:
        return np.zeros(0)
    with np.errstate(over="ignore"):
        powers = np.cumprod(np.full(count, a, dtype=np.uint64), dtype=np.uint64)  # a^1 .. a^count
        geo = np.cumsum(np.concatenate([np.ones(1, np.uint64), powers[:-1]]), dtype=np.uint64)  # sum_{j<i} a^j
        states = powers * seed + c * geo
    bits = (states >> np.uint64(32)) & np.uint64(0xFFFFFFFF)
    return (bits % np.uint64(1000)).astype(np.float64) / 1000.0 - 0.5


def fill(count: int, dtype) -> np.ndarray:
    """gemm_sweep.zig:162-174 (complex: re then im from consecutive draws)."""
    if np.dtype(dtype).kind == "c":
        v = lcg_values(2 * count)
        return (v[0::2] + 1j * v[1::2]).astype(dtype)
    return lcg_values(count).astype(dtype)


def stored_ld(rows: int, cols: int, trans: str) -> int:
    return max(1, rows if trans == "N" else cols)


def stored_cols(rows: int, cols: int, trans: str) -> int:
    return cols if trans == "N" else rows


def sample_index(slot: int, n: int):
    if n == 0:
        return None
    return [0, n // 4, n // 2, (n - 1) - (n - 1) // 4, n - 1][slot]


def check_samples(a, b, c, m, n, k, ta, tb, lda, ldb, ldc, single: bool) -> str:
    tol = (1e-3 if single else 1e-9) * max(1, k)
    if m * n <= 4096:
        cells = [(i, j) for j in range(n) for i in range(m)]
    else:
        cells = []
        for rs in range(5):
            for cs in range(5):
                i, j = sample_index(rs, m), sample_index(cs, n)
                if i is not None and j is not None and (i, j) not in cells:
                    cells.append((i, j))
    wide = np.complex128 if a.dtype.kind == "c" else np.float64
    for i, j in cells:
        arow = a[i + np.arange(k) * lda] if ta == "N" else a[np.arange(k) + i * lda]
        if ta == "C":
            arow = np.conj(arow)
        bcol = b[np.arange(k) + j * ldb] if tb == "N" else b[j + np.arange(k) * ldb]
        if tb == "C":
            bcol = np.conj(bcol)
        want = np.sum(arow.astype(wide) * bcol.astype(wide))
        if abs(complex(want) - complex(c[i + j * ldc])) > tol:
            return "check-failed"
    return "checked-ok"


def bench_one(torch, lib, kind: str, shape, ta: str, tb: str, reps: int, check: bool, host: bool):
    dtype, sym, _ = KINDS[kind]
    label, m, n, k = shape
    lda, ldb, ldc = stored_ld(m, k, ta), stored_ld(k, n, tb), max(1, m)
    a = fill(lda * stored_cols(m, k, ta), dtype)
    b = fill(ldb * stored_cols(k, n, tb), dtype)
    c = np.zeros(ldc * n, dtype=dtype)
    if host:
        pa, pb, pc = a.ctypes.data, b.ctypes.data, c.ctypes.data
        dc = None
    else:
        da, db, dc = (torch.from_numpy(x).cuda() for x in (a, b, c))
        pa, pb, pc = da.data_ptr(), db.data_ptr(), dc.data_ptr()
    one = np.ones(1, dtype=dtype)
    zero = np.zeros(1, dtype=dtype)
    if dtype in (np.float32, np.float64):  # real scalars by typed pointer, complex ones as addresses
        ptr = C.POINTER(C.c_float if dtype == np.float32 else C.c_double)
        alpha, beta = one.ctypes.data_as(ptr), zero.ctypes.data_as(ptr)
    else:
        alpha, beta = one.ctypes.data, zero.ctypes.data
    I = C.c_int32
    args = (ta.encode(), tb.encode(), C.byref(I(m)), C.byref(I(n)), C.byref(I(k)), alpha, pa, C.byref(I(lda)), pb,
            C.byref(I(ldb)), beta, pc, C.byref(I(ldc)))
    fn = getattr(lib, sym)
    fn(*args)
    if lib.zynum_gpu_last_status() != 0:
        raise RuntimeError(f"{kind} {label}: {_lib.status_string(lib.zynum_gpu_last_status())}")
    status = "unchecked"
    if check:
        out = c if host else dc.cpu().numpy()
        status = check_samples(a, b, out, m, n, k, ta, tb, lda, ldb, ldc, dtype in (np.float32, np.complex64))
    times = []
    for _ in range(reps):
        if host:
            c[:] = 0
        else:
            dc.zero_()
            torch.cuda.synchronize()
        t0 = time.perf_counter_ns()
        fn(*args)
        t1 = time.perf_counter_ns()
        times.append(max(1, t1 - t0))
    times.sort()
    p95 = min(len(times) - 1, (len(times) * 95 + 99) // 100 - 1)
    return times[0], times[len(times) // 2], times[p95], times[-1], status


def csv_row(kind, ta, tb, index, shape, library, res, reps) -> str:
    label, m, n, k = shape
    best, med, p95, mx, status = res
    gflops = KINDS[kind][2] * m * n * k / (best / 1e9) / 1e9
    return (f'{kind},{ta},{tb},{index},"{label}",{m},{n},{k},"{library}",{gflops:.6f},{best},{med},{p95},{mx},'
            f'{reps},"{status}"\n')


def parse_shape(spec: str):
    label, m, n, k = spec.split(":")
    return (label, int(m), int(n), int(k))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--csv", required=True, type=Path)
    ap.add_argument("--kind", action="append", choices=sorted(KINDS), help="repeatable; default: all four")
    ap.add_argument("--trans", action="append", help="pair such as NN, NT, CN (repeatable; default NN)")
    ap.add_argument("--shape", action="append", type=parse_shape, help="label:m:n:k (repeatable)")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--check", action="store_true")
    ap.add_argument("--host", action="store_true", help="host operands (staged by the library on every call)")
    ap.add_argument("--library", default="Zynum-MI355X")
    args = ap.parse_args(argv)
    if args.reps <= 0:
        ap.error("--reps must be positive")
    import torch

    if not torch.cuda.is_available():
        print("the sweep needs a GPU", file=sys.stderr)
        return 2
    lib = _lib.load()
    kinds = args.kind or ["sgemm", "dgemm", "cgemm", "zgemm"]
    pairs = [p.upper() for p in (args.trans or ["NN"])]
    shapes = args.shape or DEFAULT_SHAPES
    failed = 0
    with args.csv.open("w") as out:
        out.write(HEADER)
        for index, shape in enumerate(shapes):
            for pair in pairs:
                print(f"[{index + 1}/{len(shapes)}] {shape[0]} trans={pair} m={shape[1]} n={shape[2]} k={shape[3]}",
                      file=sys.stderr, flush=True)
                for kind in kinds:
                    if kind in ("sgemm", "dgemm") and "C" in pair:
                        continue  # gemm_sweep.zig:205-208
                    res = bench_one(torch, lib, kind, shape, pair[0], pair[1], args.reps, args.check, args.host)
                    failed += res[4] == "check-failed"
                    out.write(csv_row(kind, pair[0], pair[1], index, shape, args.library, res, args.reps))
    return 1 if failed else 0


if __name__ == "__main__":
    sys.exit(main())
