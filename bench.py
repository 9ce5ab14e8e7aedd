#!/usr/bin/env python3
"""DGEMM throughput on one MI355X: FP64 MFMA at 8192^3 (headline), with the
split-precision (INT8, Ozaki) form beside it and the restated reference CPU
path timed on the host as a reported baseline.

One step = one cblas_dgemm-equivalent call (request ABI, NN, M=N=K=8192,
lda=ldb=ldc=8200, alpha=1, beta=0) on inputs resident in HBM.  Prints one
JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FP64_MFMA_PEAK_TF = 78.6
INT8_PEAK_TOPS = 5000.0
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse_args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--size", type=int, default=8192, help="M=N=K (the workload is 8192)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--no-split", action="store_true", help="skip the split-precision leg")
    p.add_argument("--no-variants", action="store_true", help="skip the transpose/beta/SGEMM variant table")
    p.add_argument("--no-grouped", action="store_true", help="skip the grouped uneven-GEMM leg")
    p.add_argument("--no-level3", action="store_true", help="skip the DTRSM/DSYRK leg")
    p.add_argument("--no-complex", action="store_true", help="skip the ZGEMM/CGEMM leg")
    p.add_argument("--no-offload", action="store_true", help="skip the host-pointer offload leg")
    p.add_argument("--only-complex", action="store_true", help="headline plus the ZGEMM/CGEMM leg (profiler runs)")
    p.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline sample time")
    p.add_argument("--only-headline", action="store_true", help="plain FP64 leg only (for profiler runs)")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="before the workload, read 4 GiB once and copy 1 GiB once (known HBM bytes for "
                        "scaling the FETCH_SIZE / WRITE_SIZE counters)")
    return p.parse_args()


def read_traffic(name: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration), kept in profiles/<name>."""
    f = ROOT / "profiles" / name
    if not f.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", f.read_text())
    return float(m.group(1)) if m else None


def main() -> int:
    args = parse_args()
    if args.only_headline:
        args.no_cpu = args.no_split = args.no_variants = args.no_grouped = args.no_level3 = args.no_complex = True
        args.no_offload = True
    if args.only_complex:
        args.no_cpu = args.no_split = args.no_variants = args.no_grouped = args.no_level3 = args.no_offload = True
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is a design note, "
                                   "not a measured configuration", "n_gpus": args.gpus}))
        return 2

    import torch

    from zynum_amd import _lib, blas, synth

    lib = _lib.load()
    assert torch.cuda.is_available(), "bench needs a GPU"
    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")
    n = args.size
    ld = n + 8
    flop = 2.0 * n * n * n
    if args.calibrate_copy:
        big = torch.ones(1 << 29, dtype=torch.float64, device=dev)  # 4 GiB, larger than the 256 MB MALL
        flush = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        flush.mul_(2.0)  # evict big from the caches
        log(f"calibration read sum {float(big.sum()):.0f} (4 GiB read)")
        x = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        y = torch.empty_like(x)
        flush.mul_(2.0)
        y.copy_(x)  # 1 GiB read + 1 GiB write
        torch.cuda.synchronize()
        del big, flush, x, y

    log(f"inputs: DGEMM {n}^3 NN ld={ld}, splitmix64 seed 0")
    t0 = time.time()
    a_h = synth.matrix(0, 0, n, n, ld)
    b_h = synth.matrix(0, n * n, n, n, ld)
    A = torch.from_numpy(a_h).to(dev)
    B = torch.from_numpy(b_h).to(dev)
    Cm = torch.full((ld * n,), -999.0, dtype=torch.float64, device=dev)
    del a_h, b_h
    log(f"inputs ready in {time.time() - t0:.1f}s")

    stream = torch.cuda.Stream(device=dev)
    sh = stream.cuda_stream

    def step(mode=_lib.MODE_PLAIN, param=0, ta=False, tb=False, beta=0.0, c=None, scalar=_lib.F64, a=None, b=None):
        a = A if a is None else a
        b = B if b is None else b
        c = Cm if c is None else c
        st = blas.gemm_request(ta, tb, n, n, n, 1.0, a.data_ptr(), ld, b.data_ptr(), ld, beta, c.data_ptr(), ld,
                               scalar, stream=sh, mode=mode, mode_param=param, flags=_lib.FLAG_ASYNC)
        _lib.check(st, "bench gemm")

    # ---------------- headline: plain FP64 MFMA -----------------
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    with torch.cuda.stream(stream):
        for s, e in ev:
            s.record(stream)
            step()
            e.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    per_launch_ms = [s.elapsed_time(e) for s, e in ev]
    avg_ms = float(np.mean(per_launch_ms))
    best_ms = float(np.min(per_launch_ms))
    ms_per_step = wall * 1e3 / args.steps
    value = flop * args.steps / wall / 1e12
    achieved = flop / (avg_ms * 1e-3) / 1e12
    log(f"plain FP64: {value:.2f} TFLOP/s wall, kernel avg {avg_ms:.3f} ms ({achieved:.2f} TF), best {best_ms:.3f} ms")

    result = {
        "metric": "DGEMM TFLOP/s @8192^3: FP64-MFMA vs split-INT8 at <=FP64 error; % of peak",
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (splitmix64 seed 0, uniform [-1,1))",
        "config": {
            "workload": f"DGEMM M=N=K={n} NoTrans/NoTrans lda=ldb=ldc={ld} alpha=1 beta=0 via the request ABI "
                        "(cblas_dgemm-equivalent), inputs resident in HBM",
            "m": n, "n": n, "k": n, "ld": ld,
        },
        "roofline": {
            "bound": "mfma",
            "achieved": round(achieved, 3),
            "peak": FP64_MFMA_PEAK_TF,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP64_MFMA_PEAK_TF, 4),
            "traffic": read_traffic("dgemm_f64_8192_nn_traffic.txt"),
            "kernel": "gemm_f64_dma_kernel (one launch per step)",
            "kernel_avg_ms": round(avg_ms, 4),
            "algorithmic_flop_per_launch": flop,
        },
    }

    # torch work (allocations, copies, reductions) and the library calls share
    # one stream, so every check reads finished results
    with torch.cuda.stream(stream):
        # ---------------- variants (transposes, beta, SGEMM) -----------------
        if not args.no_variants:
            variants = {}
            C1 = torch.randn(ld * n, dtype=torch.float64, device=dev)
            for ta, tb in ((False, False), (False, True), (True, False), (True, True)):
                for beta in (0.0, 1.0):
                    name = f"{'T' if ta else 'N'}{'T' if tb else 'N'}_beta{int(beta)}"
                    step(ta=ta, tb=tb, beta=beta, c=C1)
                    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    s.record(stream)
                    for _ in range(3):
                        step(ta=ta, tb=tb, beta=beta, c=C1)
                    e.record(stream)
                    torch.cuda.synchronize()
                    variants[f"dgemm_{name}"] = round(3 * flop / (s.elapsed_time(e) * 1e-3) / 1e12, 2)
            del C1
            A32, B32 = A.float(), B.float()
            C32 = torch.randn(ld * n, dtype=torch.float32, device=dev)
            for ta, tb in ((False, False), (False, True), (True, False), (True, True)):
                for beta in (0.0, 1.0):
                    name = f"{'T' if ta else 'N'}{'T' if tb else 'N'}_beta{int(beta)}"
                    step(scalar=_lib.F32, a=A32, b=B32, c=C32, ta=ta, tb=tb, beta=beta)
                    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    s.record(stream)
                    for _ in range(5):
                        step(scalar=_lib.F32, a=A32, b=B32, c=C32, ta=ta, tb=tb, beta=beta)
                    e.record(stream)
                    torch.cuda.synchronize()
                    variants[f"sgemm_{name}"] = round(5 * flop / (s.elapsed_time(e) * 1e-3) / 1e12, 2)
            variants["sgemm_frac_of_157.3TF"] = [round(min(v for k, v in variants.items() if k.startswith("sgemm_"))
                                                       / 157.3, 4),
                                                 round(max(v for k, v in variants.items() if k.startswith("sgemm_"))
                                                       / 157.3, 4)]
            del A32, B32, C32
            result["variants_tflops"] = variants
            log(f"variants: {variants}")

        # ---------------- configuration 1: synchronous cblas_dgemm 1024^3 -----------------
        if not args.no_variants:
            try:
                result["cblas_dgemm_1024"] = config1_leg(torch, lib)
            except Exception as exc:
                result["cblas_dgemm_1024"] = {"error": str(exc)}
                log(f"cblas_dgemm 1024 leg failed: {exc}")

        # ---------------- split precision (INT8 / CRT) -----------------
        if not args.no_split:
            try:
                result["split_int8"] = split_leg(torch, lib, A, B, n, ld, stream, achieved)
            except Exception as exc:  # reported, never hidden
                result["split_int8"] = {"error": str(exc)}
                log(f"split leg failed: {exc}")

        # ---------------- grouped uneven GEMM -----------------
        if not args.no_grouped:
            try:
                result["grouped"] = grouped_leg(torch, stream, achieved)
            except Exception as exc:
                result["grouped"] = {"error": str(exc)}
                log(f"grouped leg failed: {exc}")

        # ---------------- DTRSM / DSYRK -----------------
        if not args.no_level3:
            try:
                result["level3"] = level3_leg(torch, lib, n, stream)
            except Exception as exc:
                result["level3"] = {"error": str(exc)}
                log(f"level3 leg failed: {exc}")

        # ---------------- complex GEMM -----------------
        if not args.no_complex:
            try:
                result["complex"] = complex_leg(torch, lib, n, stream, achieved)
            except Exception as exc:
                result["complex"] = {"error": str(exc)}
                log(f"complex leg failed: {exc}")

    # ---------------- host-pointer offload (GPU side) -----------------
    if not args.no_offload:
        try:
            result["offload"] = offload_leg(torch, lib)
        except Exception as exc:
            result["offload"] = {"error": str(exc)}
            log(f"offload leg failed: {exc}")

    # ---------------- CPU baseline -----------------
    if not args.no_cpu:
        edges = [] if args.no_offload or "error" in result["offload"] else OFFLOAD_EDGES
        result["cpu_baseline"] = cpu_leg(n, ld, args.cpu_seconds, edges)
        if edges:
            offload_crossover(lib, result["offload"], result["cpu_baseline"]["by_edge"])

    print(json.dumps(result), flush=True)
    return 0


def config1_leg(torch, lib, reps=20):
    """cblas_dgemm(ColMajor, NoTrans, NoTrans, 1024, 1024, 1024, 1.5, A, 1024,
    B, 1024, 0.5, C, 1024) on device operands (A, B, C uniform [-1, 1) from
    seed 0), timed by the host clock around each synchronous call: the wall
    time a BLAS caller sees, launch and synchronisation included."""
    from zynum_amd import synth

    n = 1024
    dev = torch.device("cuda:0")
    a = torch.from_numpy(synth.matrix(0, 0, n, n, n)).to(dev)
    b = torch.from_numpy(synth.matrix(0, n * n, n, n, n)).to(dev)
    c0 = torch.from_numpy(synth.matrix(0, 2 * n * n, n, n, n)).to(dev)
    c = c0.clone()
    torch.cuda.synchronize()

    def call():
        lib.cblas_dgemm(102, 111, 111, n, n, n, 1.5, a.data_ptr(), n, b.data_ptr(), n, 0.5, c.data_ptr(), n)

    call()
    if lib.zynum_gpu_last_status() != 0:
        raise RuntimeError("cblas_dgemm did not run on the GPU tier")
    times = []
    for _ in range(reps):
        c.copy_(c0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        call()
        times.append(time.perf_counter() - t0)
    times.sort()
    flop = 2.0 * n ** 3
    out = {"workload": "cblas_dgemm ColMajor NN M=N=K=1024 ld=1024 alpha=1.5 beta=0.5, device operands, "
                       f"synchronous calls, host clock, {reps} calls after 1 warm-up",
           "best_us": round(times[0] * 1e6, 1), "median_us": round(times[len(times) // 2] * 1e6, 1),
           "best_gflops": round(flop / times[0] / 1e9, 1),
           "median_gflops": round(flop / times[len(times) // 2] / 1e9, 1)}
    log(f"cblas_dgemm 1024: {out}")
    return out


def split_leg(torch, lib, A, B, n, ld, stream, plain_tf):
    """Split-precision DGEMM on the wide-range input (a_ij = u 2^e, e in
    [-16, 16]); errors of both forms against the device double-double product."""
    from zynum_amd import _lib, blas, synth

    dev = A.device
    sh = stream.cuda_stream
    log("split leg: wide-range inputs")
    Aw = torch.from_numpy(synth.wide_range(7, 0, n, n, ld)).to(dev)
    Bw = torch.from_numpy(synth.wide_range(7, n * n, n, n, ld)).to(dev)
    ref_hi = torch.zeros(ld * n, dtype=torch.float64, device=dev)
    ref_lo = torch.zeros(ld * n, dtype=torch.float64, device=dev)
    st = lib.zynum_gpu_dd_gemm(b"N", b"N", n, n, n, 1.0, Aw.data_ptr(), ld, Bw.data_ptr(), ld, 0.0, None, ld,
                               ref_hi.data_ptr(), ref_lo.data_ptr(), ld, sh)
    _lib.check(st, "dd reference")
    flop = 2.0 * n ** 3

    def run(mode, param, reps, warm=1):
        C = torch.zeros(ld * n, dtype=torch.float64, device=dev)
        for _ in range(warm):  # untimed
            st = blas.gemm_request(False, False, n, n, n, 1.0, Aw.data_ptr(), ld, Bw.data_ptr(), ld, 0.0,
                                   C.data_ptr(), ld, _lib.F64, stream=sh, mode=mode, mode_param=param,
                                   flags=_lib.FLAG_ASYNC)
            _lib.check(st, f"mode {mode}")
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record(stream)
        for _ in range(reps):
            st = blas.gemm_request(False, False, n, n, n, 1.0, Aw.data_ptr(), ld, Bw.data_ptr(), ld, 0.0,
                                   C.data_ptr(), ld, _lib.F64, stream=sh, mode=mode, mode_param=param,
                                   flags=_lib.FLAG_ASYNC)
            _lib.check(st, f"mode {mode}")
        e.record(stream)
        torch.cuda.synchronize()
        ms = s.elapsed_time(e) / reps
        v = C.view(n, ld)[:, :n]
        err = ((v - ref_hi.view(n, ld)[:, :n]) - ref_lo.view(n, ld)[:, :n]).abs()
        mag = ref_hi.view(n, ld)[:, :n].abs()
        out = {"ms": round(ms, 4), "tflops_fp64_equiv": round(flop / (ms * 1e-3) / 1e12, 3),
               "max_abs_err_vs_dd": float(err.max()), "max_rel_err_vs_dd": float((err / mag.clamp_min(1e-300)).max()),
               "normwise_err_vs_dd": float(err.max() / mag.max())}
        del C
        return out

    plain = run(_lib.MODE_PLAIN, 0, 3)
    out = {"workload": f"DGEMM {n}^3 NN, a_ij = u*2^e (u uniform [-1,1), e uniform in [-16,16]), seed 7",
           "plain_fp64": plain}
    best = None
    for moduli in (13, 14, 15, 16):
        r = run(_lib.MODE_OZAKI_CRT, moduli, 5, warm=3)
        r["int8_gemms"] = moduli
        r["int8_tops_executed"] = round(moduli * flop / (r["ms"] * 1e-3) / 1e12, 1)
        out[f"crt_{moduli}"] = r
        if r["max_abs_err_vs_dd"] <= plain["max_abs_err_vs_dd"] and (best is None or r["ms"] < best[1]["ms"]):
            best = (moduli, r)
    for slices in (4, 5, 6, 7, 8):
        r = run(_lib.MODE_OZAKI_INT8, slices, 1)
        pairs = slices * (slices + 1) // 2  # slice products with i + j <= s + 1
        r["int8_slice_products"] = pairs
        r["int8_tops_executed"] = round(pairs * flop / (r["ms"] * 1e-3) / 1e12, 1)
        out[f"slices_int8_{slices}"] = r
    for slices in (6, 7, 8):
        r = run(_lib.MODE_OZAKI_BF16, slices, 1)
        pairs = slices * (slices + 1) // 2
        r["bf16_slice_products"] = pairs
        r["bf16_tflops_executed"] = round(pairs * flop / (r["ms"] * 1e-3) / 1e12, 1)
        out[f"slices_bf16_{slices}"] = r
    if best is not None:
        out["value"] = best[1]["tflops_fp64_equiv"]
        out["unit"] = "TFLOP/s (FP64-equivalent 2MNK)"
        out["form"] = f"Ozaki scheme II, {best[0]} INT8 residue GEMMs, INT32 accumulation"
        out["speedup_vs_plain"] = round(best[1]["tflops_fp64_equiv"] / plain["tflops_fp64_equiv"], 3)
        out["error_le_plain"] = True
    return out


def grouped_leg(torch, stream, plain_tf):
    """4096 problems, M, N, K log-uniform in [16, 512] (seed 0), one persistent
    graph-replayed launch per batch vs a loop of single-GEMM launches."""
    from zynum_amd import grouped

    out = {"workload": "4096 independent GEMMs, M,N,K log-uniform integers in [16,512] (seed 0), NN, "
                       "ld = rows, alpha=1, beta=0, entries uniform [-1,1)"}
    shapes = grouped.draw_shapes(4096, 0, 16, 512)
    sh = stream.cuda_stream
    for name, dt in (("f64", torch.float64), ("f32", torch.float32)):
        b = grouped.Batch(torch, shapes, dt, seed=0)
        plan = b.plan()
        info = plan.info()
        plan.graph_run(sh)
        torch.cuda.synchronize()

        def timed(fn, reps):
            s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s.record(stream)
            for _ in range(reps):
                fn()
            e.record(stream)
            torch.cuda.synchronize()
            return s.elapsed_time(e) / reps

        g_ms = timed(lambda: plan.graph_run(sh), 10)
        l_ms = timed(lambda: b.loop(sh), 2)
        r = {"flop": b.flops, "tiles": info["tiles"], "workgroups": info["workgroups"],
             "graph_ms": round(g_ms, 4), "tflops": round(b.flops / (g_ms * 1e-3) / 1e12, 2),
             "loop_ms": round(l_ms, 4), "loop_tflops": round(b.flops / (l_ms * 1e-3) / 1e12, 2),
             "speedup_vs_loop": round(l_ms / g_ms, 2),
             "hbm_gbs_algorithmic": round(b.bytes / (g_ms * 1e-3) / 1e9, 1)}
        if name == "f64":
            r["frac_of_large_dgemm"] = round(r["tflops"] / plain_tf, 3)
        out[name] = r
        plan.close()
        del b
        log(f"grouped {name}: {r}")
    return out


def level3_leg(torch, lib, n, stream):
    """DTRSM Left/Lower/NoTrans/NonUnit N=8192 with 8192 right-hand sides
    (diagonal 8 + |u|) and DSYRK Lower NoTrans N=8192, K=4096, beta=0."""
    from zynum_amd import _lib, synth

    dev = torch.device("cuda:0")
    sh = stream.cuda_stream
    out = {}
    # --- DTRSM ---
    L = torch.from_numpy(synth.matrix(11, 0, n, n, n)).to(dev).view(n, n)  # row j = column j
    L = torch.triu(L)  # column-major lower triangle == row-major upper
    idx = torch.arange(n, device=dev)
    L[idx, idx] = 8.0 + torch.from_numpy(synth.uniform(11, n * n, n)).to(dev).abs()
    B = torch.from_numpy(synth.matrix(11, n * n + n, n, n, n)).to(dev)
    flop = float(n) * n * n
    res = {"workload": f"DTRSM L/L/N/N M=N={n}, ld={n}, alpha=1, recursive blocking down to 256-row block "
                       "solves (64-row FP64 quad solves, MFMA coupling updates)",
           "timing": "median of 5 calls after one warm-up, B restored before each call outside the timed region"}
    for mode, param, tag in ((_lib.MODE_PLAIN, 0, "plain_fp64"), (_lib.MODE_OZAKI_CRT, 14, "crt14_updates")):
        X = B.clone()
        st = lib.zynum_gpu_dtrsm(b"L", b"L", b"N", b"N", n, n, 1.0, L.data_ptr(), n, X.data_ptr(), n, sh, mode,
                                 param, 0)
        _lib.check(st, "dtrsm")
        times = []
        for _ in range(5):
            X.copy_(B)
            torch.cuda.synchronize()
            s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s.record(stream)
            st = lib.zynum_gpu_dtrsm(b"L", b"L", b"N", b"N", n, n, 1.0, L.data_ptr(), n, X.data_ptr(), n, sh, mode,
                                     param, 0)
            e.record(stream)
            torch.cuda.synchronize()
            _lib.check(st, "dtrsm")
            times.append(s.elapsed_time(e))
        ms = sorted(times)[len(times) // 2]
        # residual L X - B in double-double on the device (the dd product is
        # independent of the GEMM core), against the backward-error scale
        # n u (|L||X|) formed with the plain GEMM on absolute values
        R_hi = torch.empty_like(B)
        R_lo = torch.empty_like(B)
        st = lib.zynum_gpu_dd_gemm(b"N", b"N", n, n, n, 1.0, L.data_ptr(), n, X.data_ptr(), n, -1.0, B.data_ptr(), n,
                                   R_hi.data_ptr(), R_lo.data_ptr(), n, sh)
        _lib.check(st, "dd residual")
        R = (R_hi + R_lo).abs()
        del R_hi, R_lo
        La, Xa = L.abs(), X.abs()
        scale = torch.zeros_like(B)
        from zynum_amd import blas
        _lib.check(blas.gemm_request(False, False, n, n, n, 1.0, La.data_ptr(), n, Xa.data_ptr(), n, 0.0,
                                     scale.data_ptr(), n, _lib.F64, stream=sh), "residual scale")
        torch.cuda.synchronize()
        del La, Xa
        nrm = float(L.abs().sum(dim=1).max()) * float(X.abs().max())
        res[tag] = {"ms": round(ms, 3), "tflops": round(flop / (ms * 1e-3) / 1e12, 2),
                    "residual_max_dd": float(R.max()), "residual_normwise": float(R.max()) / nrm,
                    "residual_over_n_u_LX_max": float((R / (n * 2.0**-53 * scale)).max())}
        del X, R, scale
    out["dtrsm"] = res
    del L, B
    # --- DSYRK ---
    k = n // 2
    A = torch.from_numpy(synth.matrix(12, 0, n, k, n)).to(dev)
    ref_hi = torch.zeros(n * n, dtype=torch.float64, device=dev)
    ref_lo = torch.zeros(n * n, dtype=torch.float64, device=dev)
    st = lib.zynum_gpu_dd_gemm(b"N", b"T", n, n, k, 1.0, A.data_ptr(), n, A.data_ptr(), n, 0.0, None, n,
                               ref_hi.data_ptr(), ref_lo.data_ptr(), n, sh)
    _lib.check(st, "dd syrk reference")
    flop = float(n) * n * k
    res = {"workload": f"DSYRK Lower NoTrans N={n} K={k}, beta 0 and 1, lower-triangle tiles of the GEMM core "
                       "(best of 3 calls)"}
    low = torch.tril(torch.ones(n, n, dtype=torch.bool, device=dev)).T  # column-major lower in row-major view
    for mode, param, tag, beta in ((_lib.MODE_PLAIN, 0, "plain_fp64", 0.0), (_lib.MODE_PLAIN, 0, "plain_fp64_beta1", 1.0),
                                   (_lib.MODE_OZAKI_CRT, 14, "crt14", 0.0)):
        C = torch.zeros(n * n, dtype=torch.float64, device=dev)
        st = lib.zynum_gpu_dsyrk(b"L", b"N", n, k, 1.0, A.data_ptr(), n, beta, C.data_ptr(), n, sh, mode, param, 0)
        _lib.check(st, "dsyrk")
        C.zero_()  # beta = 1 then adds the product to exact zeros: the result is the product
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record(stream)
        st = lib.zynum_gpu_dsyrk(b"L", b"N", n, k, 1.0, A.data_ptr(), n, beta, C.data_ptr(), n, sh, mode, param, 0)
        e.record(stream)
        torch.cuda.synchronize()
        _lib.check(st, "dsyrk")
        ms = s.elapsed_time(e)
        err = ((C - ref_hi) - ref_lo).view(n, n)[low].abs().max()
        for _ in range(2):  # two more timed calls (beta = 1 accumulates; only the first is checked)
            s.record(stream)
            st = lib.zynum_gpu_dsyrk(b"L", b"N", n, k, 1.0, A.data_ptr(), n, beta, C.data_ptr(), n, sh, mode, param,
                                     0)
            e.record(stream)
            torch.cuda.synchronize()
            _lib.check(st, "dsyrk")
            ms = min(ms, s.elapsed_time(e))
        res[tag] = {"ms": round(ms, 3), "tflops": round(flop / (ms * 1e-3) / 1e12, 2),
                    "max_abs_err_vs_dd": float(err), "beta": beta}
        del C
    out["dsyrk"] = res
    del A, ref_hi, ref_lo
    out.update(level3_more(torch, lib, n, stream))
    log(f"level3: {out}")
    return out


def level3_more(torch, lib, n, stream, S=1024):
    """DSYMM Left/Lower M=N=n, DTRMM Left/Lower/NoTrans/NonUnit M=N=n and DSYR2K
    Lower/NoTrans N=n, K=n/2 (beta 0), timed over 3 calls each; errors on an
    S x S corner against the device double-double product of the expanded
    operands (torch builds the expansions for the check only)."""
    from zynum_amd import _lib, synth

    dev = torch.device("cuda:0")
    sh = stream.cuda_stream
    out = {}

    def timed(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record(stream)
        for _ in range(reps):
            fn()
        e.record(stream)
        torch.cuda.synchronize()
        return s.elapsed_time(e) / reps

    def dd_corner(x, ldx, y, ldy, k):
        """dd of X[:S, :k] Y[:k, :S] (column-major device buffers) as (hi, lo) S x S."""
        hi = torch.zeros(S * S, dtype=torch.float64, device=dev)
        lo = torch.zeros(S * S, dtype=torch.float64, device=dev)
        _lib.check(lib.zynum_gpu_dd_gemm(b"N", b"N", S, S, k, 1.0, x.data_ptr(), ldx, y.data_ptr(), ldy, 0.0, None, S,
                                         hi.data_ptr(), lo.data_ptr(), S, sh), "dd reference")
        return hi.view(S, S), lo.view(S, S)  # [j, i]

    def err_of(got_cm, ld, hi, lo):
        g = got_cm.view(-1, ld)[:S, :S]
        e = ((g - hi) - lo).abs()
        return float(e.max()), float(e.max() / hi.abs().max())

    A = torch.from_numpy(synth.matrix(13, 0, n, n, n)).to(dev)           # column-major n x n
    B = torch.from_numpy(synth.matrix(13, n * n, n, n, n)).to(dev)
    # --- DSYMM: S from the lower triangle of A ---
    C = torch.zeros(n * n, dtype=torch.float64, device=dev)
    ms = timed(lambda: _lib.check(lib.zynum_gpu_dsymm(b"L", b"L", n, n, 1.0, A.data_ptr(), n, B.data_ptr(), n, 0.0,
                                                      C.data_ptr(), n, sh, _lib.FLAG_ASYNC), "dsymm"))
    At = A.view(n, n)                        # [j, i] = A(i, j)
    low = torch.triu(At)                     # keeps A(i, j) with i >= j
    Sfull = (low + low.T - torch.diag(torch.diagonal(At))).contiguous()   # symmetric: either orientation
    hi, lo = dd_corner(Sfull, n, B, n, n)
    e, ne = err_of(C, n, hi, lo)
    out["dsymm"] = {"workload": f"DSYMM Left Lower M=N={n}, alpha 1, beta 0", "ms": round(ms, 3),
                    "tflops": round(2.0 * n ** 3 / (ms * 1e-3) / 1e12, 2),
                    "max_abs_err_vs_dd_corner": e, "normwise_err_vs_dd_corner": ne}
    del Sfull, C, hi, lo
    # --- DTRMM: lower triangle of A, B in place (a fresh copy per call) ---
    X = B.clone()
    flop = float(n) * n * n  # m^2 n
    ms_total = timed(lambda: (X.copy_(B), _lib.check(lib.zynum_gpu_dtrmm(b"L", b"L", b"N", b"N", n, n, 1.0,
                                                                          A.data_ptr(), n, X.data_ptr(), n, sh,
                                                                          _lib.FLAG_ASYNC), "dtrmm")))
    Tl = torch.triu(At).contiguous()        # [j, i] = A(i, j) for i >= j: the lower triangle, column-major
    hi, lo = dd_corner(Tl, n, B, n, n)
    e, ne = err_of(X, n, hi, lo)            # X holds the last call's result
    ms_copy = timed(lambda: X.copy_(B))
    ms = ms_total - ms_copy
    out["dtrmm"] = {"workload": f"DTRMM Left Lower NoTrans NonUnit M=N={n}, alpha 1", "ms": round(ms, 3),
                    "tflops": round(flop / (ms * 1e-3) / 1e12, 2),
                    "max_abs_err_vs_dd_corner": e, "normwise_err_vs_dd_corner": ne}
    del X, Tl, hi, lo
    # --- DSYR2K: C_lower = A B^T + B A^T, A and B n x k ---
    k = n // 2
    C = torch.zeros(n * n, dtype=torch.float64, device=dev)
    ms = timed(lambda: _lib.check(lib.zynum_gpu_dsyr2k(b"L", b"N", n, k, 1.0, A.data_ptr(), n, B.data_ptr(), n, 0.0,
                                                       C.data_ptr(), n, sh, _lib.FLAG_ASYNC), "dsyr2k"))
    Bt = B.view(n, n)[:k, :].T.contiguous()            # B(:, :k)^T as column-major k x n
    At2 = A.view(n, n)[:k, :].T.contiguous()
    h1, l1 = dd_corner(A, n, Bt, k, k)
    h2, l2 = dd_corner(B, n, At2, k, k)
    hi, lo = h1 + h2, l1 + l2
    g = C.view(n, n)[:S, :S]
    lower_mask = torch.triu(torch.ones(S, S, dtype=torch.bool, device=dev))  # [j, i] with i >= j
    e_t = ((g - hi) - lo).abs()[lower_mask]
    out["dsyr2k"] = {"workload": f"DSYR2K Lower NoTrans N={n} K={k}, alpha 1, beta 0", "ms": round(ms, 3),
                     "tflops": round(2.0 * n * n * k / (ms * 1e-3) / 1e12, 2),
                     "max_abs_err_vs_dd_corner": float(e_t.max()),
                     "normwise_err_vs_dd_corner": float(e_t.max() / hi.abs()[lower_mask].max())}
    del A, B, C, Bt, At2
    return out


def complex_leg(torch, lib, n, stream, plain_tf):
    """ZGEMM / CGEMM M=N=K=n (uniform [-1,1) real and imaginary parts) through
    the request ABI: 8MNK real flops per call; errors on a 1024 x 1024 corner
    against the device double-double product of the real and imaginary parts."""
    from zynum_amd import _lib, blas

    dev = torch.device("cuda:0")
    sh = stream.cuda_stream
    ld = n + 8
    g = torch.Generator(device=dev)
    g.manual_seed(21)

    def cplx(dtype):
        re = torch.rand((n, ld), dtype=torch.float64, device=dev, generator=g) * 2 - 1
        im = torch.rand((n, ld), dtype=torch.float64, device=dev, generator=g) * 2 - 1
        return torch.complex(re, im).to(dtype)  # row j of the tensor = column j (column-major)

    A, B = cplx(torch.complex128), cplx(torch.complex128)
    flop = 8.0 * n ** 3
    out = {"workload": f"ZGEMM/CGEMM M=N=K={n}, ld={ld}, alpha=1, beta=0, one real MFMA GEMM on the 2x2-block "
                       "expansion of op(A) per call; inputs resident in HBM"}

    def timed(scalar, ta, tb, a, b, c, reps, alpha=1.0, beta=0.0):
        def go():
            st = blas.gemm_request(ta, tb, n, n, n, alpha, a.data_ptr(), ld, b.data_ptr(), ld, beta, c.data_ptr(),
                                   ld, scalar, stream=sh, flags=_lib.FLAG_ASYNC)
            _lib.check(st, "complex gemm")
        go()
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record(stream)
        for _ in range(reps):
            go()
        e.record(stream)
        torch.cuda.synchronize()
        return s.elapsed_time(e) / reps

    def corner_error(C, ta, tb, S=1024):
        """Max error of C[:S, :S] against dd(op(A) op(B)) built from real GEMMs."""
        Ar = A[:, :S] if ta == "N" else A[:S, :n].T     # (k, S): [p, i] = op(A)(i, p)
        Ai = Ar.imag * (-1.0 if ta == "C" else 1.0)
        Ar = Ar.real
        Ar, Ai = Ar.contiguous(), Ai.contiguous()        # (k, S): column-major S x k, ld S
        Bs = B[:S, :n] if tb == "N" else B[:, :S].T      # (S, k): [j, p] = op(B)(p, j), column-major k x S
        Bi = Bs.imag * (-1.0 if tb == "C" else 1.0)
        Br, Bi = Bs.real.contiguous(), Bi.contiguous()
        parts = {}
        for name, x, y in (("rr", Ar, Br), ("ii", Ai, Bi), ("ri", Ar, Bi), ("ir", Ai, Br)):
            hi = torch.zeros(S * S, dtype=torch.float64, device=dev)
            lo = torch.zeros(S * S, dtype=torch.float64, device=dev)
            st = lib.zynum_gpu_dd_gemm(b"N", b"N", S, S, n, 1.0, x.data_ptr(), S, y.data_ptr(), n, 0.0, None, S,
                                       hi.data_ptr(), lo.data_ptr(), S, sh)
            _lib.check(st, "dd reference")
            parts[name] = (hi.view(S, S), lo.view(S, S))
        ref_re = (parts["rr"][0] - parts["ii"][0]) + (parts["rr"][1] - parts["ii"][1])
        ref_im = (parts["ri"][0] + parts["ir"][0]) + (parts["ri"][1] + parts["ir"][1])
        got = C[:S, :S].to(torch.complex128)
        err = torch.maximum((got.real - ref_re).abs(), (got.imag - ref_im).abs())
        mag = torch.maximum(ref_re.abs(), ref_im.abs())
        return float(err.max()), float(err.max() / mag.max())

    for tag, ta, tb in (("zgemm_NN", "N", "N"), ("zgemm_CN", "C", "N"), ("zgemm_NT", "N", "T")):
        C = torch.zeros((n, ld), dtype=torch.complex128, device=dev)
        ms = timed(_lib.C64, ta, tb, A, B, C, 3)
        err, nerr = corner_error(C, ta, tb)
        out[tag] = {"ms": round(ms, 3), "tflops": round(flop / (ms * 1e-3) / 1e12, 2),
                    "frac_of_78.6TF": round(flop / (ms * 1e-3) / 1e12 / 78.6, 4),
                    "max_abs_err_vs_dd_1024_corner": err, "normwise_err_vs_dd_1024_corner": nerr}
        del C
    out["zgemm_NN"]["vs_plain_dgemm"] = round(out["zgemm_NN"]["tflops"] / plain_tf, 3)
    C = torch.zeros((n, ld), dtype=torch.complex128, device=dev)
    ms = timed(_lib.C64, "N", "N", A, B, C, 3, alpha=0.6 - 0.4j, beta=-0.2 + 0.3j)
    out["zgemm_NN_complex_alpha_beta"] = {"ms": round(ms, 3), "tflops": round(flop / (ms * 1e-3) / 1e12, 2)}
    # the library's own complex GEMM on the same operands, for reference (torch.matmul -> rocBLAS/hipBLASLt)
    At, Bt = A[:, :n].contiguous(), B[:, :n].contiguous()
    torch.matmul(Bt, At)
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(2):
        torch.matmul(Bt, At)
    e.record()
    torch.cuda.synchronize()
    out["torch_matmul_complex128_tflops"] = round(flop / (s.elapsed_time(e) / 2 * 1e-3) / 1e12, 2)
    del C, At, Bt
    A32, B32 = A.to(torch.complex64), B.to(torch.complex64)
    del A, B
    C32 = torch.zeros((n, ld), dtype=torch.complex64, device=dev)
    ms = timed(_lib.C32, "N", "N", A32, B32, C32, 3)
    out["cgemm_NN"] = {"ms": round(ms, 3), "tflops": round(flop / (ms * 1e-3) / 1e12, 2),
                       "frac_of_157.3TF": round(flop / (ms * 1e-3) / 1e12 / 157.3, 4)}
    del A32, B32, C32
    log(f"complex: {out}")
    return out


OFFLOAD_EDGES = (64, 128, 192, 256, 384, 512, 768, 1024, 1536, 2048)


def offload_leg(torch, lib):
    """Host-pointer DGEMM through dgemm_ (A, B, C in host memory, staged by the
    library) and the same call on device operands, against the offload model's
    predictions, at square edges."""
    from zynum_amd import _lib, autotune

    rows = {}
    for e in OFFLOAD_EDGES:
        host = autotune.blas_call_ns(torch, lib, _lib.F64, e, e, e, True, 0.0, 5)
        devn = autotune.blas_call_ns(torch, lib, _lib.F64, e, e, e, False, 0.0, 5)
        ph = lib.zynum_gpu_offload_predict_ns(_lib.F64, e, e, e, 7, 0)
        pd = lib.zynum_gpu_offload_predict_ns(_lib.F64, e, e, e, 0, 0)
        rows[str(e)] = {"host_us": round(host / 1e3, 1), "device_us": round(devn / 1e3, 1),
                        "predicted_host_us": round(ph / 1e3, 1), "predicted_device_us": round(pd / 1e3, 1)}
        log(f"offload dgemm {e}^3: host {host / 1e3:.1f} us (model {ph / 1e3:.1f}), "
            f"device {devn / 1e3:.1f} us (model {pd / 1e3:.1f})")
    ratio = [r["predicted_host_us"] / r["host_us"] for r in rows.values()]
    return {"workload": "DGEMM NN m=n=k=edge, alpha 1, beta 0, best of 5 synchronous dgemm_ calls; host: "
                        "A, B, C in pageable host memory", "by_edge": rows,
            "model_over_measured_host": [round(min(ratio), 3), round(max(ratio), 3)]}


def offload_crossover(lib, off, cpu_by_edge):
    """Where the measured host-pointer GPU call beats the CPU path, and whether
    zynum_gpu_offload_choose, given the CPU rate at that edge, makes the same
    choice."""
    from zynum_amd import _lib

    agree, measured_edge = 0, None
    for e in OFFLOAD_EDGES:
        row = off["by_edge"][str(e)]
        cpu_us = cpu_by_edge[str(e)]["us"]
        row["cpu_us"] = cpu_us
        gpu_wins = row["host_us"] < cpu_us
        rate = 2.0 * e ** 3 / (cpu_us * 1e3)  # GFLOP/s = FLOP per ns
        row["choose"] = int(lib.zynum_gpu_offload_choose(_lib.F64, e, e, e, 7, 0, rate))
        agree += row["choose"] == int(gpu_wins)
        if gpu_wins and measured_edge is None:
            measured_edge = e
        elif not gpu_wins:
            measured_edge = None
    off["measured_crossover_edge"] = measured_edge
    off["choice_agrees_with_measurement"] = f"{agree}/{len(OFFLOAD_EDGES)}"
    log(f"offload crossover: GPU (host operands) wins from edge {measured_edge}; model choice agrees at "
        f"{agree}/{len(OFFLOAD_EDGES)} edges")


def cpu_leg(n, ld, target_s, edges=(), with_configs=True):
    """The oracle (restated reference CPU path: packed-B 16x8 FMA tiles, column
    panels over min(cores, 64) threads) on a column block of the same workload,
    and at the square edges of the offload leg."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle
    from zynum_amd import synth

    oracle.build()
    threads = oracle.threads()
    log(f"cpu baseline: {threads} threads, isa tier {oracle.isa()}")
    a = synth.matrix(0, 0, n, n, ld)
    b = synth.matrix(0, n * n, n, n, ld)
    cols = max(64, 8 * threads)
    while True:
        c = np.zeros(ld * cols)
        t0 = oracle.now()
        oracle.dgemm("N", "N", n, cols, n, 1.0, a, ld, b, ld, 0.0, c, ld)
        dt = oracle.now() - t0
        if dt >= target_s * 0.5 or cols >= n:
            break
        cols = min(n, max(cols * 2, int(cols * target_s / max(dt, 1e-3) / 8) * 8))
    rate = 2.0 * n * cols * n / dt / 1e12
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    used = min(threads, (cols + 7) // 8)
    by_edge = {}
    for e in edges:
        # the reference planner runs work below 192^3 on one thread (planner.zig:13,135-144)
        oracle.set_threads(1 if e ** 3 < 192 ** 3 else 0)
        ae, be = synth.matrix(1, 0, e, e, e), synth.matrix(1, e * e, e, e, e)
        best = math.inf
        for _ in range(3 if e <= 1024 else 1):
            ce = np.zeros(e * e)
            t0 = oracle.now()
            oracle.dgemm("N", "N", e, e, e, 1.0, ae, e, be, e, 0.0, ce, e)
            best = min(best, oracle.now() - t0)
        by_edge[str(e)] = {"us": round(best * 1e6, 1), "gflops": round(2.0 * e ** 3 / best / 1e9, 2),
                           "threads": oracle.threads()}
    oracle.set_threads(0)
    if by_edge:
        log(f"cpu baseline by edge: {by_edge}")
    configs = cpu_configs(oracle, synth) if with_configs else {}
    return {"value": round(rate, 5), "unit": "TFLOP/s", "cores": used, "kind": "port", "configs": configs,
            "sample": f"DGEMM M={n} N={cols} K={n} (a {cols}-column block of the 8192^3 NN workload, ld={ld}), "
                      f"one call, {dt:.2f}s; restatement of Zynum's portable CPU path (reference not executable: "
                      f"no Zig toolchain); host '{cpu_model}', {os.cpu_count()} logical CPUs",
            "by_edge": by_edge}


def cpu_configs(oracle, synth):
    """The restated CPU path on the other benchmark configurations: the
    1024^3 cblas_dgemm (alpha 1.5, beta 0.5; best and median of 10 calls after
    a warm-up), the loop over the 4096 uneven problems (one thread below 192^3
    as the reference planner, else all), and DTRSM / DSYRK at N = 2048 with
    the time extrapolated to N = 8192 by (8192 / 2048)^3 (the reference's
    column-substitution TRSM and dot-product SYRK are cubic)."""
    from zynum_amd import grouped

    out = {}
    n = 1024
    a, b, c0 = synth.matrix(0, 0, n, n, n), synth.matrix(0, n * n, n, n, n), synth.matrix(0, 2 * n * n, n, n, n)
    oracle.set_threads(0)
    times = []
    for r in range(11):
        c = c0.copy()
        t0 = oracle.now()
        oracle.dgemm("N", "N", n, n, n, 1.5, a, n, b, n, 0.5, c, n)
        if r:
            times.append(oracle.now() - t0)
    times.sort()
    out["cblas_dgemm_1024"] = {"best_us": round(times[0] * 1e6, 1), "median_us": round(times[len(times) // 2] * 1e6, 1),
                               "best_gflops": round(2.0 * n ** 3 / times[0] / 1e9, 1),
                               "median_gflops": round(2.0 * n ** 3 / times[len(times) // 2] / 1e9, 1),
                               "threads": oracle.threads()}
    del a, b, c0, c
    shapes = grouped.draw_shapes(4096, 0, 16, 512)
    rng = np.random.default_rng(0)
    ops = [(rng.uniform(-1, 1, int(m * k)), rng.uniform(-1, 1, int(k * nn)), np.zeros(int(m * nn)))
           for m, nn, k in shapes]
    flop = float(np.sum(2.0 * shapes[:, 0] * shapes[:, 1] * shapes[:, 2]))
    t0 = oracle.now()
    for (m, nn, k), (x, y, z) in zip(shapes, ops):
        oracle.set_threads(1 if m * nn * k < 192 ** 3 else 0)
        oracle.dgemm("N", "N", int(m), int(nn), int(k), 1.0, x, int(m), y, int(k), 0.0, z, int(m))
    dt = oracle.now() - t0
    oracle.set_threads(0)
    out["grouped_4096_loop"] = {"ms": round(dt * 1e3, 1), "tflops": round(flop / dt / 1e12, 4),
                                "workload": "4096 DGEMMs NN, M,N,K log-uniform in [16,512] (seed 0), one call each"}
    del ops
    n = 2048
    L = synth.matrix(11, 0, n, n, n).reshape(n, n)
    L = np.triu(L)  # column-major lower triangle
    L[np.diag_indices(n)] = 8.0 + np.abs(synth.uniform(11, n * n, n))
    X = synth.matrix(11, n * n + n, n, n, n)
    t0 = oracle.now()
    oracle.dtrsm("L", "L", "N", "N", n, n, 1.0, L.ravel(), n, X, n)
    dt = oracle.now() - t0
    out["dtrsm_2048"] = {"ms": round(dt * 1e3, 1), "tflops": round(float(n) ** 3 / dt / 1e12, 4),
                         "extrapolated_8192_ms": round(dt * 64 * 1e3, 1),
                         "note": "DTRSM L/L/N/N M=N=2048; 8192 time extrapolated by (8192/2048)^3, not measured"}
    k = n // 2
    A = synth.matrix(12, 0, n, k, n)
    C = np.zeros(n * n)
    t0 = oracle.now()
    oracle.dsyrk("L", "N", n, k, 1.0, A, n, 0.0, C, n)
    dt = oracle.now() - t0
    out["dsyrk_2048x1024"] = {"ms": round(dt * 1e3, 1), "tflops": round(float(n) * n * k / dt / 1e12, 4),
                              "extrapolated_8192x4096_ms": round(dt * 64 * 1e3, 1),
                              "note": "DSYRK Lower NoTrans N=2048 K=1024; 8192x4096 time extrapolated by 4^3, "
                                      "not measured"}
    log(f"cpu baseline configs: {out}")
    return out


if __name__ == "__main__":
    sys.exit(main())
