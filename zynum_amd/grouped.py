"""Host helper for grouped GEMM batches (device buffers + problem table).

Wraps ``zynum_gpu_grouped_plan_*`` / ``zynum_gpu_grouped_loop`` from
``include/zynum_gpu.h``.  Shapes are drawn log-uniformly from a seed; entries
are uniform on [-1, 1) from a seeded torch generator on the device.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib, synth


def draw_shapes(count: int, seed: int, lo: int = 16, hi: int = 512) -> np.ndarray:
    """(count, 3) integer M, N, K = round(exp(U[ln lo, ln hi])), clipped to [lo, hi]."""
    u = (synth.uniform(seed, 0, 3 * count) + 1.0) * 0.5
    v = np.rint(np.exp(math.log(lo) + u * (math.log(hi) - math.log(lo))))
    return np.clip(v, lo, hi).astype(np.int64).reshape(count, 3)


class Batch:
    """Operands of ``len(shapes)`` column-major problems carved from three flat
    device tensors (A, B, C), with ld = rows (+ ``pad``)."""

    def __init__(self, torch, shapes, dtype, seed=0, transa=False, transb=False, alpha=1.0, beta=0.0, pad=0,
                 device="cuda"):
        self.torch = torch
        self.shapes = np.asarray(shapes, dtype=np.int64)
        self.dtype = dtype
        self.scalar = _lib.F64 if dtype == torch.float64 else _lib.F32
        self.es = 8 if dtype == torch.float64 else 4
        n = len(self.shapes)
        self.problems = (_lib.GemmProblem * max(1, n))()
        offs_a, offs_b, offs_c = [], [], []
        ta = tb = tc = 0
        lds = []
        for (m, nn, k) in self.shapes:
            ra, ca = (k, m) if transa else (m, k)
            rb, cb = (nn, k) if transb else (k, nn)
            lda, ldb, ldc = max(1, ra) + pad, max(1, rb) + pad, max(1, m) + pad
            offs_a.append(ta); offs_b.append(tb); offs_c.append(tc)
            ta += int(lda * ca); tb += int(ldb * cb); tc += int(ldc * nn)
            lds.append((int(lda), int(ldb), int(ldc)))
        g = torch.Generator(device=device).manual_seed(seed)
        self.A = torch.rand(max(1, ta), generator=g, device=device, dtype=dtype) * 2 - 1
        self.B = torch.rand(max(1, tb), generator=g, device=device, dtype=dtype) * 2 - 1
        self.C = torch.rand(max(1, tc), generator=g, device=device, dtype=dtype) * 2 - 1
        self.C0 = self.C.clone()
        self.offs = (offs_a, offs_b, offs_c)
        self.lds = lds
        for i, (m, nn, k) in enumerate(self.shapes):
            p = self.problems[i]
            p.transa = _lib.TRANS if transa else _lib.NO_TRANS
            p.transb = _lib.TRANS if transb else _lib.NO_TRANS
            p.m, p.n, p.k = int(m), int(nn), int(k)
            p.lda, p.ldb, p.ldc = (int(x) for x in lds[i])
            p.alpha, p.beta = alpha, beta
            p.a = int(self.A.data_ptr() + int(offs_a[i]) * self.es)
            p.b = int(self.B.data_ptr() + int(offs_b[i]) * self.es)
            p.c = int(self.C.data_ptr() + int(offs_c[i]) * self.es)
        self.flops = float(np.sum(2.0 * self.shapes[:, 0] * self.shapes[:, 1] * self.shapes[:, 2]))
        self.bytes = float(np.sum(self.es * (self.shapes[:, 0] * self.shapes[:, 2] + self.shapes[:, 2] * self.shapes[:, 1]
                                             + self.shapes[:, 0] * self.shapes[:, 1] * (2 if beta != 0 else 1))))

    def reset_c(self):
        self.C.copy_(self.C0)

    def plan(self, tile: int = 0) -> "Plan":
        return Plan(self, tile)

    def loop(self, stream=None) -> None:
        st = _lib.load().zynum_gpu_grouped_loop(self.scalar, self.problems, len(self.shapes), stream)
        _lib.check(st, "grouped loop")

    def result(self, i: int):
        """Problem i's C as a (m, n) host array."""
        m, nn, _ = self.shapes[i]
        ldc = self.lds[i][2]
        off = self.offs[2][i]
        flat = self.C[off: off + ldc * nn].cpu().numpy()
        return flat.reshape(nn, ldc)[:, :m].T


class Plan:
    def __init__(self, batch: Batch, tile: int = 0):
        self.lib = _lib.load()
        self.handle = C.c_void_p()
        st = self.lib.zynum_gpu_grouped_plan_create(batch.scalar, batch.problems, len(batch.shapes), tile,
                                                    C.byref(self.handle))
        _lib.check(st, "grouped plan")

    def run(self, stream=None) -> None:
        _lib.check(self.lib.zynum_gpu_grouped_run(self.handle, stream), "grouped run")

    def graph_run(self, stream=None) -> None:
        _lib.check(self.lib.zynum_gpu_grouped_graph_run(self.handle, stream), "grouped graph run")

    def info(self):
        t, w, f = C.c_int64(), C.c_int64(), C.c_double()
        _lib.check(self.lib.zynum_gpu_grouped_plan_info(self.handle, C.byref(t), C.byref(w), C.byref(f)))
        return {"tiles": t.value, "workgroups": w.value, "flops": f.value}

    def close(self) -> None:
        if self.handle:
            self.lib.zynum_gpu_grouped_plan_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def main(argv=None) -> int:
    """Runs the benchmark batch (4096 problems, M, N, K log-uniform in
    [16, 512], seed 0) through one plan without graph capture, ``--reps``
    times, and prints the last run's time: a command for kernel traces and
    counter collection of the path kernels one launch at a time."""
    import argparse

    import torch

    ap = argparse.ArgumentParser(prog="python -m zynum_amd.grouped")
    ap.add_argument("--dtype", choices=("f64", "f32"), default="f64")
    ap.add_argument("--reps", type=int, default=4)
    args = ap.parse_args(argv)
    dt = torch.float64 if args.dtype == "f64" else torch.float32
    stream = torch.cuda.Stream()
    batch = Batch(torch, draw_shapes(4096, 0, 16, 512), dt, seed=0)
    plan = batch.plan()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for i in range(args.reps):
        if i == args.reps - 1:
            e0.record(stream)
        plan.run(stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    print(f"grouped {args.dtype}: {plan.info()}, last run {e0.elapsed_time(e1):.3f} ms", flush=True)
    plan.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
