"""Writes the golden fixtures under tests/golden/ (run from the repo root).

config1_dd_samples.json: the 1024^3 NN DGEMM workload (alpha 1.5, beta 0.5,
A, B, C drawn in that order from splitmix64 seed 0, see zynum_amd/synth.py) at
512 seeded (row, col) entries: the double-double value (hi, lo) of
alpha*A*B + beta*C, computed with exact TwoProd/TwoSum arithmetic, and the
oracle's FP64 value (the restated reference CPU path).
"""
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

import oracle  # noqa: E402
from zynum_amd import synth  # noqa: E402


def config1():
    n = 1024
    a = synth.matrix(0, 0, n, n, n)
    b = synth.matrix(0, n * n, n, n, n)
    c = synth.matrix(0, 2 * n * n, n, n, n)
    rs = np.random.default_rng(2024)
    rows = rs.integers(0, n, 512)
    cols = rs.integers(0, n, 512)
    hi, lo = oracle.dd_gemm_entries("N", "N", n, 1.5, a, n, b, n, 0.5, c, n, rows, cols)
    ref = c.copy()
    oracle.dgemm("N", "N", n, n, n, 1.5, a, n, b, n, 0.5, ref, n)
    return {
        "workload": "cblas_dgemm ColMajor NN m=n=k=1024 lda=ldb=ldc=1024 alpha=1.5 beta=0.5, splitmix64 seed 0",
        "rows": rows.tolist(), "cols": cols.tolist(),
        "hi": hi.tolist(), "lo": lo.tolist(),
        "oracle_fp64": ref[rows + cols * n].tolist(),
    }


if __name__ == "__main__":
    out = ROOT / "tests" / "golden" / "config1_dd_samples.json"
    out.write_text(json.dumps(config1()))
    print("wrote", out)
