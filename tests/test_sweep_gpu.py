"""The GEMM sweep harness (bench/gemm_sweep.zig) run on the GPU tier.

Bars: the CSV header is the reference's (gemm_sweep.zig:505); every row has
its 16 fields with the shape, kind and transpose pair asked for; every row's
sampled check (all entries for M N <= 4096, else the 5 x 5 grid,
gemm_sweep.zig:247-290, tolerance 1e-9 K, 1e-3 K single) reads
"checked-ok"; the statistics are ordered best <= median <= p95 <= max and
GFLOP/s equals 2MNK (8MNK complex) / best.  Device operands and host
operands (staged by the library) both run.
"""
import csv
import io

import pytest

pytestmark = pytest.mark.gpu

REFERENCE_HEADER = ("kind,transa,transb,shape_index,label,m,n,k,library,gflops,best_ns,median_ns,p95_ns,max_ns,"
                    "reps,check")
SHAPES = ["m1_n1_k1:1:1:1", "m63_n65_k17:63:65:17", "sq192:192:192:192", "m2048_n17_k257:2048:17:257",
          "m128_n128_k4096:128:128:4096"]


@pytest.mark.parametrize("host", [False, True])
def test_sweep_check_rows(torch_cuda, tmp_path, host):
    from zynum_amd import sweep

    out = tmp_path / "sweep.csv"
    argv = ["--csv", str(out), "--reps", "3", "--check", "--trans", "NN", "--trans", "TN", "--trans", "CT"]
    for s in SHAPES:
        argv += ["--shape", s]
    if host:
        argv.append("--host")
    assert sweep.main(argv) == 0
    text = out.read_text()
    assert text.splitlines()[0] == REFERENCE_HEADER
    rows = list(csv.reader(io.StringIO(text)))[1:]
    # real kinds skip conjugate pairs (gemm_sweep.zig:205-208): 2 pairs x 4 kinds + 1 pair x 2 kinds per shape
    assert len(rows) == len(SHAPES) * (2 * 4 + 2)
    seen = set()
    for r in rows:
        assert len(r) == 16
        kind, ta, tb, idx, label, m, n, k = r[:8]
        assert kind in sweep.KINDS and r[8] == "Zynum-MI355X"
        lab, sm, sn, sk = SHAPES[int(idx)].split(":")
        assert (label, m, n, k) == (lab, sm, sn, sk)
        best, med, p95, mx, reps = map(int, r[10:15])
        assert 0 < best <= med <= p95 <= mx and reps == 3
        flop = sweep.KINDS[kind][2] * int(m) * int(n) * int(k)
        assert float(r[9]) == pytest.approx(flop / best, rel=1e-5, abs=1e-6)
        assert r[15] == "checked-ok", r
        seen.add((kind, ta + tb))
    assert ("dgemm", "TN") in seen and ("zgemm", "CT") in seen and ("sgemm", "CT") not in seen
