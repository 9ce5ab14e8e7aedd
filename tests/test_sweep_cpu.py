"""CPU checks of the GEMM sweep tool: the harness's input sequence, the CSV
format of bench/gemm_sweep.zig and the sampled-entry check."""
import numpy as np

from zynum_amd import sweep


def test_lcg_closed_form_matches_the_recurrence():
    """gemm_sweep.zig:156-160 stepped one value at a time."""
    state = 0x3141592653589793
    want = []
    for _ in range(1000):
        state = (state * 6364136223846793005 + 1442695040888963407) % 2**64
        want.append(((state >> 32) % 1000) / 1000.0 - 0.5)
    assert np.array_equal(sweep.lcg_values(1000), np.array(want))


def test_complex_fill_takes_re_then_im():
    v = sweep.lcg_values(6)
    z = sweep.fill(3, np.complex128)
    assert np.array_equal(z.real, v[0::2]) and np.array_equal(z.imag, v[1::2])


def test_csv_header_and_row_format():
    """gemm_sweep.zig:505 (header) and :380-398 (quoted label, library and check; 6-decimal GFLOP/s)."""
    assert sweep.HEADER == ("kind,transa,transb,shape_index,label,m,n,k,library,gflops,best_ns,median_ns,p95_ns,"
                            "max_ns,reps,check\n")
    row = sweep.csv_row("zgemm", "C", "N", 3, ("m8_n8_k8", 8, 8, 8), "Zynum-MI355X", (1000, 1200, 1500, 1500,
                                                                                     "checked-ok"), 5)
    assert row == 'zgemm,C,N,3,"m8_n8_k8",8,8,8,"Zynum-MI355X",4.096000,1000,1200,1500,1500,5,"checked-ok"\n'


def test_default_shape_list():
    """gemm_sweep.zig:58-101: 42 shapes, first m1_n1_k1, last m640_n384_k96."""
    assert len(sweep.DEFAULT_SHAPES) == 42
    assert sweep.DEFAULT_SHAPES[0] == ("m1_n1_k1", 1, 1, 1)
    assert sweep.DEFAULT_SHAPES[-1] == ("m640_n384_k96", 640, 384, 96)
    assert all(lbl.count("_") in (0, 2) for lbl, *_ in sweep.DEFAULT_SHAPES)


def test_sampled_check_flags_a_wrong_entry():
    m, n, k = 100, 90, 7
    a, b = sweep.fill(m * k, np.float64), sweep.fill(k * n, np.float64)
    c = (a.reshape(k, m).T @ b.reshape(n, k).T).T.ravel().copy()
    assert sweep.check_samples(a, b, c, m, n, k, "N", "N", m, k, m, False) == "checked-ok"
    c[(n - 1) * m + (m - 1)] += 1e-3  # sample (m-1, n-1) is on the 5 x 5 grid
    assert sweep.check_samples(a, b, c, m, n, k, "N", "N", m, k, m, False) == "check-failed"
