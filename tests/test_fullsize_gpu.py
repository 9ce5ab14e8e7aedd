"""Parity at the full benchmark sizes of the split-precision, grouped and
Level-3 forms.

* Split precision, DGEMM 8192^3, ld 8200, wide-range inputs a = u 2^e
  (u uniform [-1, 1), e uniform in [-16, 16], seeds as in the benchmark):
  the 14-residue CRT form (Ozaki II) within its error bound
  4 sqrt(K) 2^-t ||a_i|| ||b_j|| + 2^-52 |c| against the oracle's
  double-double value on 2048 sampled entries; the CRT-14 and the 8-slice
  INT8 forms no worse (max error on the samples) than plain FP64 MFMA on the
  same samples; the 8-slice BF16 form bitwise equal to the 8-slice INT8 form
  over the whole matrix (every slice product is an integer below 2^12 and
  the FP32 partial sums stay exact).
* Grouped GEMM, the 4096-problem batch of the benchmark (M, N, K
  log-uniform in [16, 512], seed 0), FP64 and FP32: the graph replay of the
  persistent launch bitwise equals the loop of single plain GEMM launches,
  and 64 sampled problems lie within 2 (K + 2) u |A||B| of the oracle.
* DTRSM Left/Lower/NoTrans/NonUnit 8192 x 8192 on the benchmark's L
  (diagonal 8 + |u|): the residual L X - B on 2048 sampled entries, formed in
  double-double by the oracle, within 8 n u (|L||X|); a unit-diagonal
  integer triangle with integer X recovered bitwise (plain and CRT-14
  split-precision updates).
* DSYRK Lower, N = 8192, K = 4096, beta 0 and 1, integer inputs: the stored
  triangle bitwise equal to exact arithmetic, the other triangle and the
  padding untouched.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 8192
LD = 8200
MODULI = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199]


def say(msg):
    print(f"[fullsize] {msg}", flush=True)


def request(torch, a, b, c, mode, param, n=N, ld=LD):
    from zynum_amd import _lib, blas

    st = blas.gemm_request(False, False, n, n, n, 1.0, a.data_ptr(), ld, b.data_ptr(), ld, 0.0, c.data_ptr(), ld,
                           _lib.F64, mode=mode, mode_param=param)
    _lib.check(st, f"mode {mode}/{param}")
    torch.cuda.synchronize()


def crt_scale_bits(nmod, k):
    """Largest t with 2 log2(2^t + sqrt(K)/2) <= log2(P) - 1 (the Cauchy-Schwarz scaling of the CRT form)."""
    log2p = sum(math.log2(x) for x in MODULI[:nmod])
    t = math.floor((log2p - 1) / 2)
    while 2 * math.log2(2.0**t + 0.5 * math.sqrt(k)) > log2p - 1 - 1e-6:
        t -= 1
    return t


def test_split_precision_8192_wide_range(torch_cuda, oracle):
    from zynum_amd import synth

    torch = torch_cuda
    say("split: generating 8192^2 wide-range A and B (ld 8200)")
    a = synth.wide_range(7, 0, N, N, LD)
    b = synth.wide_range(7, N * N, N, N, LD)
    A = torch.from_numpy(a).cuda()
    B = torch.from_numpy(b).cuda()
    outs = {}
    for name, mode, param in (("plain", 0, 0), ("crt14", 3, 14), ("int8x8", 1, 8), ("bf16x8", 2, 8)):
        say(f"split: {name}")
        c = torch.full((LD * N,), -999.0, dtype=torch.float64, device="cuda")
        request(torch, A, B, c, mode, param)
        outs[name] = c
    assert torch.equal(outs["bf16x8"], outs["int8x8"]), "BF16 slices differ from INT8 slices"
    del outs["bf16x8"]
    pad = outs["crt14"].view(N, LD)[:, N:]
    assert bool((pad == -999.0).all()), "padding rows written"

    rng = np.random.default_rng(8192)
    rows = rng.integers(0, N, 2048)
    cols = rng.integers(0, N, 2048)
    say("split: double-double oracle on 2048 sampled entries")
    hi, lo = oracle.dd_gemm_entries("N", "N", N, 1.0, a, LD, b, LD, 0.0, None, LD, rows, cols)
    idx = torch.from_numpy(rows + cols * LD).cuda()
    err = {k: np.abs((v[idx].cpu().numpy() - hi) - lo) for k, v in outs.items()}
    say("split: max err on samples " + ", ".join(f"{k} {v.max():.3e}" for k, v in err.items()))

    Av = A.view(N, LD)[:, :N]  # Av[p, i] = A(i, p)
    Bv = B.view(N, LD)[:, :N]  # Bv[j, p] = B(p, j)
    na = torch.linalg.vector_norm(Av, dim=0).cpu().numpy()  # row norms of A
    nb = torch.linalg.vector_norm(Bv, dim=1).cpu().numpy()  # column norms of B
    t = crt_scale_bits(14, N)
    bound = 4 * math.sqrt(N) * 2.0**-t * na[rows] * nb[cols] + 2.0**-52 * np.abs(hi)
    assert np.all(err["crt14"] <= bound)
    assert err["crt14"].max() <= err["plain"].max()
    assert err["int8x8"].max() <= err["plain"].max()


@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
def test_grouped_4096_batch(torch_cuda, oracle, dtype_name):
    from zynum_amd import grouped

    torch = torch_cuda
    dt = getattr(torch, dtype_name)
    shapes = grouped.draw_shapes(4096, 0, 16, 512)
    b = grouped.Batch(torch, shapes, dt, seed=0)
    say(f"grouped {dtype_name}: loop of 4096 single launches")
    b.loop()
    torch.cuda.synchronize()
    want = b.C.clone()
    plan = b.plan()
    info = plan.info()
    assert info["tiles"] > 20000
    for rep in range(2):
        b.C.fill_(float("nan"))
        torch.cuda.synchronize()
        plan.graph_run()
        torch.cuda.synchronize()
        assert torch.equal(b.C, want), f"graph replay {rep} differs from the loop"
    plan.close()

    u = 2.0**-53 if dt == torch.float64 else 2.0**-24
    gemm = oracle.dgemm if dt == torch.float64 else oracle.sgemm
    A, B = b.A.cpu().numpy(), b.B.cpu().numpy()
    got = b.C.cpu().numpy()
    rng = np.random.default_rng(4096)
    for i in rng.choice(len(shapes), 64, replace=False):
        m, n, k = (int(x) for x in shapes[i])
        lda, ldb, ldc = b.lds[i]
        oa, ob, oc = b.offs[0][i], b.offs[1][i], b.offs[2][i]
        c = np.zeros(ldc * n, dtype=A.dtype)
        gemm("N", "N", m, n, k, 1.0, A[oa:], lda, B[ob:], ldb, 0.0, c, ldc)
        Am = A[oa: oa + lda * k].reshape(k, lda)[:, :m].T.astype(np.float64)
        Bm = B[ob: ob + ldb * n].reshape(n, ldb)[:, :k].T.astype(np.float64)
        bound = 2 * (k + 2) * u * (np.abs(Am) @ np.abs(Bm))
        g = got[oc: oc + ldc * n].reshape(n, ldc)[:, :m].T.astype(np.float64)
        w = c.reshape(n, ldc)[:, :m].T.astype(np.float64)
        assert np.all(np.abs(g - w) <= bound), f"problem {i} ({m}x{n}x{k})"


def trsm_dev(torch, m, n, L, ldl, X, ldx, mode=0, param=0):
    from zynum_amd import _lib

    st = _lib.load().zynum_gpu_dtrsm(b"L", b"L", b"N", b"N", m, n, 1.0, L.data_ptr(), ldl, X.data_ptr(), ldx, None,
                                     mode, param, 0)
    _lib.check(st, "dtrsm")
    torch.cuda.synchronize()


def test_dtrsm_8192_residual_vs_oracle(torch_cuda, oracle):
    from zynum_amd import synth

    torch = torch_cuda
    say("dtrsm: generating L (diagonal 8 + |u|) and B")
    L = synth.matrix(11, 0, N, N, N).reshape(N, N)  # L[j, i] = element (i, j)
    L = np.triu(L)  # element (i, j) kept for i >= j: the lower triangle
    L[np.diag_indices(N)] = 8.0 + np.abs(synth.uniform(11, N * N, N))
    Bh = synth.matrix(11, N * N + N, N, N, N)
    Ld = torch.from_numpy(L.ravel()).cuda()
    Lpoison = Ld.view(N, N).clone()
    Lpoison[torch.tril(torch.ones(N, N, dtype=torch.bool, device="cuda"), -1)] = float("nan")  # upper: never read
    X = torch.from_numpy(Bh.copy()).cuda()
    say("dtrsm: solve")
    trsm_dev(torch, N, N, Lpoison.view(-1), N, X, N)
    del Lpoison
    Xh = X.cpu().numpy()
    assert np.all(np.isfinite(Xh))
    rng = np.random.default_rng(55)
    rows = rng.integers(0, N, 2048)
    cols = rng.integers(0, N, 2048)
    say("dtrsm: double-double residual on 2048 sampled entries")
    flat = L.ravel()
    hi, lo = oracle.dd_gemm_entries("N", "N", N, 1.0, flat, N, Xh, N, -1.0, Bh, N, rows, cols)
    res = np.abs(hi + lo)
    Lrows = np.abs(L[:, rows]).T  # |L(i, :)| for each sampled row
    Xcols = np.abs(Xh.reshape(N, N)[cols, :])  # |X(:, j)| for each sampled column
    scale = np.einsum("ek,ek->e", Lrows, Xcols)
    say(f"dtrsm: max residual {res.max():.3e}, max residual / (n u |L||X|) "
        f"{(res / (N * 2.0**-53 * scale)).max():.3e}")
    assert np.all(res <= 8 * N * 2.0**-53 * scale)


@pytest.mark.parametrize("mode,param", [(0, 0), (3, 14)])
def test_dtrsm_8192_integer_exact(torch_cuda, mode, param):
    torch = torch_cuda
    g = torch.Generator(device="cuda").manual_seed(8192 + mode)
    T = torch.randint(-1, 2, (N, N), generator=g, device="cuda", dtype=torch.float64)
    T = torch.tril(T, -1) + torch.eye(N, dtype=torch.float64, device="cuda")  # T[i, j] = element (i, j)
    X = torch.randint(-3, 4, (N, N), generator=g, device="cuda", dtype=torch.float64)
    Bm = T @ X  # integer partial sums below 2^53: exact in any order
    Lcm = T.t().contiguous()  # column-major storage: Lcm[j, i] = element (i, j)
    Lcm[torch.triu(torch.ones(N, N, dtype=torch.bool, device="cuda"), 1).t()] = float("nan")
    Bcm = Bm.t().contiguous()
    say(f"dtrsm integer: solve mode {mode}/{param}")
    trsm_dev(torch, N, N, Lcm.view(-1), N, Bcm.view(-1), N, mode, param)
    assert torch.equal(Bcm, X.t().contiguous())


@pytest.mark.parametrize("beta", [0.0, 1.0])
def test_dsyrk_8192x4096_integer_exact(torch_cuda, beta):
    from zynum_amd import _lib

    torch = torch_cuda
    n, k, lda, ldc = N, 4096, N, N + 8
    g = torch.Generator(device="cuda").manual_seed(4096 + int(beta))
    A = torch.randint(-9, 10, (k, n), generator=g, device="cuda", dtype=torch.float64)  # A[p, i] = element (i, p)
    Cm = torch.randint(-9, 10, (n, ldc), generator=g, device="cuda", dtype=torch.float64)
    Cm[:, n:] = -777.0
    C0 = Cm.clone()
    say(f"dsyrk 8192x4096 beta {beta}")
    st = _lib.load().zynum_gpu_dsyrk(b"L", b"N", n, k, 1.0, A.data_ptr(), lda, beta, Cm.data_ptr(), ldc, None,
                                     0, 0, 0)
    _lib.check(st, "dsyrk")
    torch.cuda.synchronize()
    S = A.t() @ A  # S[i, j]: integer sums below 2^53, exact
    want = S + beta * C0[:, :n].t() if beta != 0 else S
    got = Cm[:, :n].t()  # got[i, j] = element (i, j)
    lower = torch.tril(torch.ones(n, n, dtype=torch.bool, device="cuda"))
    assert torch.equal(got[lower], want[lower])
    assert torch.equal(got[~lower], C0[:, :n].t()[~lower])
    assert bool((Cm[:, n:] == -777.0).all())
