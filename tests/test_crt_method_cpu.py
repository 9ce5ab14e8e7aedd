"""The CRT reconstruction used by the Ozaki-II kernel, emulated on the CPU.

The device kernel rebuilds C' from residues u_l as digits D_d = S_d - q P_d of
40 bits (each an exact integer below 2^53), normalizes them to balanced digits
with exact carries, and sums from the lowest digit.  This test emulates that
arithmetic (exact integers where the kernel's steps are exact, IEEE doubles
where it rounds) and checks the property the kernel relies on: the result is
exactly C' whenever C' is a double, and within one rounding otherwise.  It also
shows that summing the raw digit differences without carry normalization
(a plausible mistake) loses exactness.
"""
import math
import random

MODULI = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193, 191, 181, 179, 173]


def setup(nmod):
    P = 1
    for m in MODULI[:nmod]:
        P *= m
    nd = (P.bit_length() + 8 + 39) // 40
    e = [40 * (nd - 1 - d) for d in range(nd)]
    dig = lambda x, lo: (x >> lo) & ((1 << 40) - 1)  # noqa: E731
    h = [[dig(P // m, ed) for ed in e] for m in MODULI[:nmod]]
    y = [pow((P // m) % m, -1, m) for m in MODULI[:nmod]]
    p = [dig(P, ed) for ed in e]
    papprox = sum(math.ldexp(float(p[d]), e[d]) for d in range(nd))
    inv = [math.ldexp(1.0, e[d]) / papprox for d in range(nd)]
    return P, nd, e, h, y, p, inv


def reconstruct(C, nmod, cfg, normalize=True):
    P, nd, e, h, y, p, inv = cfg
    u = [(y[l] * (C % MODULI[l])) % MODULI[l] for l in range(nmod)]
    S = [sum(h[l][d] * u[l] for l in range(nmod)) for d in range(nd)]
    assert all(s < 2**53 for s in S)
    xq = 0.0
    for d in range(nd):
        xq = float(S[d]) * inv[d] + xq
    q = round(xq)
    D = [S[d] - q * p[d] for d in range(nd)]
    assert all(abs(x) < 2**53 for x in D)
    if normalize:
        for d in range(nd - 1, 0, -1):
            c = round(D[d] * 2.0**-40)
            D[d] -= c * 2**40
            D[d - 1] += c
    cp = 0.0
    for d in range(nd - 1, -1, -1):
        cp = math.ldexp(float(D[d]), e[d]) + cp
    return cp


def representable(rng, P):
    sh = rng.randint(0, max(0, P.bit_length() - 60))
    C = rng.randint(-2**52, 2**52) << sh
    while abs(C) >= P // 2:
        C //= 4
    return C


def test_reconstruction_exact_when_representable():
    rng = random.Random(1)
    for nmod in (8, 14, 16, 20):
        cfg = setup(nmod)
        for _ in range(3000):
            C = representable(rng, cfg[0])
            assert reconstruct(C, nmod, cfg) == float(C)


def test_reconstruction_one_rounding_otherwise():
    rng = random.Random(2)
    for nmod in (10, 14, 18):
        cfg = setup(nmod)
        P = cfg[0]
        for _ in range(3000):
            C = rng.randint(-(P // 2) + 1, P // 2 - 1)
            assert abs(reconstruct(C, nmod, cfg) - C) <= abs(C) * 2.0**-52


def test_unnormalized_digits_lose_exactness():
    rng = random.Random(3)
    cfg = setup(14)
    bad = sum(reconstruct(C, 14, cfg, normalize=False) != float(C)
              for C in (representable(rng, cfg[0]) for _ in range(2000)))
    assert bad > 0


def fma_exact(a, b, c):
    """fma(a, b, c) of doubles: the exact a*b + c rounded once (Fraction -> float rounds to nearest even)."""
    from fractions import Fraction

    return float(Fraction(a) * Fraction(b) + Fraction(c))


def sym_mod(v, m, two_step=True):
    """The residue extraction of the plane kernel, in IEEE doubles: one exact
    reduction step, a second one when |r| > m, then the symmetric fix-up."""
    inv_m = 1.0 / m
    r = fma_exact(-float(round(v * inv_m)), float(m), v)
    if two_step and abs(r) > m:
        r = fma_exact(-float(round(r * inv_m)), float(m), r)
    ri = int(r)
    half = m >> 1
    if ri >= half + (m & 1):
        ri -= m
    if ri < -half:
        ri += m
    return ri


def exact_sym_mod(v, m):
    r = int(v) % m
    half = m >> 1
    return r - m if r >= half + (m & 1) else r


def test_residue_extraction_exact_for_every_modulus_count():
    """Scaled inputs are integers with |a'| <= 2^(alpha+1); for every modulus
    count the kernel's reduction returns the exact symmetric residue (the
    largest alpha, 20 moduli, needs the second step)."""
    rng = random.Random(11)
    for nmod in (8, 14, 16, 20):
        P = 1
        for m in MODULI[:nmod]:
            P *= m
        alpha = (P.bit_length() - 2) // 2
        for _ in range(400):
            e = rng.randint(0, alpha + 1)
            v = float(rng.randint(-2**52, 2**52)) * 2.0 ** max(0, e - 52)
            v = float(round(v))
            for m in MODULI[:nmod]:
                assert sym_mod(v, m) == exact_sym_mod(v, m), (nmod, v, m)


def test_single_reduction_step_is_not_enough_at_20_moduli():
    """Planted mistake: dropping the conditional second step gives wrong
    residues for the largest scaled inputs (|a'| near 2^79)."""
    rng = random.Random(12)
    bad = 0
    for _ in range(200):
        v = float(rng.randint(2**52, 2**53 - 1)) * 2.0**26
        for m in MODULI[:20]:
            if sym_mod(v, m, two_step=False) != exact_sym_mod(v, m):
                bad += 1
    assert bad > 0


def balanced(v, m):
    r = v % m
    return r - m if r > (m - 1) // 2 else r


def test_two_step_odd_residue_is_exact():
    """The plane kernel's residue of an integer-valued double v for an odd
    modulus m: t = v - m rint(v inv_m), then r = t - m rint(t inv_m), both
    differences exact (fma of integers with an integer result).  Emulated with
    IEEE doubles for the products and roundings and exact integers for the
    differences, r is the balanced residue in [-(m-1)/2, (m-1)/2] for every v up
    to 2^100 in magnitude, including values next to multiples of m and to
    half-multiples, where a one-step rounding of v / m can go either way."""
    rnd = random.Random(17)
    cases = []
    for m in MODULI[1:]:  # the odd moduli
        inv = 1.0 / m
        for bits in (8, 30, 53, 56, 78, 100):
            for _ in range(400):
                v = rnd.randrange(-(1 << bits), 1 << bits)
                if bits > 53:
                    v = int(float(v))  # an integer-valued double
                cases.append((v, m, inv))
            for k in (1, 3, 12345, (1 << 40) + 7):
                for d in (-1, 0, 1):
                    cases.append((k * m + d, m, inv))
                    cases.append((-(k * m) + d, m, inv))
                    cases.append((k * m + (m - 1) // 2 + d, m, inv))
    for v, m, inv in cases:
        t = v - m * int(round(float(v) * inv))  # rint = round half to even, as Python's round on floats
        r = t - m * int(round(float(t) * inv))
        assert r == balanced(v, m), (v, m, t, r)


def test_one_step_residue_needs_the_second_step():
    """A plausible shortcut, one rounded quotient, leaves |r| above (m-1)/2 for large v."""
    rnd = random.Random(3)
    m, inv = 173, 1.0 / 173
    bad = 0
    for _ in range(2000):
        v = int(float(rnd.randrange(-(1 << 78), 1 << 78)))
        r = v - m * int(round(float(v) * inv))
        bad += abs(r) > (m - 1) // 2
    assert bad > 0
