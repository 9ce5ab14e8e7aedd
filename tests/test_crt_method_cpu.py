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


def exact_double(x):
    """An integer the kernel computes with one exact FMA: it must be a double."""
    assert float(x) == x, x
    return int(x)


def split_residue(v, m, parts):
    """The residue extraction of the plane kernel, in IEEE doubles: a = v is
    split once into a2 2^64 + a1 2^32 + a0 (truncated quotients, exact
    differences; a2 = 0 for two parts), then t = a2 (2^64 mod m) +
    a1 (2^32 mod m) + a0 and one rounded quotient for an odd m, a0 mod 256 as
    a signed byte for m = 256."""
    v = float(v)
    a2 = math.trunc(v * 2.0**-64) if parts == 3 else 0
    rest = exact_double(int(v) - a2 * 2**64)
    a1 = math.trunc(float(rest) * 2.0**-32)
    a0 = exact_double(rest - a1 * 2**32)
    if m == 256:
        u = exact_double(a0 - 256 * math.floor(float(a0) * 2.0**-8))
        assert 0 <= u < 256
        return u - 256 if u > 127 else u
    t = exact_double(a1 * (2**32 % m) + a0)
    if parts == 3:
        t = exact_double(a2 * (2**64 % m) + t)
    q = round(float(t) * (1.0 / m))  # rint: round half to even, as Python's round on floats
    return t - m * q


def exact_sym_mod(v, m):
    r = int(v) % m
    half = m >> 1
    return r - m if r >= half + (m & 1) else r


def parts_for(nmod):
    P = 1
    for m in MODULI[:nmod]:
        P *= m
    alpha = (P.bit_length() - 2) // 2
    return alpha, (3 if alpha + 1 >= 63 else 2)


def test_residue_extraction_exact_for_every_modulus_count():
    """Scaled inputs are integers with |a'| <= 2^(alpha+1); for every modulus
    count the kernel's split reduction returns the exact balanced residue
    (three parts from 16 moduli on, where |a'| passes 2^63)."""
    rng = random.Random(11)
    for nmod in (8, 14, 16, 20):
        alpha, parts = parts_for(nmod)
        for _ in range(400):
            e = rng.randint(0, alpha + 1)
            v = float(rng.randint(-2**52, 2**52)) * 2.0 ** max(0, e - 52)
            v = float(round(v))
            for m in MODULI[:nmod]:
                assert split_residue(v, m, parts) == exact_sym_mod(v, m), (nmod, v, m)


def test_two_parts_are_not_enough_at_20_moduli():
    """Planted mistake: the two-part split for the largest scaled inputs
    (|a'| near 2^79) leaves t far above 2^53, so its sums are not exact and
    the residues are wrong."""
    rng = random.Random(12)
    bad = 0
    for _ in range(200):
        v = float(rng.randint(2**52, 2**53 - 1)) * 2.0**26
        for m in MODULI[1:20]:
            try:
                bad += split_residue(v, m, 2) != exact_sym_mod(v, m)
            except AssertionError:
                bad += 1
    assert bad > 0


def balanced(v, m):
    r = v % m
    return r - m if r > (m - 1) // 2 else r


def test_split_residue_exact_next_to_multiples():
    """The split reduction at the values where a rounded quotient can go
    either way: next to multiples and to half-multiples of every odd modulus,
    at every magnitude up to 2^79 and with either part count that covers it."""
    rnd = random.Random(17)
    cases = []
    for m in MODULI[1:]:  # the odd moduli
        for bits in (8, 30, 53, 56, 62, 70, 79):
            for _ in range(200):
                v = int(float(rnd.randrange(-(1 << bits), 1 << bits)))
                cases.append(v)
            for k in (1, 3, 12345, (1 << 40) + 7, (1 << 60) // m):
                for d in (-1, 0, 1):
                    for v in (k * m + d, -(k * m) + d, k * m + (m - 1) // 2 + d):
                        if float(v) == v:
                            cases.append(v)
    for v in cases:
        for m in MODULI:
            for parts in ((2, 3) if abs(v) < 2**62 else (3,)):
                assert split_residue(v, m, parts) == balanced(v, m), (v, m, parts)


