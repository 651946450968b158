"""The exact phase's carry fix-up, restated in Python integers and checked against the exact W on
the CPU: W = floor(G V / 2^110) mod 2^64 formed from product columns 2..5, plus, when bits
[98,110) of those columns are all ones, the carry that columns 0..1 add at bit 110, formed as
(A + floor(T01 / 2^64)) >> 46 (A = bits [64,110) of the partial product, T01 = columns 0..1)."""
import random

M64 = (1 << 64) - 1


def _limbs(x):
    return [(x >> (32 * i)) & 0xFFFFFFFF for i in range(4)]


def _phase_word(G, V):
    g, v = _limbs(G), _limbs(V)
    P = sum(g[i] * v[j] << (32 * (i + j)) for i in range(4) for j in range(4) if i + j >= 2)
    W = (P >> 110) & M64
    s3lo = (P >> 96) & 0xFFFFFFFF
    hit = ((s3lo >> 2) & 0xFFF) == 0xFFF
    if hit:
        tlo = (g[0] * v[2] + g[1] * v[1] + g[2] * v[0]) & 0xFFFFFFFF
        c0, c1 = g[0] * v[0], g[0] * v[1] + g[1] * v[0]
        bint = ((c0 >> 32) + c1) >> 32
        A = ((s3lo & 0x3FFF) << 32) | tlo
        W = (W + ((A + bint) >> 46)) & M64
    return W, hit


def test_fixup_exact_on_random_pairs():
    rng = random.Random(7)
    hits = carries = 0
    for _ in range(200_000):
        G, V = rng.getrandbits(128), rng.getrandbits(128)
        W, hit = _phase_word(G, V)
        exact = ((G * V) >> 110) & M64
        assert W == exact
        if hit:
            hits += 1
            g, v = _limbs(G), _limbs(V)
            P = sum(g[i] * v[j] << (32 * (i + j)) for i in range(4) for j in range(4) if i + j >= 2)
            carries += ((P >> 110) & M64) != exact
    # about 200,000 / 4096 = 49 fix-up cases, about half of them with a carry
    assert hits >= 20 and carries >= 5
