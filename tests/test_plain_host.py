"""Host-side facts the schoolbook kernel relies on (no GPU).

The kernel keeps sums of products of canonical residues in 64-bit
accumulators and folds acc -> hi * (2^32 mod p) + lo after every four
products (cumodp_amd/csrc/accum.h).  That is safe when the largest folded
value plus four products stays below 2^64, for every modulus p < 2^31.
"""
import random


def folded_max(p):
    c = (1 << 32) % p
    return (2**32 - 1) * c + 2**32 - 1


def four_terms(p):
    return folded_max(p) + 4 * (p - 1) ** 2 < 2**64


def test_four_products_between_folds_fit_every_modulus():
    # the tightest cases: p just below 2^31 (k = floor(2^32/p) = 2), just
    # above 2^32/3 and 2^32/4 (c = 2^32 mod p close to p), and small p
    edges = [(1 << 31) - 1, (1 << 31) - 3, (1 << 32) // 3 + 1, (1 << 32) // 3 + 3,
             (1 << 30) + 1, (1 << 30) + 3, 2013265921, 469762049, 3, 5, 7, 257]
    for p in edges:
        assert four_terms(p), p
    for q in range((1 << 31) - 1, (1 << 31) - 400000, -2):
        assert four_terms(q), q
    rnd = random.Random(5)
    for _ in range(200000):
        q = rnd.randrange(3, 1 << 31) | 1
        assert four_terms(q), q


def test_five_products_would_not_fit():
    """Four is the limit: a fifth product overflows for the benchmark prime."""
    p = 2013265921
    assert folded_max(p) + 5 * (p - 1) ** 2 >= 2**64
