"""Pins the CPU oracle (oracle/) before anything is compared against it.

Each check is chosen so a plausible slip in the oracle (a dropped term, a
wrong sign or index, a transposed operand) fails: exhaustive arithmetic on a
small prime, the DFT definition by brute force, closed-form products, the
reference's own gtest vectors and Maple fixtures, and evaluation of tree
levels at random points.  CPU only.
"""
import ctypes
import hashlib
import json
import os
import random

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
P1, P2 = 469762049, 2013265921
CRT_PRIMES = [962592769, 950009857, 940572673]


def test_arith_exhaustive_small_prime(orc):
    p = 257
    for a in range(p):
        for b in range(0, p, 7):
            assert orc.add_mod(a, b, p) == (a + b) % p
            assert orc.sub_mod(a, b, p) == (a - b) % p
            assert orc.mul_mod(a, b, p) == (a * b) % p
        assert orc.neg_mod(a, p) == (-a) % p


@pytest.mark.parametrize("p", [P1, P2] + CRT_PRIMES)
def test_arith_large_primes(orc, p):
    rnd = random.Random(p)
    for _ in range(2000):
        a, b = rnd.randrange(p), rnd.randrange(p)
        assert orc.mul_mod(a, b, p) == a * b % p
        assert orc.add_mod(a, b, p) == (a + b) % p
        assert orc.sub_mod(a, b, p) == (a - b) % p
        if a:
            assert orc.inv_mod(a, p) * a % p == 1
        e = rnd.randrange(-1000, 1 << 30)
        assert orc.pow_mod(a or 1, e, p) == pow(a or 1, e, p)


@pytest.mark.parametrize("p", [P1, P2, 257] + CRT_PRIMES)
def test_fourier_degree_and_roots(orc, p):
    K = orc.fourier_degree(p)
    assert (p - 1) % (1 << K) == 0 and ((p - 1) >> K) % 2 == 1
    for e in range(1, K + 1):
        w = orc.primitive_root(e, p)
        assert pow(w, 1 << e, p) == 1 and pow(w, 1 << (e - 1), p) == p - 1
    # libc-seeded search is deterministic
    assert orc.primitive_root(K, p) == orc.primitive_root(K, p)


@pytest.mark.parametrize("p", [P1, P2, 257, 962592769])
def test_stockham_is_the_dft(orc, p):
    """The Stockham restatement equals Y_j = sum_i X_i w^(ij) (natural order)."""
    rng = np.random.default_rng(p)
    for k in range(1, 11):
        if k > orc.fourier_degree(p):
            break
        n = 1 << k
        w = orc.primitive_root(k, p)
        X = rng.integers(0, p, n)
        Y = orc.dft_naive(X, w, p)
        assert (orc.stockham(X, w, p) == Y).all()
        assert (orc.gsnn_fft(X, w, p) == Y).all()
        assert (orc.inverse_stockham(Y, w, p) == X).all()


def test_stockham_equals_gsnn_large(orc):
    rng = np.random.default_rng(5)
    for k in (12, 14, 16):
        X = rng.integers(0, P1, 1 << k)
        w = orc.primitive_root(k, P1)
        assert (orc.stockham(X, w, P1) == orc.gsnn_fft(X, w, P1)).all()


def test_list_stockham_rows(orc):
    rng = np.random.default_rng(6)
    m, k = 5, 8
    X = rng.integers(0, P1, m << k)
    w = orc.primitive_root(k, P1)
    Y = orc.list_stockham(X, m, k, w, P1).reshape(m, -1)
    for r in range(m):
        assert (Y[r] == orc.dft_naive(X[r << k:(r + 1) << k], w, P1)).all()


def test_twiddle_powers(orc):
    w = orc.primitive_root(10, P1)
    W = orc.get_powers(512, w, P1)
    assert [int(x) for x in W[:20]] == [pow(w, j, P1) for j in range(20)]
    assert int(W[511]) == pow(w, 511, P1)


def test_products_closed_forms(orc):
    # check_polymul's inputs (src/fft_aux.cu:843-844): F = 1 (d1+1 ones), G_i = i
    d1, d2 = 700, 500
    F = np.ones(d1 + 1, np.int32)
    G = np.arange(d2 + 1, dtype=np.int32)
    for H in (orc.plain_polymul(F, G, P1), orc.ntt_polymul(F, G, P1)):
        for k in (0, 1, 250, 500, 700, 1000, 1200):
            lo, hi = max(0, k - d1), min(k, d2)
            assert int(H[k]) == (hi * (hi + 1) // 2 - (lo - 1) * lo // 2) % P1
    # F = G = 1: coefficient k counts the pairs i + j = k
    F = np.ones(300, np.int32)
    H = orc.ntt_polymul(F, F, P2)
    assert [int(x) for x in H] == [min(k + 1, 599 - k) for k in range(599)]
    # test_poly_mul inputs (src/stockham.cu:637-650): F_i = i, G_i = i+1, p = 257
    d = 64
    F = np.arange(d + 1, dtype=np.int32)
    G = np.arange(1, d + 2, dtype=np.int32)
    want = [sum(i * (k - i + 1) for i in range(max(0, k - d), min(k, d) + 1)) % 257
            for k in range(2 * d + 1)]
    assert [int(x) for x in orc.ntt_polymul(F, G, 257)] == want
    assert [int(x) for x in orc.plain_polymul(F, G, 257)] == want


@pytest.mark.parametrize("p", [P1, P2, 940572673])
def test_plain_equals_ntt_random(orc, p):
    rng = np.random.default_rng(p + 1)
    for lf, lg in ((1, 1), (1, 300), (513, 1), (1000, 777), (2048, 2048)):
        F, G = rng.integers(0, p, lf), rng.integers(0, p, lg)
        a = orc.plain_polymul(F, G, p)
        assert (a == orc.ntt_polymul(F, G, p)).all()
    # one column of the big product against Python integers
    F, G = rng.integers(0, p, 900), rng.integers(0, p, 900)
    H = orc.plain_polymul(F, G, p)
    k = 777
    assert int(H[k]) == sum(int(F[i]) * int(G[k - i]) for i in range(k + 1)) % p


def test_host_wrapper_semantics(orc):
    rng = np.random.default_rng(9)
    F, G = rng.integers(0, P1, 200).astype(np.int32), rng.integers(0, P1, 100).astype(np.int32)
    full = orc.plain_polymul(F, G, P1)                       # s = 299, n = 512
    H = np.full(600, -7, np.int32)
    assert orc.stockham_poly_mul_host(599, H, F, G, P1) == 0
    assert (H[:299] == full).all() and (H[299:512] == 0).all() and (H[512:] == -7).all()
    H = np.full(600, -7, np.int32)
    assert orc.stockham_poly_mul_host(99, H, F, G, P1) == 0
    assert (H[:100] == full[:100]).all() and (H[100:] == -7).all()
    H = np.full(10, -7, np.int32)                            # n = 128 < 256: refused
    assert orc.stockham_poly_mul_host(9, H, F[:60], G[:60], P1) == 1
    assert (H == -7).all()


# ---- integer multiplication ---------------------------------------------------
def _kats():
    with open(os.path.join(GOLD, "tc_kats.json")) as f:
        return json.load(f)


def _words(vals, W=None):
    import oracle
    if W is None:
        m = max(abs(int(v)) for v in vals)
        W = max(1, (m.bit_length() + 1 + 63) // 64)
    return oracle.ints_to_words([int(v) for v in vals], W)


def test_bit_width_kats(orc):
    for c in _kats()["bit_width"]:
        a, b = [int(x) for x in c["a"]], [int(x) for x in c["b"]]
        W = max(_words(a).shape[1], _words(b).shape[1])
        got = max(orc.bit_width(_words(a, W)), orc.bit_width(_words(b, W)))
        assert got == c["bits"], c


def test_limb_split_kats(orc):
    for c in _kats()["limb_split"]:
        got = orc.limb_split(_words(c["a"]), c["K"], c["M"], c["prime"])
        assert [int(x) for x in got] == c["expected"], c


def test_scale_x_kats(orc):
    for c in _kats()["scale_x"]:
        got = orc.scale_x(c["A"], c["K"], c["theta"], c["prime"])
        assert [int(x) for x in got] == c["expected"], c


def test_integer_product_kats(orc):
    for c in _kats()["products"]:
        a, b = _words(c["a"], 1), _words(c["b"], 1)
        assert orc.words_to_ints(orc.zmul_mw(a, b, 2)) == c["c"]
        assert orc.words_to_ints(orc.two_convolution_zmul(a, b, 2)) == c["c"]
        ai, bi = np.array(c["a"], np.int64), np.array(c["b"], np.int64)
        assert orc.words_to_ints(orc.zmul_i64(ai, bi)) == c["c"]


def _fixtures():
    with open(os.path.join(GOLD, "bench_fixtures.json")) as f:
        idx = json.load(f)
    data = np.load(os.path.join(GOLD, "bench_fixtures.npz"))
    return idx, data


def _digest(cw):
    return hashlib.sha256(np.ascontiguousarray(cw, np.uint64).tobytes()).hexdigest()


def _check_fixture(e, data, c):
    key = e["name"]
    if e["full_c"]:
        assert (c == data[key + "__c"]).all(), key
    else:
        idx = data[key + "__cidx"]
        assert (c[idx] == data[key + "__csample"]).all(), key
    assert _digest(c) == e["c_sha256"], key


def test_maple_fixtures_schoolbook(orc):
    idx, data = _fixtures()
    for e in idx:
        if e["na"] * e["nb"] * e["wa"] * e["wb"] > (1 << 27):
            continue      # the two largest run in test_maple_fixtures_large
        c = orc.zmul_mw(data[e["name"] + "__a"], data[e["name"] + "__b"], e["wc"])
        _check_fixture(e, data, c)


def test_maple_fixtures_large(orc):
    idx, data = _fixtures()
    for e in idx:
        if e["na"] * e["nb"] * e["wa"] * e["wb"] <= (1 << 27):
            continue
        if e["wa"] == 1 and e["wb"] == 1:
            a = data[e["name"] + "__a"].view(np.int64).ravel()
            b = data[e["name"] + "__b"].view(np.int64).ravel()
            c3 = orc.zmul_i64(a, b)
            c = c3[:, :e["wc"]]
            assert (c3[:, e["wc"]:] == 0).all() if e["wc"] < 3 else True
        else:
            c = orc.zmul_mw(data[e["name"] + "__a"], data[e["name"] + "__b"], e["wc"])
        _check_fixture(e, data, c)


def test_maple_fixtures_two_convolution(orc):
    idx, data = _fixtures()
    for e in idx:
        if e["na"] > 4096 and e["wa"] == 1:
            continue       # covered by the schoolbook; keeps the CPU suite short
        c = orc.two_convolution_zmul(data[e["name"] + "__a"], data[e["name"] + "__b"], e["wc"])
        _check_fixture(e, data, c)


def test_two_convolution_random_signed(orc):
    rng = np.random.default_rng(11)
    for bits in (3, 16, 34, 48, 63):
        for na, nb in ((1, 1), (7, 300), (257, 129)):
            lim = 1 << bits
            a = rng.integers(-lim + 1, lim, na, dtype=np.int64)
            b = rng.integers(-lim + 1, lim, nb, dtype=np.int64)
            A, B = orc.ints_to_words(a.tolist(), 1), orc.ints_to_words(b.tolist(), 1)
            want = orc.words_to_ints(orc.zmul_i64(a, b))
            assert orc.words_to_ints(orc.two_convolution_zmul(A, B, 3)) == want
    # extreme magnitudes: all coefficients -(2^63-1) and 2^63-1
    a = np.full(64, -(2 ** 63 - 1), np.int64)
    b = np.full(64, 2 ** 63 - 1, np.int64)
    A, B = orc.ints_to_words(a.tolist(), 1), orc.ints_to_words(b.tolist(), 1)
    assert orc.words_to_ints(orc.two_convolution_zmul(A, B, 3)) == \
        [-(min(k + 1, 127 - k)) * (2 ** 63 - 1) ** 2 for k in range(127)]


def test_crt_constants(orc):
    """include/two_convolution_poly_mul.h:55-65 are the CRT constants of the primes."""
    p1, p2, p3 = CRT_PRIMES
    assert p1 * p2 == 914472618826924033 and p2 * p3 == 893553310574837761
    assert p3 * p1 == 905388453748801537
    assert (-412536414 * p2 * p3) % p1 == 1
    assert (316662352 * p3 * p1) % p2 == 1
    assert (89582607 * p1 * p2) % p3 == 1


# ---- subproduct tree ----------------------------------------------------------
def test_subproduct_levels_by_evaluation(orc):
    rng = np.random.default_rng(12)
    pts = rng.integers(0, P1, 1 << 9).astype(np.int32)
    for lev in (0, 1, 3, 6, 9):
        T = orc.subproduct_level(pts, lev, P1)
        assert T.shape == ((1 << 9) >> lev, (1 << lev) + 1)
        assert (T[:, -1] == 1).all()
        for j in (0, T.shape[0] - 1):
            block = [int(x) for x in pts[j << lev:(j + 1) << lev]]
            for r in (3, 123456789):
                want = 1
                for x in block:
                    want = want * (r - x) % P1
                got = 0
                for coef in reversed(T[j].tolist()):
                    got = (got * r + coef) % P1
                assert got == want


def test_splitmix_matches_numpy_stream(orc):
    import cumodp_amd.workload as wl
    st = ctypes.c_uint64(12345)
    c = [orc.lib().orc_splitmix64(ctypes.byref(st)) for _ in range(8)]
    assert [int(x) for x in wl.SplitMix64(12345).next(8)] == c


def test_poly_eval_closed_forms(orc):
    """Horner evaluation against the definition and the tree's roots."""
    p = 469762049
    rng = np.random.default_rng(7)
    F = rng.integers(0, p, 40).astype(np.int32)
    X = rng.integers(0, p, 25).astype(np.int32)
    want = [sum(int(F[t]) * pow(int(x), t, p) for t in range(len(F))) % p for x in X]
    assert [int(v) for v in orc.poly_eval(F, X, p)] == want
    # the root of the subproduct tree vanishes at every point
    pts = rng.integers(0, p, 16).astype(np.int32)
    root = orc.subproduct_level(pts, 4, p)[0]
    assert (orc.poly_eval(root, pts, p) == 0).all()
    # constant and zero polynomials
    assert (orc.poly_eval(np.array([5], np.int32), X, p) == 5).all()
    assert (orc.poly_eval(np.zeros(3, np.int32), X, p) == 0).all()


def test_polymul_items_equals_per_item_products(orc):
    """The OpenMP batch driver (the parallel CPU-baseline leg) gives each
    item's schoolbook / NTT product, whatever the thread count."""
    rng = np.random.default_rng(31)
    p = 469762049
    shapes = [(0, 0), (3, 900), (64, 64), (700, 5), (1500, 1400), (63, 2000)]
    F = [rng.integers(0, p, a + 1) for a, _ in shapes]
    G = [rng.integers(0, p, b + 1) for _, b in shapes]
    desc, fo, go, ho = [], 0, 0, 0
    for (a, b) in shapes:
        desc.append([fo, go, ho, a, b])
        fo, go, ho = fo + a + 1, go + b + 1, ho + a + b + 1
    Fa, Ga = np.concatenate(F), np.concatenate(G)
    for threads in (1, 4):
        H = orc.polymul_items(np.array(desc), Fa, Ga, ho, p, plain_cut=64, threads=threads)
        for (f0, g0, h0, a, b), f, g in zip(desc, F, G):
            assert np.array_equal(H[h0:h0 + a + b + 1], orc.plain_polymul(f, g, p))
