"""GPU parity of the NTT and NTT-multiply path against the CPU oracle.

Every comparison is bit-exact (integer residues).  Full-size products are
also checked through the size-independent identity H(r) = F(r) G(r) mod p at
random points r.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P1, P2 = 469762049, 2013265921
CRT = [962592769, 950009857, 940572673]


def eval_mod(c, r, p):
    """sum_i c_i r^i mod p, vectorised."""
    c = np.asarray(c, np.int64).astype(np.uint64)
    n = len(c)
    B = 4096
    small = np.empty(B, np.uint64)
    x = 1
    for j in range(B):
        small[j] = x
        x = x * r % p
    nb = (n + B - 1) // B
    big = np.empty(nb, np.uint64)
    y, rb = 1, pow(r, B, p)
    for a in range(nb):
        big[a] = y
        y = y * rb % p
    pw = ((big[:, None] * small[None, :]) % np.uint64(p)).ravel()[:n]
    return int(((c * pw) % np.uint64(p)).sum() % np.uint64(p))


def rand_poly(rng, n, p):
    return rng.integers(0, p, n, dtype=np.int64).astype(np.int32)


def test_workload1_exact(cm, orc):
    w = cm.workload1()
    want = orc.plain_polymul(w["F"], w["G"], w["p"])
    assert (cm.poly_mul(w["F"], w["G"], w["p"]) == want).all()
    H = np.zeros(2047, np.int32)
    cm.stockham_poly_mul_host(2046, H, 1023, w["F"], 1023, w["G"], w["p"])
    assert (H == want).all()


@pytest.mark.parametrize("p", [P1, P2] + CRT)
def test_multiply_size_sweep(cm, orc, p):
    rng = np.random.default_rng(p % 1000)
    for logn in range(8, 21):
        n = 1 << logn
        lf = n // 2 + int(rng.integers(0, n // 4))
        lg = n - lf + 1 - int(rng.integers(0, n // 8))
        F, G = rand_poly(rng, lf, p), rand_poly(rng, lg, p)
        got = cm.poly_mul(F, G, p)
        assert (got == orc.ntt_polymul(F, G, p)).all(), (p, logn)


def test_multiply_small_and_edge_lengths_dev(cm, orc):
    rng = np.random.default_rng(3)
    for lf, lg in [(1, 1), (1, 7), (2, 2), (9, 8), (16, 1), (17, 16), (100, 29), (129, 128),
                   (4097, 1), (4096, 4097), (8193, 8192), (1, 40000), (70000, 3)]:
        F, G = rand_poly(rng, lf, P1), rand_poly(rng, lg, P1)
        dF, dG = cm.DeviceArray.from_host(F), cm.DeviceArray.from_host(G)
        dH = cm.DeviceArray(lf + lg - 1)
        cm.fftmul_uni_dev(lf + lg - 2, dH, lf - 1, dF, lg - 1, dG, P1)
        got = dH.download()
        assert (got == orc.poly_mul(F, G, P1)).all(), (lf, lg)


def test_host_wrapper_semantics(cm, orc):
    rng = np.random.default_rng(4)
    F, G = rand_poly(rng, 200, P1), rand_poly(rng, 100, P1)      # s = 299, n = 512
    full = orc.plain_polymul(F, G, P1)
    H = np.full(600, -7, np.int32)
    cm.stockham_poly_mul_host(599, H, 199, F, 99, G, P1)
    assert (H[:299] == full).all() and (H[299:512] == 0).all() and (H[512:] == -7).all()
    H = np.full(600, -7, np.int32)
    assert cm.cumodp_fftmul_uni(99, H, 199, F, 99, G, P1) == cm.CUMODP_SUCCESS
    assert (H[:100] == full[:100]).all() and (H[100:] == -7).all()
    # H aliases F (modpn calls cumodp_fftmul_uni(n-1, A, degA, A, degB, B, p))
    A = np.zeros(512, np.int32)
    A[:200] = F
    assert cm.lib().cumodp_fftmul_uni(511, A.ctypes.data_as(C.c_void_p), 199,
                                      A.ctypes.data_as(C.c_void_p), 99,
                                      G.ctypes.data_as(C.c_void_p), P1) == 0
    assert (A[:299] == full).all() and (A[299:] == 0).all()
    # n < 256: refused, H untouched
    H = np.full(40, -7, np.int32)
    cm.stockham_poly_mul_host(39, H, 19, F[:20], 19, G[:20], P1)
    assert (H == -7).all()


@pytest.mark.parametrize("k2", [12, 13, 14])
def test_four_step_splits(cm, orc, k2):
    """Every supported column/row split of the four-step multiply is exact."""
    rng = np.random.default_rng(k2)
    p = P2
    try:
        for logn in (15, 19, 22):
            if logn - k2 > 13 or logn - k2 < 1:
                continue
            assert cm.lib().cumodp_tune_set_split(logn, k2) == 0
            lf, lg = (1 << logn) // 2, (1 << logn) // 2 - 3
            F, G = rand_poly(rng, lf, p), rand_poly(rng, lg, p)
            assert (cm.poly_mul(F, G, p) == orc.ntt_polymul(F, G, p)).all(), (logn, k2)
    finally:
        for logn in (15, 19, 22):
            cm.lib().cumodp_tune_set_split(logn, 13 if logn >= 21 else 12)


def extreme_operands(rng, lf, lg, p, mix):
    """All p - 1, or a random mix of 0, 1, p - 2 and p - 1."""
    if mix:
        vals = np.array([0, 1, p - 2, p - 1], np.int64)
        return (vals[rng.integers(0, 4, lf)].astype(np.int32),
                vals[rng.integers(0, 4, lg)].astype(np.int32))
    return (np.full(lf, p - 1, np.int64).astype(np.int32),
            np.full(lg, p - 1, np.int64).astype(np.int32))


@pytest.mark.parametrize("p", [998244353, 1004535809, P2])
def test_extreme_residues_multiply(cm, orc, p):
    """Operands made of the largest residues (all p - 1, then a mix of 0, 1,
    p - 2 and p - 1) through the one-workgroup multiply (2^9, 2^12), the
    four-step multiply (2^14 .. 2^20) and the list multiply's mixed tiles.
    Below 2^30 these kernels keep their values in [0, 2p) between
    butterflies; these inputs drive every sum and difference to the top of
    that range.  998244353 and 1004535809 are the NTT primes just below
    2^30."""
    rng = np.random.default_rng(p % 977)
    for logn in (9, 12, 14, 17, 20):
        n = 1 << logn
        for mix in (False, True):
            F, G = extreme_operands(rng, n // 2, n // 2 - 1, p, mix)
            got = cm.poly_mul(F, G, p)
            assert (got == orc.ntt_polymul(F, G, p)).all(), (p, logn, mix)
    degs = [(300, 20), (1000, 1000), (2047, 2048), (5000, 3000), (40, 700)]
    pairs = [extreme_operands(rng, a + 1, b + 1, p, i % 2 == 1) for i, (a, b) in enumerate(degs)]
    items = (cm.MulItem * len(degs))()
    fo = go = ho = 0
    for i, (a, b) in enumerate(degs):
        items[i].f_off, items[i].g_off, items[i].h_off = fo, go, ho
        items[i].df, items[i].dg = a, b
        fo += a + 1; go += b + 1; ho += a + b + 1
    dF = cm.DeviceArray.from_host(np.concatenate([f for f, _ in pairs]))
    dG = cm.DeviceArray.from_host(np.concatenate([g for _, g in pairs]))
    dH = cm.DeviceArray(ho)
    cm.list_fftmul_dev(items, dF, dG, dH, p)
    H = dH.download()
    for i, (a, b) in enumerate(degs):
        got = H[items[i].h_off:items[i].h_off + a + b + 1]
        assert (got == orc.poly_mul(pairs[i][0], pairs[i][1], p)).all(), (p, a, b)


@pytest.mark.parametrize("p", [P1, P2])
def test_list_multiply_lds_2e14_class(cm, orc, p):
    """600 products of 2^14 points (df + dg + 1 in (8192, 16384], uneven)
    in one list: a class that large runs one product per 1024-thread
    workgroup in LDS.  Bit-exact against the one-call-per-product form
    (four-step at this size) and, on a sample, against the oracle."""
    rng = np.random.default_rng(14 + p % 7)
    m = 600
    df = rng.integers(4096, 8000, m)
    dg = rng.integers(np.maximum(8192 - df, 300), 16383 - df)   # above the plain cutoff
    assert ((df + dg + 1 > 8192) & (df + dg + 1 <= 16384)).all()
    items, nf, ng, nh = cm.make_items(df, dg)
    F, G = rand_poly(rng, nf, p), rand_poly(rng, ng, p)
    dF, dG = cm.DeviceArray.from_host(F), cm.DeviceArray.from_host(G)
    dH1, dH2 = cm.DeviceArray(nh), cm.DeviceArray(nh)
    cm.list_fftmul_dev(items, dF, dG, dH1, p)
    cm.list_fftmul_dev(items, dF, dG, dH2, p, per_call=True)
    H1, H2 = dH1.download(), dH2.download()
    assert np.array_equal(H1, H2)
    for i in (0, 1, 299, m - 1):
        it = items[i]
        f = F[it.f_off:it.f_off + it.df + 1]
        g = G[it.g_off:it.g_off + it.dg + 1]
        assert (H1[it.h_off:it.h_off + it.df + it.dg + 1] == orc.ntt_polymul(f, g, p)).all(), i


def test_workload2_full_size(cm, orc):
    """The 2^24-point multiply mod 2013265921, bit-exact against the oracle."""
    w = cm.workload2()
    F, G, p = w["F"], w["G"], w["p"]
    dF, dG = cm.DeviceArray.from_host(F), cm.DeviceArray.from_host(G)
    dH = cm.DeviceArray(len(F) + len(G) - 1)
    cm.fftmul_uni_dev(len(F) + len(G) - 2, dH, len(F) - 1, dF, len(G) - 1, dG, p)
    H = dH.download()
    for r in (2, 1234567, p - 5):
        assert eval_mod(H, r, p) == eval_mod(F, r, p) * eval_mod(G, r, p) % p
    assert (H == orc.ntt_polymul(F, G, p)).all()


@pytest.mark.parametrize("logn", [25, 26])
def test_largest_transforms_evaluation_identity(cm, orc, logn):
    """2^25- and 2^26-point multiplies (2^14-word rows, the largest supported
    size) checked at random points: H(r) = F(r) G(r) mod p, plus a product
    with operands longer than half the transform (no zero upper half)."""
    rng = np.random.default_rng(logn)
    p = P2
    n = 1 << logn
    for lf, lg in [(n // 2, n // 2), (n // 2 + 12345, n // 2 - 12346)]:
        F, G = rand_poly(rng, lf, p), rand_poly(rng, lg, p)
        dF, dG = cm.DeviceArray.from_host(F), cm.DeviceArray.from_host(G)
        dH = cm.DeviceArray(lf + lg - 1)
        try:
            cm.fftmul_uni_dev(lf + lg - 2, dH, lf - 1, dF, lg - 1, dG, p)
            H = dH.download()
        finally:
            for d in (dF, dG, dH):
                d.free()
        pts = np.array([3, 987654321, p - 2], np.int32)
        hv, fv, gv = orc.poly_eval(H, pts, p), orc.poly_eval(F, pts, p), orc.poly_eval(G, pts, p)
        for i in range(len(pts)):
            assert int(hv[i]) == int(fv[i]) * int(gv[i]) % p, (logn, lf, lg, int(pts[i]))


@pytest.mark.parametrize("p", [P1, P2, 257])
def test_forward_transform_raw_outputs(cm, orc, p):
    """cumodp_fft_uni returns the reference's natural-order DFT for the root w."""
    rng = np.random.default_rng(5)
    for k in range(4, 19):
        if k > orc.fourier_degree(p):
            break
        n = 1 << k
        w = orc.primitive_root(k, p)
        X = rand_poly(rng, n, p)
        want = orc.stockham(X, w, p) if k <= 12 else orc.gsnn_fft(X, w, p)
        got = cm.fft_uni(X, w, p)
        assert (got == want).all(), (p, k)
        assert (cm.invfft_uni(got, w, p) == X).all(), (p, k)


def test_list_transforms(cm, orc):
    rng = np.random.default_rng(6)
    for m, k in [(1, 8), (3, 8), (7, 10), (5, 13), (3, 15), (2, 17)]:
        X = rand_poly(rng, m << k, P1)
        w = orc.primitive_root(k, P1)
        want = orc.list_stockham(X, m, k, w, P1) if k <= 10 else np.concatenate(
            [orc.gsnn_fft(X[r << k:(r + 1) << k], w, P1) for r in range(m)])
        assert (cm.list_stockham_host(X, m, k, w, P1) == want).all(), (m, k)
        Y = X.copy()
        cm.cxx("list_inv_stockham_host")(Y.ctypes.data_as(C.c_void_p), m, k, w, P1)
        assert (Y == X).all()


def test_device_entry_points(cm, orc):
    rng = np.random.default_rng(7)
    k, p = 11, P1
    n = 1 << k
    w = orc.primitive_root(k, p)
    # get_powers_binary
    dW = cm.DeviceArray(n // 2)
    cm.cxx("get_powers_binary")(k - 1, C.c_void_p(dW.ptr), w, p)
    assert (dW.download() == orc.get_powers(n // 2, w, p)).all()
    # stockham_dev with the table, and with the root
    X = rand_poly(rng, n, p)
    dX = cm.DeviceArray.from_host(X)
    cm.cxx("stockham_dev_W")(C.c_void_p(dX.ptr), n, k, C.c_void_p(dW.ptr), p)
    Y = dX.download()
    assert (Y == orc.stockham(X, w, p)).all()
    # inverse with the table of inverse powers
    winv = orc.inv_mod(w, p)
    dWi = cm.DeviceArray(n // 2)
    cm.cxx("get_powers_binary")(k - 1, C.c_void_p(dWi.ptr), winv, p)
    cm.cxx("inverse_stockham_dev_W")(C.c_void_p(dX.ptr), n, k, C.c_void_p(dWi.ptr), p)
    assert (dX.download() == X).all()
    cm.cxx("stockham_dev_w")(C.c_void_p(dX.ptr), n, k, w, p)
    cm.cxx("inverse_stockham_dev_w")(C.c_void_p(dX.ptr), n, k, w, p)
    assert (dX.download() == X).all()
    # stockham_poly_mul_dev is the cyclic product of length n; G is overwritten
    for e in (9, 15):
        m = 1 << e
        F, G = rand_poly(rng, m, p), rand_poly(rng, m, p)
        lin = orc.ntt_polymul(F, G, p).astype(np.int64)
        cyc = lin[:m].copy()
        cyc[:m - 1] += lin[m:]
        cyc %= p
        dF, dG = cm.DeviceArray.from_host(F), cm.DeviceArray.from_host(G)
        cm.cxx("stockham_poly_mul_dev")(m, e, C.c_void_p(dF.ptr), C.c_void_p(dG.ptr), p)
        assert (dF.download() == cyc).all(), e
    # pointwise / scale / expand / reset
    A, B = rand_poly(rng, 1000, p), rand_poly(rng, 1000, p)
    dA, dB = cm.DeviceArray.from_host(A), cm.DeviceArray.from_host(B)
    cm.cxx("pointwise_mul_dev")(1000, 10, C.c_void_p(dA.ptr), C.c_void_p(dB.ptr), p)
    assert (dA.download() == (A.astype(np.int64) * B % p)).all()
    cm.cxx("scale_vector_dev")(12345, 1000, C.c_void_p(dB.ptr), p)
    assert (dB.download() == (B.astype(np.int64) * 12345 % p)).all()
    cm.cxx("expand_to_fft_dev")(1000, 300, C.c_void_p(dB.ptr))
    got = dB.download()
    assert (got[:300] == (B[:300].astype(np.int64) * 12345 % p)).all() and (got[300:] == 0).all()
    cm.cxx("reset_vector_dev")(300, C.c_void_p(dB.ptr))
    assert (dB.download() == 0).all()


def biv_oracle(orc, X, em, wm, en, wn, p):
    m, n = 1 << em, 1 << en
    Y = orc.list_stockham(X, m, en, wn, p).reshape(m, n)
    Z = orc.list_stockham(np.ascontiguousarray(Y.T).ravel(), n, em, wm, p).reshape(n, m)
    return np.ascontiguousarray(Z.T).ravel()


def test_bivariate_transforms(cm, orc):
    rng = np.random.default_rng(8)
    p = P1
    for em, en in [(4, 8), (8, 8), (6, 11), (9, 5), (3, 14)]:
        wm, wn = orc.primitive_root(em, p), orc.primitive_root(en, p)
        X = rand_poly(rng, 1 << (em + en), p)
        Y = X.copy()
        assert cm.lib().cumodp_fft_bivariate(Y.ctypes.data_as(C.c_void_p), em, wm, en, wn, p) == 0
        assert (Y == biv_oracle(orc, X, em, wm, en, wn, p)).all(), (em, en)
        assert cm.lib().cumodp_invfft_bivariate(Y.ctypes.data_as(C.c_void_p), em, wm, en, wn,
                                                p) == 0
        assert (Y == X).all(), (em, en)


def test_list_bivariate_transforms(cm, orc):
    """list_bivariate_stockham_dev (include/list_stockham.h:94-95): q arrays
    of 2^em x 2^en, each equal to the bivariate DFT of that array, through
    both overloads (power tables and roots), with em != en and em = 1."""
    rng = np.random.default_rng(18)
    for p in (P1, P2):
        for q, em, en in [(3, 4, 9), (2, 9, 5), (5, 1, 6), (1, 6, 14), (7, 3, 4)]:
            wm, wn = orc.primitive_root(em, p), orc.primitive_root(en, p)
            m, n = 1 << em, 1 << en
            X = rand_poly(rng, q * m * n, p)
            want = np.concatenate([biv_oracle(orc, X[b * m * n:(b + 1) * m * n], em, wm, en, wn,
                                              p) for b in range(q)])
            d = cm.DeviceArray.from_host(X)
            cm.cxx("list_bivariate_stockham_dev_w")(C.c_void_p(d.ptr), q, em, wm, en, wn, p)
            assert (d.download() == want).all(), (p, q, em, en, "roots")
            dWm = cm.DeviceArray.from_host(orc.get_powers(max(m // 2, 1), wm, p))
            dWn = cm.DeviceArray.from_host(orc.get_powers(n // 2, wn, p))
            d2 = cm.DeviceArray.from_host(X)
            cm.cxx("list_bivariate_stockham_dev_W")(C.c_void_p(d2.ptr), q, em,
                                                    C.c_void_p(dWm.ptr), en,
                                                    C.c_void_p(dWn.ptr), p)
            assert (d2.download() == want).all(), (p, q, em, en, "tables")
            for a in (d, d2, dWm, dWn):
                a.free()


def test_bivariate_multiply(cm, orc):
    """bi_stockham_poly_mul_dev is the 2-D cyclic convolution."""
    rng = np.random.default_rng(9)
    p, em, en = P1, 5, 6
    m, n = 1 << em, 1 << en
    F = np.zeros((m, n), np.int64)
    G = np.zeros((m, n), np.int64)
    F[:m // 2, :n // 2] = rng.integers(0, p, (m // 2, n // 2))
    G[:m // 2, :n // 2] = rng.integers(0, p, (m // 2, n // 2))
    want = np.zeros((m, n), object)
    Fo, Go = F.astype(object), G.astype(object)
    for a in range(m // 2):
        for b in range(n // 2):
            if F[a, b]:
                want[a:a + m // 2, b:b + n // 2] += Fo[a, b] * Go[:m // 2, :n // 2]
    want = (want % p).astype(np.int64)
    dF = cm.DeviceArray.from_host(F.astype(np.int32).ravel())
    dG = cm.DeviceArray.from_host(G.astype(np.int32).ravel())
    cm.cxx("bi_stockham_poly_mul_dev")(m, em, n, en, C.c_void_p(dF.ptr), C.c_void_p(dG.ptr), p)
    assert (dF.download().reshape(m, n) == want).all()


def test_list_multiply_uneven(cm, orc):
    rng = np.random.default_rng(10)
    p = P1
    degs = [(0, 0), (5, 3), (31, 31), (100, 2000), (4095, 1), (4096, 4095), (3000, 3000),
            (8000, 190), (9000, 9000), (20000, 12000), (64, 64)] + \
        [(int(a), int(b)) for a, b in rng.integers(0, 600, (40, 2))]
    Fs = [rand_poly(rng, a + 1, p) for a, _ in degs]
    Gs = [rand_poly(rng, b + 1, p) for _, b in degs]
    items = (cm.MulItem * len(degs))()
    fo = go = ho = 0
    for i, (a, b) in enumerate(degs):
        items[i].f_off, items[i].g_off, items[i].h_off = fo, go, ho
        items[i].df, items[i].dg = a, b
        fo += a + 1; go += b + 1; ho += a + b + 1
    dF = cm.DeviceArray.from_host(np.concatenate(Fs))
    dG = cm.DeviceArray.from_host(np.concatenate(Gs))
    dH = cm.DeviceArray(ho)
    cm.list_fftmul_dev(items, dF, dG, dH, p)
    H = dH.download()
    for i, (a, b) in enumerate(degs):
        got = H[items[i].h_off:items[i].h_off + a + b + 1]
        assert (got == orc.poly_mul(Fs[i], Gs[i], p)).all(), (a, b)
