"""Schoolbook products: mulCUDA (src/opt_plain_mul.cu:258-384) and the
batched cumodp_list_plainmul_dev, against the oracle's plain product
(uint64 % p, any modulus) and, below 2^30, modpn's plain multiply.

The kernel folds its 64-bit accumulators after every four products; the
moduli include those with the largest folded bound relative to 2^64 (just
below 2^31, and just above 2^30 where 2^32 mod p is close to p) and moduli the
NTT cannot use (no 2-power root of unity).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def is_prime(n):
    if n < 2:
        return False
    for q in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % q == 0:
            return n == q
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


P_TWO = next(q for q in range((1 << 30) + 1, (1 << 30) + 1000, 2) if is_prime(q))
P_TOP = next(q for q in range((1 << 31) - 1, (1 << 31) - 1000, -2) if is_prime(q))
MODULI = [2013265921, 469762049, P_TOP, P_TWO, 1000003, 65537, 257, 3]


def _rand(rng, n, p):
    return rng.integers(0, p, n, dtype=np.int64).astype(np.int32)


@pytest.mark.parametrize("p", MODULI)
def test_mulcuda_small_and_uneven(cm, orc, p):
    rng = np.random.default_rng(p % 997)
    for n, m in [(1, 1), (1, 7), (7, 1), (2, 300), (300, 2), (64, 64), (127, 129),
                 (256, 255), (1000, 31), (5, 2000)]:
        A, B = _rand(rng, n, p), _rand(rng, m, p)
        got = cm.mulCUDA(A, B, p)
        assert (got == orc.plain_polymul(A, B, p)).all(), (p, n, m)


@pytest.mark.parametrize("p", [2013265921, P_TWO, 1000003])
def test_plain_dev_long_operands(cm, orc, p):
    """Several 256-output chunks and several 512-term passes of B, both operand
    orders, and an output pointer that is not 16-byte aligned."""
    rng = np.random.default_rng(7)
    for n, m in [(3000, 1700), (1700, 3000), (4097, 513)]:
        A, B = _rand(rng, n, p), _rand(rng, m, p)
        want = orc.plain_polymul(A, B, p)
        dA, dB = cm.DeviceArray.from_host(A), cm.DeviceArray.from_host(B)
        dC = cm.DeviceArray(n + m + 3)
        try:
            for off in (0, 1, 3):
                cm.plain_mul_dev(dC.addr(off), dA, n, dB, m, p)
                cm.synchronize()
                assert (dC.download()[off:off + n + m - 1] == want).all(), (p, n, m, off)
        finally:
            for d in (dA, dB, dC):
                d.free()


@pytest.mark.parametrize("p", [2013265921, P_TOP, P_TWO])
def test_plain_all_max_residues(cm, orc, p):
    """Operands of p - 1 everywhere: the largest products, 2500 per output."""
    A = np.full(2500, p - 1, np.int32)
    B = np.full(2600, p - 1, np.int32)
    dA, dB = cm.DeviceArray.from_host(A), cm.DeviceArray.from_host(B)
    dC = cm.DeviceArray(5099)
    try:
        cm.plain_mul_dev(dC, dA, 2500, dB, 2600, p)
        cm.synchronize()
        assert (dC.download() == orc.plain_polymul(A, B, p)).all()
    finally:
        for d in (dA, dB, dC):
            d.free()


def test_list_plain_equals_list_ntt(cm, orc):
    """The first 512 pairs of the uneven list workload (degrees 32..16384 mod
    469762049): schoolbook list == NTT list, and == the oracle on a sample."""
    w = cm.workload4()
    p = w["p"]
    items, nf, ng, nh = cm.make_items(w["df"][:512], w["dg"][:512])
    dF, dG = cm.DeviceArray.from_host(w["F"][:nf]), cm.DeviceArray.from_host(w["G"][:ng])
    dH1, dH2 = cm.DeviceArray(nh), cm.DeviceArray(nh)
    try:
        cm.list_plainmul_dev(items, dF, dG, dH1, p)
        cm.list_fftmul_dev(items, dF, dG, dH2, p)
        cm.synchronize()
        H1, H2 = dH1.download(), dH2.download()
    finally:
        for d in (dF, dG, dH1, dH2):
            d.free()
    assert (H1 == H2).all()
    for it in list(items)[:40]:
        f = w["F"][it.f_off:it.f_off + it.df + 1]
        g = w["G"][it.g_off:it.g_off + it.dg + 1]
        assert (H1[it.h_off:it.h_off + it.df + it.dg + 1] == orc.poly_mul(f, g, p)).all()


def test_mulcuda_matches_modpn_plain(cm, modpn):
    """Below modpn's FFT cut-off (degree 138) its EX_Mont_Mul is its plain
    multiply (modpn/src/FMUL.c:181-193)."""
    p = 469762049
    rng = np.random.default_rng(138)
    for n, m in [(1, 1), (40, 60), (100, 38), (69, 69)]:
        A, B = _rand(rng, n, p), _rand(rng, m, p)
        assert (cm.mulCUDA(A, B, p) == modpn.mul(A, B, p)).all()


def test_plain_rejects_bad_arguments(cm):
    L = cm.lib()
    A = np.ones(4, np.int32)
    C_out = np.full(7, -5, np.int32)
    cm.mulCUDA(A, A, 10, C_out)            # even modulus: refused, C untouched
    assert (C_out == -5).all()
    cm.mulCUDA(A, A, 1 << 31, C_out)       # modulus beyond 2^31
    assert (C_out == -5).all()
    assert L.cumodp_plain_mul_dev(None, None, 0, None, 4, 7, None) != cm.CUMODP_SUCCESS
    items = (cm.MulItem * 1)()
    items[0].df = -1
    assert L.cumodp_list_plainmul_dev(items, 1, None, None, None, 7, None) != cm.CUMODP_SUCCESS


def test_timing_prints_are_opt_in(cm):
    """CUMODP_VERBOSE=1 turns on the reference's stdout timing lines
    (src/stockham.cu:600, src/opt_plain_mul.cu:367-369); by default the
    library prints nothing.  One child process per setting."""
    import os
    import subprocess
    import sys
    code = ("import numpy as np, cumodp_amd as cm\n"
            "F = np.arange(1, 301, dtype=np.int32); G = np.arange(2, 202, dtype=np.int32)\n"
            "H = np.zeros(500, np.int32)\n"
            "cm.stockham_poly_mul_host(499, H, 299, F, 199, G, 469762049)\n"
            "cm.mulCUDA(F[:10], G[:5], 469762049)\n"
            "C = np.zeros(14, np.int32); fn = cm.cxx('mulCUDA'); fn.restype = None\n"
            "fn(cm._ptr(F), cm._ptr(G), cm._ptr(C), 10, 5, 469762049, 0)\n")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    outs = {}
    for v in ("0", "1"):
        env = dict(os.environ, CUMODP_VERBOSE=v)
        r = subprocess.run([sys.executable, "-c", code], cwd=root, env=env, capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        outs[v] = r.stdout
    assert outs["0"] == ""
    lines = outs["1"].splitlines()
    assert lines[0].startswith("FFT-based: ")
    assert len(lines) == 2 and lines[1].split()[:2] == ["10", "5"]   # verify=1 call is silent
