"""The oracle against the reference's own CPU multiply (modpn, built from its
C sources into oracle/_ref), and modpn's fallback when no GPU is usable.

modpn's arithmetic is exact for primes below 2^30 only (signed 32-bit
intermediates in (-p, 2p), modpn/src/utilityFunctions.c:81-106), so these
checks use 30-bit and smaller Fourier primes; the 31-bit benchmark prime is
pinned by the oracle's other checks.
"""
import numpy as np
import pytest

PRIMES = [469762049, 998244353, 919601153]   # 7*2^26+1, 119*2^23+1, 877*2^20+1

# (degA, degB): modpn's plain/FFT cut-off is degree 138 (modpn/src/FMUL.h:36)
SIZES = [(1, 0), (0, 7), (68, 69), (69, 68), (137, 0), (100, 38), (255, 255),
         (1000, 23), (4095, 4096), (30000, 17)]


def _rand(rng, n, p):
    return rng.integers(0, p, n).astype(np.int32)


@pytest.mark.parametrize("p", PRIMES)
def test_modpn_products_equal_oracle(modpn, orc, p):
    rng = np.random.default_rng(p % 1000)
    for da, db in SIZES:
        A, B = _rand(rng, da + 1, p), _rand(rng, db + 1, p)
        want = orc.poly_mul(A, B, p)
        assert (modpn.mul(A, B, p) == want).all(), ("EX_Mont_Mul", p, da, db)
        assert (modpn.fft_mul(A, B, p) == want).all(), ("EX_Mont_FFTMul", p, da, db)
        got = modpn.kn_fftmul(A, B, p, cuda_tag=False)
        assert (got[:len(want)] == want).all() and not got[len(want):].any(), (p, da, db)


def test_modpn_extremes_equal_oracle(modpn, orc):
    """All-(p-1) operands and a small prime (257 = 2^8+1) at its largest transform."""
    p = 469762049
    A = np.full(3000, p - 1, np.int32)
    assert (modpn.fft_mul(A, A[:2000], p) == orc.poly_mul(A, A[:2000], p)).all()
    rng = np.random.default_rng(3)
    A, B = _rand(rng, 100, 257), _rand(rng, 157, 257)
    assert (modpn.fft_mul(A, B, 257) == orc.poly_mul(A, B, 257)).all()


def test_modpn_falls_back_to_cpu_without_gpu(modpn, orc, cm):
    """With CUDA_TAG set and a 2^17-point product, modpn asks
    can_call_fftmul_uni (this library) first; without a usable GPU the answer
    is no and its CPU FFT runs (modpn/src/FMUL.c:1943-1963)."""
    if cm.lib().global_memory_in_bytes() != 0:
        pytest.skip("a GPU is visible: the GPU branch is covered by test_gpu_modpn.py")
    p = 469762049
    rng = np.random.default_rng(17)
    A, B = _rand(rng, 70000, p), _rand(rng, 60000, p)
    got = modpn.kn_fftmul(A, B, p, cuda_tag=True)
    want = orc.poly_mul(A, B, p)
    assert (got[:len(want)] == want).all() and not got[len(want):].any()


def test_modpn_cpu_arithmetic_is_30_bit(modpn, orc):
    """Documents why the checks above avoid 31-bit primes: modpn's own CPU
    product mod 2013265921 differs from the exact one."""
    p = 2013265921
    rng = np.random.default_rng(31)
    A, B = _rand(rng, 60, p), _rand(rng, 50, p)
    assert not (modpn.mul(A, B, p) == orc.poly_mul(A, B, p)).all()
