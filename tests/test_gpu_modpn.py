"""modpn, the reference's CPU library, calling this library through the C ABI.

oracle/_ref/libmodpn_fmul.so is modpn's FMUL.c (and the three files it needs)
compiled from the reference's sources and linked against libcumodp_amd.so.
With CUDA_TAG set, EX_KN_Mont_FFTMul_OPT2_AS_GENE_1 hands transforms of 2^16
points and more to cumodp_fftmul_uni(n - 1, A, degA, A, degB, B, p), output
aliasing the first operand (modpn/src/FMUL.c:1943-1955).  The product must be
the one modpn's own CPU FFT computes, word for word over all n words.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand(rng, n, p):
    return rng.integers(0, p, n).astype(np.int32)


@pytest.mark.parametrize("p,da,db", [
    (469762049, 40000, 25535),      # 2^16-point transform
    (469762049, 100000, 150000),    # 2^18
    (998244353, 1 << 20, 1 << 19),  # 2^21
    (919601153, 300000, 200000),    # 2^19 (877*2^20+1)
])
def test_modpn_gpu_branch_matches_its_cpu_fft(cm, orc, modpn, p, da, db):
    rng = np.random.default_rng(da ^ db)
    A, B = _rand(rng, da + 1, p), _rand(rng, db + 1, p)
    assert cm.lib().can_call_fftmul_uni(da, db) == 1
    cm.profile_enable(True)
    cm.profile_reset()
    try:
        gpu = modpn.kn_fftmul(A, B, p, cuda_tag=True)
        launched = cm.profile_read()
    finally:
        cm.profile_enable(False)
    assert launched, "modpn did not reach the GPU branch"
    cpu = modpn.kn_fftmul(A, B, p, cuda_tag=False)
    assert gpu.shape == cpu.shape and (gpu == cpu).all(), (p, da, db)
    want = orc.poly_mul(A, B, p)
    assert (gpu[:len(want)] == want).all() and not gpu[len(want):].any()


def test_modpn_gets_exact_31_bit_products_through_the_gpu(cm, orc, modpn):
    """For 2013265921 modpn's CPU arithmetic is out of range, but its GPU
    branch (this library) returns the exact product."""
    p = 2013265921
    rng = np.random.default_rng(31)
    A, B = _rand(rng, 70001, p), _rand(rng, 50000, p)
    got = modpn.kn_fftmul(A, B, p, cuda_tag=True)
    want = orc.poly_mul(A, B, p)
    assert (got[:len(want)] == want).all() and not got[len(want):].any()


def test_workload1_equals_modpn(cm, modpn):
    """The 2048-point multiply mod 469762049 (two degree-1023 operands) equals
    the reference's own CPU multiply."""
    w = cm.workload1()
    assert (cm.poly_mul(w["F"], w["G"], w["p"]) == modpn.mul(w["F"], w["G"], w["p"])).all()


def test_list_multiply_sample_equals_modpn(cm, modpn):
    """The first 64 pairs of the 4096-pair uneven list multiply (batched
    cumodp_list_fftmul_dev) equal modpn's CPU multiply pair by pair."""
    w = cm.workload4()
    p = w["p"]
    items, nf, ng, nh = cm.make_items(w["df"][:64], w["dg"][:64])
    dF = cm.DeviceArray.from_host(w["F"][:nf])
    dG = cm.DeviceArray.from_host(w["G"][:ng])
    dH = cm.DeviceArray(nh)
    try:
        cm.list_fftmul_dev(items, dF, dG, dH, p)
        cm.synchronize()
        H = dH.download()
    finally:
        for d in (dF, dG, dH):
            d.free()
    for it in items:
        f = w["F"][it.f_off:it.f_off + it.df + 1]
        g = w["G"][it.g_off:it.g_off + it.dg + 1]
        assert (H[it.h_off:it.h_off + it.df + it.dg + 1] == modpn.mul(f, g, p)).all()
