"""cumodp_fftmul_uni_dev_noalloc: the non-allocating multiply.

Its products must equal cumodp_fftmul_uni_dev's bit for bit. Two calls run
in a row on the same scratch, the second with a smaller input, so a value
left over from the first call would show. The free device memory must not
change across calls that reach the four-step path on a stream that has never
had library scratch. Sizes beyond 2^22 and primes whose tables are not built
yet are rejected.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZE_TOO_LARGE = 0x202
ASSUMPTION_ERROR = 0x002


def _free_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(0), C.c_size_t(0)
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _operands(cm, seed, lf, lg, p):
    rng = cm.SplitMix64(seed)
    return cm.uniform_mod(rng, lf, p), cm.uniform_mod(rng, lg, p)


def _both(cm, F, G, p, stream):
    """(allocating product, non-allocating product) of F and G on stream."""
    L = cm.lib()
    dF, dG = cm.DeviceArray.from_host(F), cm.DeviceArray.from_host(G)
    nout = len(F) + len(G) - 1
    dA, dB = cm.DeviceArray(nout), cm.DeviceArray(nout)
    try:
        cm.fftmul_uni_dev(nout - 1, dA, len(F) - 1, dF, len(G) - 1, dG, p, stream)
        dB.upload(np.full(nout, 0x7fffffff, np.uint32))
        cm.fftmul_uni_dev(nout - 1, dB, len(F) - 1, dF, len(G) - 1, dG, p, stream, noalloc=True)
        assert L.cumodp_stream_synchronize(C.c_void_p(stream)) == 0
        return dA.download(), dB.download()
    finally:
        for d in (dF, dG, dA, dB):
            d.free()


@pytest.mark.parametrize("p", [2013265921, 469762049])
def test_noalloc_equals_allocating_two_calls_in_a_row(cm, orc, p):
    L = cm.lib()
    s = L.cumodp_stream_create()
    try:
        # 2^22-point transform (the largest), then 2^16, then an LDS size
        for seed, lf, lg in ((71, 1 << 21, (1 << 21) - 5), (72, 40000, 20001), (73, 1000, 777)):
            F, G = _operands(cm, seed, lf, lg, p)
            want, got = _both(cm, F, G, p, s)
            assert np.array_equal(got, want), (p, lf, lg)
            if lf <= 40000:
                assert np.array_equal(got, orc.ntt_polymul(F, G, p)), (p, lf, lg)
    finally:
        L.cumodp_stream_destroy(C.c_void_p(s))


def test_noalloc_allocates_nothing(cm):
    L = cm.lib()
    p = 2013265921
    F, G = _operands(cm, 81, 30000, 30000, p)
    nout = len(F) + len(G) - 1
    dF, dG, dH = cm.DeviceArray.from_host(F), cm.DeviceArray.from_host(G), cm.DeviceArray(nout)
    dR = cm.DeviceArray(nout)
    # tables of this prime and size (2^16) are built by an allocating call on
    # the default stream
    cm.fftmul_uni_dev(nout - 1, dR, len(F) - 1, dF, len(G) - 1, dG, p)
    cm.synchronize()
    s = L.cumodp_stream_create()
    try:
        # first use of the new stream: an LDS-sized product (no scratch)
        cm.fftmul_uni_dev(30, dH, 15, dF, 15, dG, p, s, noalloc=True)
        assert L.cumodp_stream_synchronize(C.c_void_p(s)) == 0
        before = _free_bytes()
        for _ in range(3):
            cm.fftmul_uni_dev(nout - 1, dH, len(F) - 1, dF, len(G) - 1, dG, p, s, noalloc=True)
        assert L.cumodp_stream_synchronize(C.c_void_p(s)) == 0
        after = _free_bytes()
        assert after == before
        assert np.array_equal(dH.download(), dR.download())
    finally:
        L.cumodp_stream_destroy(C.c_void_p(s))
        for d in (dF, dG, dH, dR):
            d.free()


def test_noalloc_rejects_large_sizes_and_unbuilt_tables(cm):
    L = cm.lib()
    p = 2013265921
    fn = L.cumodp_fftmul_uni_dev_noalloc
    dH = cm.DeviceArray(16)
    # df + dg + 1 = 2^22 + 1 needs a 2^23-point transform: rejected before any
    # pointer is touched
    assert fn((1 << 22), dH.addr(), 1 << 21, dH.addr(), 1 << 21, dH.addr(), p, None) == \
        SIZE_TOO_LARGE
    assert b"2^22" in L.cumodp_last_error()
    # a prime no other call has used: its tables do not exist
    q = 167772161   # 5 * 2^25 + 1
    F, G = _operands(cm, 91, 9000, 9000, q)
    nout = len(F) + len(G) - 1
    dF, dG, dA, dB = (cm.DeviceArray.from_host(F), cm.DeviceArray.from_host(G),
                      cm.DeviceArray(nout), cm.DeviceArray(nout))
    try:
        assert fn(nout - 1, dB.addr(), len(F) - 1, dF.addr(), len(G) - 1, dG.addr(), q,
                  None) == ASSUMPTION_ERROR
        assert b"not" in L.cumodp_last_error()
        cm.fftmul_uni_dev(nout - 1, dA, len(F) - 1, dF, len(G) - 1, dG, q)
        cm.fftmul_uni_dev(nout - 1, dB, len(F) - 1, dF, len(G) - 1, dG, q, noalloc=True)
        cm.synchronize()
        assert np.array_equal(dA.download(), dB.download())
    finally:
        for d in (dH, dF, dG, dA, dB):
            d.free()
