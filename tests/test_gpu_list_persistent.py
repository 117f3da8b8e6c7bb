"""The list multiply's persistent launch against its one-launch-per-class
form and against the oracle.

Every task kind of the persistent kernel is covered:
* tiles of every transform size 2^4 .. 2^14;
* 2^15-point products (two 2^14-point halves recombined), with operands
  longer and shorter than 2^14 on either side;
* schoolbook products, including outputs longer than one 4096-word task and
  (cut-off 1024) shorter operands longer than one 512-term pass;
* 2^16-point products after the persistent launch (batched four-step);
* 60,000 small products in one call (tens of thousands of workgroups).
Both primes below 2^30 (lazy butterflies) and above (canonical) run.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _degrees(rng):
    degs = [(0, 0), (7, 8), (15, 16), (40, 30), (100, 90), (200, 300), (600, 400),
            (1000, 1000), (2000, 2047), (4000, 4095), (8000, 8191), (16383, 1),
            (16383, 300), (300, 16383), (9000, 12000), (16383, 16383), (20000, 5000),
            (3000, 29000), (16384, 16384), (30000, 2767), (40000, 20000), (257, 20000),
            (250, 9000), (256, 256), (100, 12000), (5000, 256)]
    degs += [(int(a), int(b)) for a, b in rng.integers(0, 3000, (60, 2))]
    degs += [(int(a), int(b)) for a, b in rng.integers(8000, 16384, (12, 2))]
    return degs


def _run(cm, L, items, dF, dG, nh, p, persistent, cutoff=None):
    old_p, old_c = L.cumodp_tune_get_list_persistent(), L.cumodp_tune_get_list_plain()
    assert L.cumodp_tune_set_list_persistent(persistent) == 0
    if cutoff is not None:
        assert L.cumodp_tune_set_list_plain(cutoff) == 0
    try:
        dH = cm.DeviceArray(nh)
        dH.upload(np.full(nh, 0x7fffffff, np.int32))
        cm.list_fftmul_dev(items, dF, dG, dH, p)
        cm.synchronize()
        H = dH.download()
        dH.free()
        return H
    finally:
        L.cumodp_tune_set_list_persistent(old_p)
        L.cumodp_tune_set_list_plain(old_c)


@pytest.mark.parametrize("p", [469762049, 2013265921])
@pytest.mark.parametrize("cutoff", [None, 1024])
def test_persistent_equals_class_launches_and_oracle(cm, orc, p, cutoff):
    L = cm.lib()
    rng = np.random.default_rng(31 if p < (1 << 30) else 32)
    degs = _degrees(rng)
    df = np.array([a for a, _ in degs])
    dg = np.array([b for _, b in degs])
    items, nf, ng, nh = cm.make_items(df, dg)
    F = rng.integers(0, p, nf).astype(np.int64).astype(np.int32)
    G = rng.integers(0, p, ng).astype(np.int64).astype(np.int32)
    # all p - 1 in a few products: the top of the lazy range
    for i in (15, 18, 20):
        it = items[i]
        F[it.f_off:it.f_off + it.df + 1] = p - 1
        G[it.g_off:it.g_off + it.dg + 1] = p - 1
    dF, dG = cm.DeviceArray.from_host(F), cm.DeviceArray.from_host(G)
    try:
        H1 = _run(cm, L, items, dF, dG, nh, p, 1, cutoff)
        H0 = _run(cm, L, items, dF, dG, nh, p, 0, cutoff)
    finally:
        dF.free(), dG.free()
    assert np.array_equal(H1, H0)
    for i, (a, b) in enumerate(degs):
        it = items[i]
        f = F[it.f_off:it.f_off + a + 1]
        g = G[it.g_off:it.g_off + b + 1]
        got = H1[it.h_off:it.h_off + a + b + 1]
        assert np.array_equal(got, orc.poly_mul(f, g, p)), (p, a, b)


def test_persistent_repeated_calls_and_graph(cm, orc):
    """The task counter starts at zero in every call and every replay."""
    import ctypes as C
    L = cm.lib()
    p = 469762049
    rng = np.random.default_rng(33)
    degs = _degrees(rng)
    items, nf, ng, nh = cm.make_items([a for a, _ in degs], [b for _, b in degs])
    F = rng.integers(0, p, nf).astype(np.int32)
    G = rng.integers(0, p, ng).astype(np.int32)
    dF, dG, dH = cm.DeviceArray.from_host(F), cm.DeviceArray.from_host(G), cm.DeviceArray(nh)
    s = L.cumodp_stream_create()
    try:
        first = None
        for _ in range(3):
            dH.upload(np.zeros(nh, np.int32))
            cm.list_fftmul_dev(items, dF, dG, dH, p, s)
            assert L.cumodp_stream_synchronize(C.c_void_p(s)) == 0
            H = dH.download()
            first = H if first is None else first
            assert np.array_equal(H, first)
        g = cm.Graph.capture(s, lambda: cm.list_fftmul_dev(items, dF, dG, dH, p, s))
        for _ in range(3):
            dH.upload(np.zeros(nh, np.int32))
            g.launch(s)
            assert L.cumodp_stream_synchronize(C.c_void_p(s)) == 0
            assert np.array_equal(dH.download(), first)
        g.destroy()
    finally:
        L.cumodp_stream_destroy(C.c_void_p(s))
        for d in (dF, dG, dH):
            d.free()
    it = items[15]
    assert np.array_equal(first[it.h_off:it.h_off + it.df + it.dg + 1],
                          orc.poly_mul(F[it.f_off:it.f_off + it.df + 1],
                                       G[it.g_off:it.g_off + it.dg + 1], p))


def test_persistent_many_small_products(cm, orc):
    """60,000 products of degree below 300 (schoolbook and tile tasks of
    every small size, tens of thousands of workgroups, a descriptor blob of
    several MiB): the persistent launch equals the per-class launches, and
    a spread sample equals the oracle."""
    L = cm.lib()
    p = 469762049
    rng = np.random.default_rng(35)
    n = 60000
    df = rng.integers(0, 300, n)
    dg = rng.integers(0, 300, n)
    items, nf, ng, nh = cm.make_items(df, dg)
    F = rng.integers(0, p, nf).astype(np.int32)
    G = rng.integers(0, p, ng).astype(np.int32)
    dF, dG = cm.DeviceArray.from_host(F), cm.DeviceArray.from_host(G)
    try:
        H1 = _run(cm, L, items, dF, dG, nh, p, 1)
        H0 = _run(cm, L, items, dF, dG, nh, p, 0)
    finally:
        dF.free(), dG.free()
    assert np.array_equal(H1, H0)
    for i in np.linspace(0, n - 1, 40).astype(int):
        it = items[int(i)]
        f = F[it.f_off:it.f_off + it.df + 1]
        g = G[it.g_off:it.g_off + it.dg + 1]
        assert np.array_equal(H1[it.h_off:it.h_off + it.df + it.dg + 1], orc.poly_mul(f, g, p)), i
