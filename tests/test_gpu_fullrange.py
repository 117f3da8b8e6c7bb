"""GPU parity at the top of the serving and list-multiply workloads.

The serving workload draws degrees up to 65536, so the two-convolution's
y-transforms reach 2^17 points; the list-multiply workload is 4096 pairs.
These tests run those sizes through the C ABI:

* single and batched integer products at degrees 32768..65536 and 16, 32,
  48 and 64-bit coefficients, against the oracle's two-convolution
  restatement (itself pinned against the exact schoolbook and the
  reference's a*b = c fixtures in test_oracle.py), each product the same
  alone and in the batch;
* the serving loop over the first 40 requests of the full serving workload
  (10 of them above degree 32768), every result equal to the oracle and to
  the same product computed alone;
* the whole 4096-pair list multiply: batched equals one call per pair, and
  256 sampled pairs equal the oracle's product mod p.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (na, nb, bits): y-transforms of 2^16 and 2^17 points
TOP = [(65537, 65537, 64), (50001, 40001, 48), (32769, 65537, 32), (65537, 33001, 16),
       (32769, 32769, 64), (65537, 65, 48)]


def _operands(seed):
    from cumodp_amd.workload import SplitMix64, signed_uniform
    rng = SplitMix64(seed)
    out = []
    for na, nb, bits in TOP:
        bound = 1 << (bits - 1)
        out.append((signed_uniform(rng, na, bound), signed_uniform(rng, nb, bound), bits))
    return out


def _oracle_all(orc, pairs, wcs):
    """The oracle's exact products, computed on worker threads (the C
    oracle releases the GIL)."""
    def one(k):
        a, b, _ = pairs[k]
        return orc.two_convolution_zmul(a.view(np.uint64).reshape(-1, 1),
                                        b.view(np.uint64).reshape(-1, 1), wcs[k])
    with ThreadPoolExecutor(max_workers=8) as ex:
        return list(ex.map(one, range(len(pairs))))


@pytest.fixture(scope="module")
def top_products(cm, orc):
    from cumodp_amd.serving import out_words
    pairs = _operands(4242)
    wcs = [out_words(bits, len(a), len(b)) for a, b, bits in pairs]
    return pairs, wcs, _oracle_all(orc, pairs, wcs)


def test_two_convolution_top_range_single(cm, top_products):
    pairs, wcs, want = top_products
    for k, (a, b, bits) in enumerate(pairs):
        print(f"single product {len(a)} x {len(b)} at {bits} bits", flush=True)
        got = cm.two_convolution_poly_mul(a, b, wcs[k])
        assert np.array_equal(got, want[k]), (len(a), len(b), bits)


def test_two_convolution_top_range_batched(cm, top_products):
    """All top-range products in one cumodp_two_convolution_list_dev call,
    with small products mixed in; each equals the oracle and the same
    product computed alone."""
    from test_gpu_serving import _run_list
    pairs, wcs, want = top_products
    from cumodp_amd.workload import SplitMix64, signed_uniform
    rng = SplitMix64(4343)
    small = [(signed_uniform(rng, n, 1 << 31), signed_uniform(rng, m, 1 << 31), 32)
             for n, m in [(100, 70), (3000, 4000), (1, 9000)]]
    mix = [pairs[0], small[0], pairs[1], pairs[2], small[1], pairs[3], pairs[4], small[2],
           pairs[5]]
    A = [x[0] for x in mix]
    B = [x[1] for x in mix]
    bits = [x[2] for x in mix]
    got = _run_list(cm, A, B, bits)
    top_at = {0: 0, 2: 1, 3: 2, 5: 3, 6: 4, 8: 5}
    for r in range(len(mix)):
        if r in top_at:
            assert np.array_equal(got[r], want[top_at[r]]), r
        alone = cm.two_convolution_poly_mul(A[r], B[r], got[r].shape[1])
        assert np.array_equal(alone, got[r]), r


def test_serving_full_degree_range(cm, orc):
    """The serving loop over the first 40 requests of the serving workload
    (degrees log-uniform in [64, 65536], 16..64-bit coefficients): all at
    once, then Poisson-paced; every result equals the oracle's exact product
    and the product computed alone."""
    from cumodp_amd import serving
    w = cm.workload5()
    n = 40
    big = int((np.maximum(w["deg_a"][:n], w["deg_b"][:n]) > 32768).sum())
    assert big >= 4, big
    Q = serving.make_requests(w, count=n)
    srv = serving.ZmulServer(slots=2, slot_bytes=256 << 20)
    try:
        print(f"serving {n} requests, {big} above degree 32768", flush=True)
        done0, st0 = srv.run(Q, np.zeros(n))
        first = [Q.result(r).copy() for r in range(n)]
        Q.OUT.words[:] = 0
        arr = np.cumsum(w["exp_gaps"][:n]) * 5e-4
        done1, st1 = srv.run(Q, arr)
        second = [Q.result(r).copy() for r in range(n)]
        pairs = [(Q.a(r).copy(), Q.b(r).copy(), int(Q.bits[r])) for r in range(n)]
        wcs = [int(Q.wc[r]) for r in range(n)]
    finally:
        srv.close()
    assert np.isfinite(done0).all() and np.isfinite(done1).all()
    assert st0["batches"] >= 1 and st1["batches"] >= 1
    want = _oracle_all(orc, pairs, wcs)
    for r in range(n):
        assert np.array_equal(first[r], want[r]), r
        assert np.array_equal(second[r], want[r]), r
        alone = cm.two_convolution_poly_mul(pairs[r][0], pairs[r][1], wcs[r])
        assert np.array_equal(alone, want[r]), r
    Q.free()


def test_workload4_full_batch(cm, orc):
    """The whole 4096-pair list multiply (degrees log-uniform in
    [32, 16384] mod 469762049): one batched call equals one call per pair
    word for word, and 256 pairs spread over the batch equal the oracle."""
    w = cm.workload4()
    p = w["p"]
    items, nf, ng, nh = cm.make_items(w["df"], w["dg"])
    dF, dG = cm.DeviceArray.from_host(w["F"]), cm.DeviceArray.from_host(w["G"])
    dH1, dH2 = cm.DeviceArray(nh), cm.DeviceArray(nh)
    cm.list_fftmul_dev(items, dF, dG, dH1, p)
    cm.list_fftmul_dev(items, dF, dG, dH2, p, per_call=True)
    cm.synchronize()
    H1, H2 = dH1.download(), dH2.download()
    for d in (dF, dG, dH1, dH2):
        d.free()
    assert np.array_equal(H1, H2)
    idx = np.linspace(0, len(items) - 1, 256).astype(np.int64)

    def check(k):
        it = items[int(k)]
        f = w["F"][it.f_off:it.f_off + it.df + 1]
        g = w["G"][it.g_off:it.g_off + it.dg + 1]
        return bool(np.array_equal(H1[it.h_off:it.h_off + it.df + it.dg + 1],
                                   orc.poly_mul(f, g, p)))
    with ThreadPoolExecutor(max_workers=8) as ex:
        ok = list(ex.map(check, idx))
    assert all(ok), [int(idx[i]) for i in range(len(idx)) if not ok[i]][:8]
