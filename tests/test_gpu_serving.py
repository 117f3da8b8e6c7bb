"""GPU parity of the batched integer multiply and of the serving loop.

Every product of a batch must equal the exact product. This is checked
against the oracle's schoolbook over Z, the same closed form the
two-convolution tests use. It must also be bit-identical to the same
product computed alone through the single-product entry point, so a
result does not depend on what it was batched with.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _items_for(cm, A, B, bits):
    from cumodp_amd.serving import out_words
    n = len(A)
    items = (cm.ZmulItem * n)()
    off = coff = 0
    for r in range(n):
        it = items[r]
        it.a_off, it.b_off = off, off + len(A[r])
        off += len(A[r]) + len(B[r])
        it.na, it.nb = len(A[r]), len(B[r])
        it.wa = it.wb = 1
        it.wc = out_words(int(bits[r]), len(A[r]), len(B[r]))
        it.bits = int(bits[r])
        it.c_off = coff
        coff += (len(A[r]) + len(B[r]) - 1) * it.wc
    packed = np.concatenate([np.concatenate([A[r], B[r]]) for r in range(n)]).view(np.uint64)
    return items, packed, coff


def _run_list(cm, A, B, bits):
    items, packed, nout = _items_for(cm, A, B, bits)
    D = cm.DeviceArray.from_host(packed, np.uint64)
    O = cm.DeviceArray(nout, np.uint64)
    rc = cm.lib().cumodp_two_convolution_list_dev(items, len(items), D.addr(), D.addr(), O.addr(),
                                                  None)
    assert rc == 0, cm.lib().cumodp_last_error()
    cm.synchronize()
    out = O.download()
    res = []
    for r in range(len(items)):
        it = items[r]
        res.append(out[it.c_off:it.c_off + (it.na + it.nb - 1) * it.wc].reshape(-1, it.wc))
    return res


def _signed(rng, n, bits):
    from cumodp_amd.workload import signed_uniform
    return signed_uniform(rng, n, 1 << (bits - 1))


def test_list_mixed_sizes_exact_and_alone(cm, orc):
    from cumodp_amd.workload import SplitMix64
    rng = SplitMix64(501)
    sizes = [(1, 1), (1, 700), (37, 5), (64, 64), (300, 1200), (2048, 2049), (4000, 100),
             (5000, 5000), (129, 8191), (16, 16), (2, 3), (1000, 1000)]
    bitsv = [1, 16, 32, 48, 64, 64, 16, 32, 48, 64, 2, 33]
    A = [_signed(rng, na, b) for (na, _), b in zip(sizes, bitsv)]
    B = [_signed(rng, nb, b) for (_, nb), b in zip(sizes, bitsv)]
    # zero polynomial and all-extreme operands
    A[9] = np.zeros(16, np.int64)
    A[4] = np.full(300, -(2 ** 63) + 1, np.int64)
    B[4] = np.full(1200, 2 ** 63 - 1, np.int64)
    got = _run_list(cm, A, B, bitsv)
    for r in range(len(A)):
        want = orc.zmul_i64(A[r], B[r])[:, :got[r].shape[1]]
        assert (got[r] == want).all(), (r, sizes[r], bitsv[r])
        alone = cm.two_convolution_poly_mul(A[r], B[r], got[r].shape[1])
        assert (alone == got[r]).all(), r


def test_list_merged_classes_same_products(cm, orc):
    """Grouping every width of one transform length under the widest plan
    (cumodp_tune_set_zmul_merge(1)) gives the same words as exact plans."""
    from cumodp_amd.workload import SplitMix64
    rng = SplitMix64(502)
    sizes = [(100, 120), (90, 140), (200, 50), (1000, 900), (1100, 800), (3, 4), (60, 70)]
    bitsv = [16, 64, 48, 1, 63, 32, 20]
    A = [_signed(rng, na, b) for (na, _), b in zip(sizes, bitsv)]
    B = [_signed(rng, nb, b) for (_, nb), b in zip(sizes, bitsv)]
    L = cm.lib()
    before = L.cumodp_tune_get_zmul_merge()
    try:
        assert L.cumodp_tune_set_zmul_merge(0) == 0
        exact = _run_list(cm, A, B, bitsv)
        assert L.cumodp_tune_set_zmul_merge(1) == 0
        merged = _run_list(cm, A, B, bitsv)
        assert L.cumodp_tune_set_zmul_merge(2) == cm.CUMODP_ASSUMPTION_ERROR
    finally:
        L.cumodp_tune_set_zmul_merge(before)
    for r in range(len(A)):
        assert (merged[r] == exact[r]).all(), r
        assert (exact[r] == orc.zmul_i64(A[r], B[r])[:, :exact[r].shape[1]]).all(), r


def test_list_rejects_narrow_output(cm):
    items = (cm.ZmulItem * 1)()
    items[0].na = items[0].nb = 4
    items[0].wa = items[0].wb = items[0].wc = 1
    items[0].bits = 40
    assert cm.lib().cumodp_two_convolution_list_dev(items, 1, None, None, None, None) == \
        cm.CUMODP_ASSUMPTION_ERROR


def test_serving_loop_exact(cm, orc):
    from cumodp_amd import serving
    w = cm.workload5(requests=160, lo=64, hi=3000)
    Q = serving.make_requests(w)
    srv = serving.ZmulServer(slots=2, slot_bytes=64 << 20)
    try:
        # all at once, then paced arrivals with the same data
        done0, st0 = srv.run(Q, np.zeros(Q.n))
        first = [Q.result(r).copy() for r in range(Q.n)]
        Q.OUT.words[:] = 0
        arr = np.cumsum(w["exp_gaps"][:Q.n]) * 2e-4
        done1, st1 = srv.run(Q, arr)
    finally:
        srv.close()
    assert np.isfinite(done0).all() and np.isfinite(done1).all()
    assert (done1 >= arr).all()
    for r in range(Q.n):
        want = orc.zmul_i64(Q.a(r), Q.b(r))[:, :int(Q.wc[r])]
        assert (first[r] == want).all(), r
        assert (Q.result(r) == want).all(), r
    assert st0["batches"] >= 1 and st1["batches"] >= 1
    Q.free()
