"""Host-side logic of the serving loop (no GPU): output widths, admission by
footprint, latency summary."""
import types

import numpy as np

from cumodp_amd import serving


def test_out_words_cover_the_product_bound():
    # 2*bits + ceil(log2(min(na, nb))) + 1 bits must fit
    assert serving.out_words(16, 65, 70) == 1        # 32 + 7 + 1 = 40
    assert serving.out_words(32, 1000, 10) == 2      # 64 + 4 + 1 = 69
    assert serving.out_words(48, 65537, 65537) == 2  # 96 + 17 + 1 = 114
    assert serving.out_words(64, 65537, 65537) == 3  # 128 + 17 + 1 = 146
    assert serving.out_words(64, 1, 1) == 3          # 128 + 0 + 1 = 129
    assert serving.out_words(31, 1, 1) == 1          # 62 + 1 = 63


def _fake_requests(na, nb, bits):
    na, nb, bits = (np.asarray(x, np.int64) for x in (na, nb, bits))
    Q = types.SimpleNamespace(na=na, nb=nb, bits=bits, nc=na + nb - 1)
    Q.wc = np.array([serving.out_words(int(b), int(x), int(y)) for b, x, y in zip(bits, na, nb)])
    return Q


def _server(cap):
    s = object.__new__(serving.ZmulServer)
    s.slot_bytes = cap
    s._plans = {}
    s.max_batch_s = None
    s.sec_per_unit = None
    s.age_factor = 0.0
    s._age = 0.0
    # scratch estimate without the library: 2 x 8 x Ly words per prime, 2 primes
    s._scratch_bytes = lambda bits, ny: 4 * 2 * 8 * (1 << max(4, (int(ny) - 1).bit_length())) * 2
    return s


def test_admission_respects_the_footprint_cap():
    Q = _fake_requests([100] * 10, [100] * 10, [64] * 10)
    one = 8 * (200 + 199 * 3) + 4 * 2 * 8 * 256 * 2
    s = _server(3 * one)
    assert s._fit(Q, 0, 10) == 3
    assert s._fit(Q, 3, 10) == 6
    assert s._fit(Q, 9, 10) == 10
    # a request larger than the cap is admitted alone
    s = _server(one // 2)
    assert s._fit(Q, 0, 10) == 1
    # only arrived requests are considered
    s = _server(100 * one)
    assert s._fit(Q, 2, 5) == 5


def test_admission_respects_the_service_time_cap():
    """Once a rate is learnt, a batch stops at the estimated time cap; the
    first request is always admitted; before any batch is measured only the
    footprint cap applies."""
    Q = _fake_requests([100] * 10, [100] * 10, [64] * 10)
    s = _server(1 << 40)
    s.max_batch_s = 1e-3
    assert s._fit(Q, 0, 10) == 10           # no rate yet
    u = s._units(64, 199)
    s._learn(4 * u, 4e-3)                   # 1 ms per request
    assert abs(s.sec_per_unit * u - 1e-3) < 1e-12
    assert s._fit(Q, 0, 10) == 1
    s.max_batch_s = 3.5e-3
    assert s._fit(Q, 0, 10) == 3
    assert s._fit(Q, 8, 10) == 10
    # the rate is an exponentially weighted average of the batches served
    s._learn(u, 2e-3)
    assert abs(s.sec_per_unit * u - (0.8 * 1e-3 + 0.2 * 2e-3)) < 1e-12


def test_latency_summary():
    arr = np.array([0.0, 0.001, 0.002, 0.003])
    done = arr + np.array([0.002, 0.004, 0.001, 0.003])
    s = serving.latency_summary(done, arr)
    assert abs(s["p50_ms"] - 2.5) < 1e-9
    assert abs(s["max_ms"] - 4.0) < 1e-9


def test_admission_matches_request_by_request_counting():
    """The prefix-sum admission admits exactly the requests that adding them
    one at a time under both caps would admit."""
    rng = np.random.default_rng(7)
    n = 300
    Q = _fake_requests(rng.integers(1, 5000, n), rng.integers(1, 5000, n),
                       rng.choice([16, 32, 48, 64], n))
    for cap_bytes, cap_s in ((1 << 22, None), (1 << 24, 2e-4), (1 << 40, 5e-5), (1 << 21, 1e-3)):
        s = _server(cap_bytes)
        s.max_batch_s = cap_s
        s.sec_per_unit = 1e-12
        for i0 in range(0, n - 1, 17):
            i1 = min(n, i0 + int(rng.integers(1, 120)))
            total, work, j = 0, 0.0, i0
            while j < i1:
                b = 8 * (int(Q.na[j] + Q.nb[j]) + int(Q.nc[j] * Q.wc[j])) + \
                    s._scratch_bytes(Q.bits[j], Q.nc[j])
                u = s._units(Q.bits[j], Q.nc[j])
                if j > i0 and (total + b > cap_bytes or
                               (cap_s is not None and work + u > cap_s / s.sec_per_unit)):
                    break
                total, work, j = total + b, work + u, j + 1
            assert s._fit(Q, i0, i1) == j, (cap_bytes, cap_s, i0, i1)


def test_service_time_cap_grows_with_queue_age():
    """With age_factor a, the cap is max(max_batch_s, a x the oldest waiting
    request's age): a backlog is served in larger batches."""
    Q = _fake_requests([100] * 10, [100] * 10, [64] * 10)
    s = _server(1 << 40)
    s.max_batch_s = 1e-3
    u = s._units(64, 199)
    s._learn(u, 1e-3)                       # 1 ms per request
    s.age_factor = 0.5
    s._age = 0.0
    assert s._fit(Q, 0, 10) == 1            # young queue: the fixed cap
    s._age = 6.5e-3                         # oldest request waited 6.5 ms
    assert s._fit(Q, 0, 10) == 3            # cap 3.25 ms: three requests
    s.age_factor = 0.0
    assert s._fit(Q, 0, 10) == 1
