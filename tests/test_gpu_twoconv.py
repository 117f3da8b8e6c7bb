"""GPU parity of the two-convolution integer multiplier (exact over Z).

Checked against the reference's own vectors (the gtest products of
unit_tests/test_two_convolution_poly_mul.cpp:112-192 and the 27 Maple
a*b = c fixtures of benchmark/*.test) and against the oracle's exact
schoolbook on seeded signed inputs, including the 16384-term workload with
coefficients in (-10^10, 10^10).
"""
import ctypes as C
import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_gtest_products(cm, orc):
    with open(os.path.join(GOLD, "tc_kats.json")) as f:
        k = json.load(f)
    for c in k["products"]:
        got = cm.two_convolution_poly_mul(np.array(c["a"], np.int64), np.array(c["b"], np.int64), 2)
        assert orc.words_to_ints(got) == c["c"]


def test_scale_x_kats(cm):
    with open(os.path.join(GOLD, "tc_kats.json")) as f:
        k = json.load(f)
    for c in k["scale_x"]:
        d = cm.DeviceArray.from_host(np.array(c["A"], np.int32))
        assert cm.lib().cumodp_scale_x_argument_dev(d.addr(), len(c["A"]), c["K"], c["theta"],
                                                    c["prime"], None) == 0
        assert [int(x) for x in d.download()] == c["expected"]


def test_maple_fixtures(cm):
    with open(os.path.join(GOLD, "bench_fixtures.json")) as f:
        idx = json.load(f)
    data = np.load(os.path.join(GOLD, "bench_fixtures.npz"))
    for e in idx:
        key = e["name"]
        got = cm.two_convolution_poly_mul(data[key + "__a"], data[key + "__b"], e["wc"])
        if e["full_c"]:
            assert (got == data[key + "__c"]).all(), key
        else:
            assert (got[data[key + "__cidx"]] == data[key + "__csample"]).all(), key
        assert hashlib.sha256(got.tobytes()).hexdigest() == e["c_sha256"], key


def test_workload3_exact(cm, orc):
    w = cm.workload3()
    got = cm.two_convolution_poly_mul(w["a"], w["b"], 3)
    want = orc.zmul_i64(w["a"], w["b"])
    assert (got == want).all()


@pytest.mark.parametrize("bits", [1, 16, 32, 48, 63])
def test_signed_random(cm, orc, bits):
    rng = np.random.default_rng(bits)
    for na, nb in [(1, 1), (1, 100), (333, 17), (4096, 4000), (20000, 3000)]:
        lim = (1 << bits) - 1
        a = rng.integers(-lim, lim, na, dtype=np.int64, endpoint=True)
        b = rng.integers(-lim, lim, nb, dtype=np.int64, endpoint=True)
        got = cm.two_convolution_poly_mul(a, b, 3)
        assert (got == orc.zmul_i64(a, b)).all(), (bits, na, nb)


@pytest.mark.parametrize("bits", [20, 40, 62])
def test_one_word_fused_equals_two_word(cm, orc, bits):
    """One-word coefficients take the fused path (x transforms inside the
    split and CRT kernels); the same values given as two-word
    two's-complement coefficients take the column-pass path.  Both equal the
    oracle's exact product."""
    rng = np.random.default_rng(100 + bits)
    lim = (1 << bits) - 1
    for na, nb in [(1, 1), (7, 300), (1000, 999), (4096, 2500)]:
        a = rng.integers(-lim, lim, na, dtype=np.int64)
        b = rng.integers(-lim, lim, nb, dtype=np.int64)
        one = cm.two_convolution_poly_mul(a, b, 3)
        a2 = np.stack([a.view(np.uint64), np.where(a < 0, np.uint64(2**64 - 1), np.uint64(0))], 1)
        b2 = np.stack([b.view(np.uint64), np.where(b < 0, np.uint64(2**64 - 1), np.uint64(0))], 1)
        two = cm.two_convolution_poly_mul(a2, b2, 3)
        assert np.array_equal(one, two), (bits, na, nb)
        want = orc.zmul_i64(a, b)
        assert (one == want).all(), (bits, na, nb)


def test_extreme_and_zero(cm, orc):
    top = 2 ** 63 - 1
    a = np.array([-top, top, 0, -1, 1, -top], np.int64)
    b = np.array([top, -top, -top, 0, 7], np.int64)
    assert (cm.two_convolution_poly_mul(a, b, 3) == orc.zmul_i64(a, b)).all()
    z = np.zeros(50, np.int64)
    assert (cm.two_convolution_poly_mul(z, a, 3) == 0).all()


def test_device_entry_point(cm, orc):
    w = cm.workload3(n=4096)
    A = cm.DeviceArray.from_host(w["a"].view(np.uint64), np.uint64)
    B = cm.DeviceArray.from_host(w["b"].view(np.uint64), np.uint64)
    n = 2 * 4096 - 1
    Cd = cm.DeviceArray(n * 2, np.uint64)
    assert cm.lib().cumodp_two_convolution_poly_mul_dev(A.addr(), 4096, 1, B.addr(), 4096, 1,
                                                        Cd.addr(), 2, 34, None) == 0
    cm.synchronize()
    want = orc.zmul_i64(w["a"], w["b"])[:, :2]
    assert (Cd.download().reshape(n, 2) == want).all()
