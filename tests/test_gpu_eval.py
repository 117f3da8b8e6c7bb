"""GPU parity of fast multipoint evaluation (values of F at 2^k points).

The oracle is the definition, Horner's rule. Evaluation at the 2^k
points is checked exactly for k <= 12 and on a sample of points for
2^16.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 6, 7, 9, 12])
def test_fast_evaluation_exact(cm, orc, k):
    from cumodp_amd.workload import SplitMix64, uniform_mod
    p = 469762049
    rng = SplitMix64(600 + k)
    n = 1 << k
    pts = uniform_mod(rng, n, p)
    F = uniform_mod(rng, n, p)
    got = cm.fast_evaluation(F, pts, p)
    assert (got == orc.poly_eval(F, pts, p)).all()


@pytest.mark.parametrize("k", [6, 11, 13])
def test_fast_evaluation_exact_31bit(cm, orc, k):
    """The same at p = 2013265921 (above 2^30: the canonical butterflies;
    the fused tree, subinverse and remainder kernels at full 31-bit width),
    with coefficients and points drawn near p."""
    p = 2013265921
    rng = np.random.default_rng(900 + k)
    n = 1 << k
    pts = (p - 1 - rng.integers(0, 1 << 20, n)).astype(np.int64).astype(np.int32)
    pts = np.unique(pts)
    while len(pts) < n:                                   # distinct points, topped up
        pts = np.unique(np.concatenate([pts, rng.integers(0, p, n).astype(np.int32)]))
    pts = pts[:n]
    F = (p - 1 - rng.integers(0, 3, n)).astype(np.int64).astype(np.int32)
    got = cm.fast_evaluation(F, pts, p)
    assert (got == orc.poly_eval(F, pts, p)).all()


def test_fast_evaluation_2e16_sampled(cm, orc):
    w = cm.workload4_tree()
    p, pts = w["p"], w["points"]
    from cumodp_amd.workload import SplitMix64, uniform_mod
    F = uniform_mod(SplitMix64(616), len(pts), p)
    got = cm.fast_evaluation(F, pts, p)
    idx = np.linspace(0, len(pts) - 1, 512).astype(int)
    assert (got[idx] == orc.poly_eval(F, pts[idx], p)).all()


def test_fast_evaluation_edge_cases(cm, orc):
    p = 2013265921
    n = 1 << 10
    rng = np.random.default_rng(3)
    pts = rng.integers(0, p, n).astype(np.int32)
    # repeated points, zero and the point 0 itself
    pts[5] = pts[6] = pts[700]
    pts[9] = 0
    for F in (np.zeros(n, np.int32), np.full(1, 77, np.int32),
              rng.integers(0, p, 300).astype(np.int32)):
        got = cm.fast_evaluation(F, pts, p)
        assert (got == orc.poly_eval(F, pts, p)).all()
    # F = the tree root vanishes at every point
    root = orc.subproduct_level(pts, 10, p)[0][:n]    # monic x^n dropped: evaluate x^n too
    full = orc.subproduct_level(pts, 10, p)[0]
    assert (orc.poly_eval(full, pts, p) == 0).all()
    got = cm.fast_evaluation(root, pts, p)
    xn = np.array([pow(int(x), n, p) for x in pts], np.int64)
    assert ((got.astype(np.int64) + xn) % p == 0).all()
