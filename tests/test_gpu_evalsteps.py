"""The reference-named evaluation drivers (fastEvaluation, onlySubproductTree,
onlyTrees) and their struct PolyEvalSteps in the reference's layout
(include/fastPolyEvaluation.h:73-99, src/fastPolyEvaluation.cu:512-656,
:1371-1470).

What pins each array:
* Ml/Mr: the oracle's subproduct levels (closed form prod (x - x_i) per
  block), split into the two halves, the leading 1 dropped from level 8 up,
  and level 0 as (-x, 1) in 32-bit words, as the reference's
  leavesSubproductTree writes it.
* InvMl/InvMr: the defining identity rev(M) * I = 1 mod x^(2^i), checked with
  the oracle's multiply.
* fL/fR: each remainder R of a level-i node has 2^i coefficients and agrees
  with F at the node's 2^i distinct points, which fixes it (interpolation);
  level 0 holds the values, checked against Horner's rule.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _distinct_points(seed, n, p):
    rng = np.random.default_rng(seed)
    pts = np.unique(rng.integers(0, p, 2 * n))
    rng.shuffle(pts)
    assert len(pts) >= n
    return pts[:n].astype(np.int64).astype(np.int32)


def _check_tree(r, pts, k, p, orc):
    n, nh = 1 << k, 1 << (k - 1)
    lo = np.stack([-pts[:nh].astype(np.int64), np.ones(nh, np.int64)], 1)
    hi = np.stack([-pts[nh:].astype(np.int64), np.ones(nh, np.int64)], 1)
    assert np.array_equal(r["Ml"][0], lo.astype(np.int32).ravel())
    assert np.array_equal(r["Mr"][0], hi.astype(np.int32).ravel())
    for i in range(1, k):
        lev = orc.subproduct_level(pts, i, p)
        half = nh >> i
        for name, rows in (("Ml", lev[:half]), ("Mr", lev[half:])):
            if i >= 8:
                assert (rows[:, -1] == 1).all()
                rows = rows[:, :-1]
            assert np.array_equal(r[name][i], rows.ravel()), (name, i)
    assert set(r["Ml"]) == set(range(k)) and set(r["Mr"]) == set(range(k))


def _check_inverses(r, pts, k, p, orc, levels):
    nh = 1 << (k - 1)
    assert set(r["InvMl"]) == set(levels) and set(r["InvMr"]) == set(levels)
    for i in levels:
        h = 1 << i
        lev = orc.subproduct_level(pts, i, p)
        half = nh >> i
        for name, rows in (("InvMl", lev[:half]), ("InvMr", lev[half:])):
            inv = r[name][i].reshape(half, h)
            assert ((inv >= 0) & (inv < p)).all()
            for c in range(half):
                prod = orc.ntt_polymul(rows[c][::-1].copy(), inv[c], p)[:h]
                want = np.zeros(h, prod.dtype)
                want[0] = 1
                assert np.array_equal(prod, want), (name, i, c)


def _check_remainders(r, F, pts, k, p, orc, levels):
    nh = 1 << (k - 1)
    vals = orc.poly_eval(F, pts, p)
    assert set(r["fL"]) == set(levels) and set(r["fR"]) == set(levels)
    for j in levels:
        h = 1 << (j - 1)
        for name, base in (("fL", 0), ("fR", nh)):
            R = r[name][j].reshape(nh // h, h)
            assert ((R >= 0) & (R < p)).all()
            for c in range(nh // h):
                x = pts[base + c * h: base + (c + 1) * h]
                assert np.array_equal(orc.poly_eval(R[c], x, p), vals[base + c * h:
                                                                     base + (c + 1) * h]), (name, j, c)
    assert np.array_equal(np.concatenate([r["fL"][1], r["fR"][1]]), vals)


@pytest.mark.parametrize("k,p", [(3, 469762049), (9, 469762049), (11, 469762049),
                                 (10, 2013265921)])
def test_fast_evaluation_steps_flag1(cm, orc, k, p):
    pts = _distinct_points(700 + k, 1 << k, p)
    F = cm.uniform_mod(cm.SplitMix64(710 + k), 1 << k, p)
    r = cm.eval_steps("fastEvaluation", pts, p, F, flag=1)
    _check_tree(r, pts, k, p, orc)
    _check_inverses(r, pts, k, p, orc, list(range(8, k)))
    _check_remainders(r, F, pts, k, p, orc, list(range(1, k + 1)))


@pytest.mark.parametrize("k", [2, 9, 12])
def test_only_subproduct_tree_and_only_trees(cm, orc, k):
    p = 469762049
    pts = _distinct_points(720 + k, 1 << k, p)
    r = cm.eval_steps("onlySubproductTree", pts, p)
    _check_tree(r, pts, k, p, orc)
    _check_inverses(r, pts, k, p, orc, [8] if k > 8 else [])
    assert not r["fL"] and not r["fR"]
    t = cm.eval_steps("onlyTrees", pts, p)
    _check_tree(t, pts, k, p, orc)
    _check_inverses(t, pts, k, p, orc, list(range(8, k)))
    assert not t["fL"] and not t["fR"]


def test_fast_evaluation_flag0_returns_values_only(cm, orc):
    k, p = 12, 469762049
    pts = _distinct_points(731, 1 << k, p)
    F = cm.uniform_mod(cm.SplitMix64(732), 1 << k, p)
    r = cm.eval_steps("fastEvaluation", pts, p, F, flag=0)
    assert set(r["fL"]) == {1} and set(r["fR"]) == {1}
    assert not r["Ml"] and not r["InvMl"]
    vals = orc.poly_eval(F, pts, p)
    assert np.array_equal(np.concatenate([r["fL"][1], r["fR"][1]]), vals)


def test_eval_steps_errors_leave_every_pointer_null(cm):
    pts = np.arange(8, dtype=np.int32)
    assert cm.eval_steps("onlyTrees", pts[:1], 469762049) is None     # k = 0
    assert cm.eval_steps("onlySubproductTree", pts, 1000) is None     # even modulus


@pytest.mark.parametrize("p", [469762049, 2013265921])
def test_fast_evaluation_reduces_noncanonical_coefficients(cm, orc, p):
    """F given as 32-bit words at or above p is evaluated as F mod p."""
    k = 10
    pts = _distinct_points(740, 1 << k, p)
    F = cm.uniform_mod(cm.SplitMix64(741), 1 << k, p).astype(np.uint64)
    big = (F + p).astype(np.uint32).view(np.int32)        # every word in [p, 2^32)
    got = cm.fast_evaluation(big, pts, p)
    assert np.array_equal(got, orc.poly_eval(F.astype(np.int64).astype(np.int32), pts, p))
