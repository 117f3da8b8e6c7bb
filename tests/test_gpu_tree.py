"""GPU parity of the subproduct tree against the oracle (exact residues).

Small trees are compared level by level with the oracle's products; the
2^16-point tree of the benchmark workload is checked at full size by the
closed form: every node vanishes at its own points and agrees with
prod (r - x_j) at random r.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
P1 = 469762049


def test_tree_levels_exact(cm, orc):
    rng = np.random.default_rng(1)
    for k in (1, 3, 6, 10):
        pts = rng.integers(0, P1, 1 << k).astype(np.int32)
        levels = cm.subproduct_tree(pts, P1)
        assert len(levels) == k + 1
        for i in range(k + 1):
            assert (levels[i] == orc.subproduct_level(pts, i, P1)).all(), (k, i)


def horner(c, r, p):
    acc = 0
    for x in reversed([int(v) for v in c]):
        acc = (acc * r + x) % p
    return acc


def test_workload4_tree_full_size(cm, orc):
    w = cm.workload4_tree()
    pts, p = w["points"], w["p"]
    k = 16
    levels = cm.subproduct_tree(pts, p)
    root = levels[k][0]
    assert root[-1] == 1 and len(root) == (1 << k) + 1
    # the root vanishes at sampled points, and matches the product at random r
    for j in (0, 1, 12345, (1 << k) - 1):
        assert horner(root, int(pts[j]), p) == 0
    for r in (5, 1234567):
        want = 1
        for x in pts.tolist():
            want = want * (r - x) % p
        assert horner(root, r, p) == want
    # every level against the oracle at a few nodes of each level
    for i in range(1, k + 1):
        for j in (0, levels[i].shape[0] - 1):
            node = levels[i][j]
            block = pts[j << i:(j + 1) << i].tolist()
            r = 987654321 % p
            want = 1
            for x in block:
                want = want * (r - x) % p
            assert horner(node, r, p) == want, (i, j)
    # level 10 exactly against the oracle
    assert (levels[10] == orc.subproduct_level(pts, 10, p)).all()
