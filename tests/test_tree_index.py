"""The library's host-side tree builder against the oracle (no GPU needed).

The tree topology (2:1 refinement, halo allocation) and the reference's
caches -- leaves, faces, active / prolongation / boundary lists per level, in
pre-order -- are built on the host by libbmg before any GPU work; they must
equal the oracle's entry for entry (grid/tree.h:556-684).
"""
import numpy as np
import pytest

import basilisk_amd as bmg
from oracle.tree import OracleTree
from tests import tree_cases


def _both(dim, refine, origin=(-0.5, -0.5, -0.5)):
    g = bmg.Tree(dim, 1.0, origin)
    o = OracleTree(dim, 1.0, origin)
    refine(g)
    refine(o)
    return g, o


def _assert_same(g, o):
    assert g.depth() == o.depth()
    assert g.nleaves() == o.nleaves()
    la, xa = g.leaves()
    lb, xb = o.leaves()
    assert np.array_equal(la, lb) and np.array_equal(xa, xb)
    for which in ("active", "prolongation", "boundary"):
        for l in range(o.depth() + 1):
            assert np.array_equal(g.cache(which, l), o.cache(which, l)), (which, l)
    assert np.array_equal(g.faces(), o.faces())
    fa, fb = g.face_centres(), o.face_centres()
    assert np.array_equal(fa[0], fb[0]) and np.array_equal(fa[1], fb[1])
    for s in range(2 * g.dim):
        assert np.array_equal(g.bc_face_centres(s), o.bc_face_centres(s))


@pytest.mark.parametrize("dim,depth", [(2, 6), (2, 9), (3, 5), (3, 6)])
def test_circle_index(dim, depth):
    g, o = _both(dim, lambda t: t.refine(tree_cases.circle_refine(depth)))
    _assert_same(g, o)


@pytest.mark.parametrize("dim,depth,seed", [(2, 8, 1), (3, 6, 2)])
def test_spheres_index(dim, depth, seed):
    def refine(t):
        tree_cases.spheres_tree(t, depth, 3, 6, seed, 0.02, 0.1)
    g, o = _both(dim, refine, origin=(0.0, 0.0, 0.0))
    _assert_same(g, o)


def test_uniform_index():
    g, o = _both(2, lambda t: t.refine(lambda x, y, z, lev: lev < 6))
    assert g.nleaves() == 4 ** 6
    _assert_same(g, o)


def test_refine_marks_validated():
    g = bmg.Tree(2)
    with pytest.raises(ValueError):
        g.refine_marked(np.ones(3, np.uint8))
    assert g.refine_marked(np.ones(1, np.uint8)) == 1
    assert g.nleaves() == 4
