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


@pytest.mark.parametrize("dim,depth", [(2, 8), (3, 6)])
def test_shell_refine_equals_direct_condition(dim, depth):
    """The fast refinement condition of the adaptive workloads (cells
    enumerated per sphere) builds the same tree as the direct per-leaf test
    |dist - r| < band h + size sqrt(dim)."""
    from basilisk_amd import configs
    rng = np.random.default_rng(9)
    c = rng.uniform(0.1, 0.9, (7, dim))
    r = rng.uniform(0.02, 0.1, 7)
    h = 1.0 / (1 << depth)

    def direct(x, y, z, lev):
        p = np.stack([x, y, z][:dim], axis=1)
        size = 1.0 / (1 << lev.astype(np.int64))
        near = np.zeros(len(x), bool)
        for s in range(len(r)):
            near |= np.abs(np.sqrt(((p - c[s]) ** 2).sum(1)) - r[s]) < 2.0 * h + size * np.sqrt(dim)
        return (lev < 3) | ((lev < depth) & near)
    a = OracleTree(dim, 1.0, (0.0, 0.0, 0.0))
    b = OracleTree(dim, 1.0, (0.0, 0.0, 0.0))
    a.refine(direct)
    b.refine(configs.shell_refine(dim, depth, 3, c, r, 2.0))
    assert a.nleaves() == b.nleaves()
    assert np.array_equal(a.leaves()[1], b.leaves()[1])


def test_faces_ref():
    """test/faces.c: init_grid (4), refine (level == 2 && x*x + y*y <
    sq(0.25)) on [-0.5, 0.5]^2, then the face centres of foreach_face(x),
    foreach_face(y) and foreach_face(): faces.ref line for line, from the
    library's face cache and from the oracle's."""
    import os
    ref_path = os.path.join(os.path.dirname(__file__), "golden", "ref", "faces.ref")
    with open(ref_path) as f:
        ref = [ln for ln in f.read().split("\n") if ln]
    for t in (bmg.Tree(2, 1.0, (-0.5, -0.5)), OracleTree(2, 1.0, (-0.5, -0.5))):
        t.refine(lambda x, y, z, lev: lev < 2)
        t.refine(lambda x, y, z, lev: (lev == 2) & (x * x + y * y < 0.25 ** 2))
        xyz, dd = t.face_centres()
        line = lambda p: "%g %g" % (p[0], p[1])
        got = ([line(p) for p, d in zip(xyz, dd) if d == 0] + [line(p) for p, d in zip(xyz, dd) if d == 1]
               + [line(p) for p in xyz])
        assert got == ref
