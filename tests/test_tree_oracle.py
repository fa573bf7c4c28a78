"""The adaptive-tree CPU oracle against the reference's known-answer logs and
the tree invariants of grid/tree.h:748-793 (tree_check) and check_two_one.
"""
import numpy as np
import pytest

from oracle.oracle import OracleGrid
from oracle.tree import OracleTree, GH
from tests import cases, tree_cases


def _ref_block(name, depth):
    return [l for l in cases.read_ref(name)
            if l.startswith("residual %d " % depth) or l.startswith("max error %d " % depth)]


@pytest.mark.parametrize("depth", [7, 8, 9, 10])
def test_circle_ref(depth):
    """test/circle.c (2D): every residual (%.1g) and the max error (%g)."""
    t = OracleTree(2, 1.0, (-0.5, -0.5, -0.5))
    assert tree_cases.circle_lines(t, depth) == _ref_block("circle.ref", depth)


@pytest.mark.parametrize("depth", [7, 8, 9, 10])
def test_circle_variable_ref(depth):
    """test/circle-variable.c: alpha = x + y + 2 on faces, restricted by
    face averages and prolongated into halos by refine_face_x."""
    t = OracleTree(2, 1.0, (-0.5, -0.5, -0.5))
    assert tree_cases.circle_lines(t, depth, variable=True) == _ref_block("circle-variable.ref", depth)


@pytest.mark.parametrize("depth", [5, 6, 7])
def test_circle_3d_ref(depth):
    """test/circle.c on an octree (circle.3D.ref)."""
    t = OracleTree(3, 1.0, (-0.5, -0.5, -0.5))
    assert tree_cases.circle_lines(t, depth) == _ref_block("circle.3D.ref", depth)


@pytest.mark.parametrize("smoother", ["gs", "jacobi", "rbgs"])
def test_full_tree_matches_uniform_oracle(smoother):
    """A tree refined to one level is `full`: the tree oracle then equals the
    uniform-grid oracle in tree mode bit for bit (two restatements of
    test/poisson.c, poisson.ref), for each smoother, so the Fourier and
    fixed-point pins of the uniform oracle's sweeps carry over."""
    depth = 7
    t = OracleTree(2, 1.0, (-0.5, -0.5, -0.5))
    t.set_smoother(smoother)
    t.refine(lambda x, y, z, lev: lev < depth)
    assert t.nleaves() == 4 ** depth

    def f(x, y, z):
        return np.cos(3. * cases.PI * x) * np.cos(3. * cases.PI * y)
    lev, xyz = t.leaves()
    b = -18. * cases.PI ** 2 * f(xyz[:, 0], xyz[:, 1], 0)
    t.set_b(b)
    t.set_a(np.zeros_like(b))
    for s in range(4):
        t.set_bc_function(s, cases.BC_DIRICHLET, f)
    g = OracleGrid(2, depth, origin=(-0.5, -0.5, -0.5))
    g.set_tree()
    g.set_smoother(smoother)
    g.set_bc_function(cases.BC_DIRICHLET, f)
    X, Y, _ = cases.mesh(g)
    g.set_b(-18. * cases.PI ** 2 * f(X, Y, 0).reshape(g.shape()))
    g.set_a(np.zeros(g.shape()))
    assert t.residual() == g.residual()
    for _ in range(3):
        t.mg_cycle(4, 0, depth)
        g.mg_cycle(4, 0)
        assert t.residual() == g.residual()
    # leaf order (Morton) -> grid index
    n = 1 << depth
    D = 1.0 / n
    i = np.rint((xyz[:, 0] + 0.5) / D - 0.5).astype(int)
    j = np.rint((xyz[:, 1] + 0.5) / D - 0.5).astype(int)
    assert np.array_equal(t.get_a(), g.get_a()[i, j])


def _tree_check(t):
    """tree_check (grid/tree.h:748-793) and check_two_one (1587-1620)."""
    dim = t.dim
    depth = t.depth()
    metas = [t.level_meta(l) for l in range(depth + 1)]
    ring = [np.array(o) for o in np.ndindex(*(3,) * dim)]
    for l, (alloc, flags, nb, pid) in enumerate(metas):
        idx = np.argwhere(alloc)
        for p in idx:
            p = tuple(p)
            leaf, active = flags[p] & 2, flags[p] & 1
            if leaf:
                assert pid[p] >= 0
            if pid[p] >= 0:
                # local cells are active or prolongation
                assert active or (not leaf and nb[p] == 0)
            # refined neighbours in the 3^d ring
            cnt = 0
            for o in ring:
                q = tuple(np.array(p) + o - 1)
                if min(q) < 0 or max(q) >= alloc.shape[0] or not alloc[q]:
                    continue
                if not (flags[q] & 2) and nb[q] and pid[q] >= 0:
                    cnt += 1
            assert nb[p] == cnt, (l, p)
            c = tuple(2 * x - GH for x in p)
            if nb[p] == 0 and l < depth and min(c) >= 0 and max(c) < metas[l + 1][0].shape[0]:
                assert not metas[l + 1][0][c]
    # 2:1: every leaf's parent-level 3^d ring is active (check_two_one)
    for l in range(1, depth + 1):
        alloc, flags, nb, pid = metas[l]
        palloc, pflags, _, ppid = metas[l - 1]
        for p in np.argwhere(alloc & ((flags & 2) > 0)):
            par = (p + GH) // 2
            for o in ring:
                q = tuple(par + o - 1)
                if ppid[q] >= 0:
                    assert pflags[q] & 1, (l, tuple(p))


@pytest.mark.parametrize("dim,depth", [(2, 7), (3, 5)])
def test_tree_invariants(dim, depth):
    t = OracleTree(dim, 1.0, (-0.5, -0.5, -0.5))
    t.refine(tree_cases.circle_refine(depth))
    _tree_check(t)
    # every leaf is in exactly one level list, active lists are Morton ordered
    lev, _ = t.leaves()
    assert len(lev) == t.nleaves()
    total = sum(len(t.cache("active", l)) for l in range(t.depth() + 1))
    parents = total - t.nleaves()
    assert parents * (1 << dim) == total - 1


def test_refine_order_independent():
    """The fixed point of refine() does not depend on the order in which the
    leaves are refined (each pass refines the marked leaves in cache order;
    refining them in reverse order gives the same caches)."""
    a = OracleTree(2, 1.0, (-0.5, -0.5, -0.5))
    b = OracleTree(2, 1.0, (-0.5, -0.5, -0.5))
    cond = tree_cases.circle_refine(8)
    a.refine(cond)
    while True:
        lev, xyz = b.leaves()
        mark = np.asarray(cond(xyz[:, 0], xyz[:, 1], xyz[:, 2], lev), dtype=np.uint8)
        if not mark.any():
            break
        # one leaf at a time, last first
        for k in np.nonzero(mark)[0][::-1]:
            m = np.zeros_like(mark)
            m[k] = 1
            b.lib.ot_refine_marked(b.t, m.ctypes.data_as(__import__("ctypes").c_void_p))
            lev2, _ = b.leaves()
            if len(lev2) != len(lev):
                break
    for which in ("active", "prolongation", "boundary"):
        for l in range(a.depth() + 1):
            assert np.array_equal(a.cache(which, l), b.cache(which, l))
    assert np.array_equal(a.faces(), b.faces())


def test_restriction_c_face_vector_invariant():
    """test/restriction.c (assert-only): every face of a non-trivial tree set
    to 1, then the restriction of the face vector (face_average below the
    leaves, refine_face_x into the halos, halo_face at resolution changes):
    every active cell of every level sees 1 on its four faces."""
    from oracle.tree import GH, OracleTree, T_G
    import numpy as np
    depth = 4
    t = OracleTree(2, 1.0, (-0.7, -0.7))
    t.refine(lambda x, y, z, lev: (lev < depth - 2)
             | (lev <= depth * (1.0 - np.sqrt((x - 0.1) ** 2 + (y - 0.1) ** 2))))
    assert t.depth() > 2 and t.nleaves() < 4 ** t.depth()
    t.set_alpha_function(lambda x, y, z, d: np.ones(len(x)))
    t.restriction_alpha_lambda()
    checked = 0
    for l in range(t.depth() + 1):
        n = (1 << l) + 2 * GH
        ax = np.ctypeslib.as_array(t.lib.ot_level_field(t.t, l, T_G + 2), shape=(n, n))
        ay = np.ctypeslib.as_array(t.lib.ot_level_field(t.t, l, T_G + 3), shape=(n, n))
        for i, j, _ in t.cache("active", l):
            assert ax[i, j] == 1.0 and ax[i + 1, j] == 1.0, (l, i, j)
            assert ay[i, j] == 1.0 and ay[i, j + 1] == 1.0, (l, i, j)
            checked += 1
    assert checked > 50
