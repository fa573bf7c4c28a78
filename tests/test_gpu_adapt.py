"""adapt_wavelet on the GPU tree against the tree oracle.

The wavelet estimate, the field updates (refine_bilinear into new leaves, the
children's average into coarsened cells) and the tree boundary run on the
GPU; the refinement and coarsening passes over the topology on the host.
After every call the adapted tree (depth, leaves, every cache) and the leaf
values of the user scalars, a and b equal the oracle's bit for bit, and so do the refined /
coarsened counts.  The oracle's adaptation is pinned by halo.ref
(test_adapt_oracle.py); the GPU tree reproduces its x, y, level columns too.
"""
import os

import numpy as np
import pytest

import basilisk_amd as bmg
from oracle.tree import GH, OracleTree

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(__file__)


def _same_tree(g, o):
    assert g.depth() == o.depth()
    assert g.nleaves() == o.nleaves()
    la, xa = g.leaves()
    lb, xb = o.leaves()
    assert np.array_equal(la, lb) and np.array_equal(xa, xb)
    for which in ("active", "prolongation", "boundary"):
        for l in range(o.depth() + 1):
            assert np.array_equal(g.cache(which, l), o.cache(which, l)), (which, l)


def _same_fields(g, o):
    assert np.array_equal(g.get_s(), o.get_s())
    assert np.array_equal(g.get_a(), o.get_a())
    assert np.array_equal(g.get_b(), o.get_b())


def test_halo_ref_gpu():
    print("halo.ref on the GPU tree", flush=True)
    g = bmg.Tree(2, 1.0, (-0.5, -0.5))
    o = OracleTree(2, 1.0, (-0.5, -0.5))
    for t in (g, o):
        t.refine(lambda x, y, z, lev: lev < 5)
        _, xyz = t.leaves()
        t.set_s(np.exp(-(xyz[:, 0] ** 2 + xyz[:, 1] ** 2) / 0.01))
    rounds = 0
    while True:
        sg = g.adapt_wavelet(1e-2, 5)
        so = o.adapt_wavelet(1e-2, 5)
        assert sg == so
        _same_tree(g, o)
        _same_fields(g, o)
        rounds += 1
        assert rounds < 20
        if not sg[1]:
            break
    with open(os.path.join(HERE, "golden", "ref", "halo.ref")) as f:
        ref = [ln.split()[:3] for ln in f.read().split("\n") if ln]
    got = []
    for l in range(g.depth()):
        D = 1.0 / (1 << l)
        for i, j, _ in g.cache("prolongation", l):
            got.append(["%g" % ((i - GH + 0.5) * D - 0.5), "%g" % ((j - GH + 0.5) * D - 0.5), "%d" % l])
    assert got == ref


def _bumps(xyz, centres, w):
    v = np.zeros(len(xyz))
    for c in centres:
        v += np.exp(-((xyz - c) ** 2).sum(1) / (w * w))
    return v


@pytest.mark.parametrize("dim,base,maxlevel,minlevel", [(2, 4, 9, 2), (2, 5, 7, 1), (3, 3, 6, 2)])
def test_adapt_rounds_bitexact(dim, base, maxlevel, minlevel):
    """A moving pair of bumps: every call refines ahead and coarsens behind;
    s, a (with a Dirichlet side) and b follow the mesh."""
    print("adapt rounds", dim, base, maxlevel, minlevel, flush=True)
    rng = np.random.default_rng(11 * dim + base)
    g = bmg.Tree(dim, 1.0, (0.0, 0.0, 0.0))
    o = OracleTree(dim, 1.0, (0.0, 0.0, 0.0))
    for t in (g, o):
        t.refine(lambda x, y, z, lev: lev < base)
    totf = totc = 0
    for rnd in range(5):
        c = rng.uniform(0.2, 0.8, (2, 3))[:, :dim]
        for t in (g, o):
            _, xyz = t.leaves()
            p = xyz[:, :dim]
            t.set_s(_bumps(p, c, 0.06))
            if rnd == 0:
                t.set_a(np.sin(3 * p[:, 0]) + p[:, 1] ** 2)
                t.set_b(np.cos(2 * p.sum(1)))
            t.set_bc_function(bmg.LEFT, bmg.BC_DIRICHLET, lambda x, y, z: 1.0 + y)
        sg = g.adapt_wavelet(3e-3, maxlevel, minlevel)
        so = o.adapt_wavelet(3e-3, maxlevel, minlevel)
        print("  round", rnd, sg, g.nleaves(), g.depth(), flush=True)
        assert sg == so
        totf += sg[0]
        totc += sg[1]
        _same_tree(g, o)
        _same_fields(g, o)
    assert totf > 0 and totc > 0


@pytest.mark.parametrize("dim,base,maxlevel", [(2, 5, 8), (3, 3, 6)])
def test_poisson_after_adapt_bitexact(dim, base, maxlevel):
    """The solver runs on the adapted tree: poisson() with Dirichlet BCs,
    alpha from a function and a lambda constant equals the oracle's."""
    print("poisson after adapt", dim, base, maxlevel, flush=True)
    g = bmg.Tree(dim, 1.0, (0.0, 0.0, 0.0), smoother="rbgs")
    o = OracleTree(dim, 1.0, (0.0, 0.0, 0.0))
    o.set_smoother("rbgs")
    for t in (g, o):
        t.refine(lambda x, y, z, lev: lev < base)
    for rnd in range(3):
        for t in (g, o):
            _, xyz = t.leaves()
            t.set_s(_bumps(xyz[:, :dim], [np.full(dim, 0.35 + 0.1 * rnd)], 0.05))
        assert g.adapt_wavelet(2e-3, maxlevel) == o.adapt_wavelet(2e-3, maxlevel)
    _same_tree(g, o)
    for t in (g, o):
        _, xyz = t.leaves()
        p = xyz[:, :dim]
        t.set_b(np.cos(3 * p[:, 0]) * np.sin(2 * p[:, 1]))
        t.set_a(np.zeros(len(p)))
        for side in range(2 * dim):
            t.set_bc_function(side, bmg.BC_DIRICHLET, lambda x, y, z: x - y)
        t.set_alpha_function(lambda x, y, z, d: 1.0 + 0.5 * x + 0.25 * d)
        t.set_lambda(-2.0)
        t.set_niter(30)
    sg = g.poisson(tolerance=1e-7)
    so = o.poisson(tolerance=1e-7)
    for k in ("i", "resb", "resa", "nrelax", "minlevel"):
        assert sg[k] == so[k], k
    assert np.array_equal(g.get_a(), o.get_a())


def test_adapt_errors():
    print("adapt errors", flush=True)
    g = bmg.Tree(2, 1.0, (0.0, 0.0, 0.0))
    g.refine(lambda x, y, z, lev: lev < 3)
    with pytest.raises(bmg.BMGError):
        g.adapt_wavelet(1e-2, 5)  # s not set
    g.set_s(np.zeros(g.nleaves()))
    with pytest.raises(bmg.BMGError):
        g.adapt_wavelet(1e-2, 99)
    with pytest.raises(bmg.BMGError):
        g.adapt_wavelet(-1.0, 5)
    # a constant field coarsens everything down to minlevel
    nf, nc = g.adapt_wavelet(1e-2, 5, 2)
    assert nf == 0 and nc > 0
    assert g.depth() == 2 and g.nleaves() == 16


def test_coarsen_ref_gpu():
    """test/coarsen.c on the GPU tree (init_grid (64), refine (level < 8),
    unrefine (level > 5)): depth 6 and 4096 leaves as in coarsen.ref, and
    the oracle's caches."""
    print("case coarsen.ref", flush=True)
    g = bmg.Tree(2, 1.0, (0.0, 0.0, 0.0))
    o = OracleTree(2, 1.0, (0.0, 0.0, 0.0))
    for t in (g, o):
        t.refine(lambda x, y, z, lev: lev < 8)
        _, xyz = t.leaves()
        t.set_a(xyz[:, 0] - xyz[:, 1])
        t.set_s(np.zeros(len(xyz)))
    ng = g.unrefine(lambda x, y, z, lev: lev > 5)
    no = o.unrefine(lambda x, y, z, lev: lev > 5)
    assert ng == no
    assert (g.depth(), g.nleaves()) == (6, 4096)
    _same_tree(g, o)
    _same_fields(g, o)


@pytest.mark.parametrize("dim", [2, 3])
def test_unrefine_region_bitexact(dim):
    print("case unrefine", dim, flush=True)
    g = bmg.Tree(dim, 1.0, (0.0, 0.0, 0.0))
    o = OracleTree(dim, 1.0, (0.0, 0.0, 0.0))
    deep = 7 if dim == 2 else 5
    for t in (g, o):
        t.refine(lambda x, y, z, lev: (lev < 3) | ((lev < deep) & (x < 0.6)))
        _, xyz = t.leaves()
        t.set_a(np.sin(4 * xyz[:, 0]) + xyz[:, 1])
        t.set_b(np.cos(3 * xyz[:, 1]))
        t.set_s(xyz[:, 0] ** 2)
    cond = lambda x, y, z, lev: (x > 0.3) & (lev > 2)
    ng, no = g.unrefine(cond), o.unrefine(cond)
    assert ng == no and ng > 0
    _same_tree(g, o)
    _same_fields(g, o)
    # and the solver runs on the coarsened tree
    for t in (g, o):
        t.set_alpha_const(1.0)
        t.set_lambda(-1.0)
    o.set_smoother("rbgs")
    sg, so = g.poisson(tolerance=1e-7), o.poisson(tolerance=1e-7)
    for k in ("i", "resb", "resa", "nrelax"):
        assert sg[k] == so[k], k
    assert np.array_equal(g.get_a(), o.get_a())


@pytest.mark.parametrize("dim,base,maxlevel,minlevel", [(2, 4, 8, 2), (3, 3, 6, 2)])
def test_adapt_list_bitexact(dim, base, maxlevel, minlevel):
    """adapt_wavelet with two estimated user scalars (struct Adapt's slist
    and max[]) and a third user scalar carried along, then unrefine(): the
    tree, the counts and every field (user scalars 0, 1, 3, a and b) equal
    the oracle's, whose list semantics test_adapt_oracle.py pins."""
    print("adapt list", dim, base, maxlevel, minlevel, flush=True)
    rng = np.random.default_rng(7 * dim + base)
    g = bmg.Tree(dim, 1.0, (0.0, 0.0, 0.0))
    o = OracleTree(dim, 1.0, (0.0, 0.0, 0.0))
    for t in (g, o):
        t.refine(lambda x, y, z, lev: lev < base)
    tot = np.zeros(2, int)
    for rnd in range(4):
        c0, c1 = rng.uniform(0.2, 0.8, (2, 3))[:, :dim]
        for t in (g, o):
            _, xyz = t.leaves()
            p = xyz[:, :dim]
            t.set_s(_bumps(p, [c0], 0.06), k=0)
            t.set_s(_bumps(p, [c1], 0.05), k=1)
            if rnd == 0:
                t.set_s(np.sin(4 * p[:, 0]) * np.cos(3 * p[:, 1]), k=3)
                t.set_a(np.sin(3 * p[:, 0]) + p[:, 1] ** 2)
                t.set_b(np.cos(2 * p.sum(1)))
        sg = g.adapt_wavelet([3e-3, 2e-3], maxlevel, minlevel, scalars=[0, 1])
        so = o.adapt_wavelet([3e-3, 2e-3], maxlevel, minlevel, scalars=[0, 1])
        print("  round", rnd, sg, g.nleaves(), flush=True)
        assert sg == so
        tot += sg
        _same_tree(g, o)
        for k in (0, 1, 3):
            assert np.array_equal(g.get_s(k=k), o.get_s(k)), k
        assert np.array_equal(g.get_a(), o.get_a())
        assert np.array_equal(g.get_b(), o.get_b())
    assert tot[0] > 0 and tot[1] > 0
    cond = lambda x, y, z, lev: (lev > maxlevel - 2) & (x < 0.5)
    assert g.unrefine(cond) == o.unrefine(cond)
    _same_tree(g, o)
    for k in (0, 1, 3):
        assert np.array_equal(g.get_s(k=k), o.get_s(k)), k
    # a scalar not set cannot be estimated; indices outside the list fail
    with pytest.raises(bmg.BMGError):
        g.adapt_wavelet([1e-2], maxlevel, scalars=[2])
    with pytest.raises(ValueError):
        g.set_s(np.zeros(g.nleaves()), k=bmg.USER_SCALARS)


@pytest.mark.parametrize("dim,base,maxlevel", [(2, 4, 7), (3, 3, 5)])
def test_refine_carries_fields_bitexact(dim, base, maxlevel):
    """refine (cond) (grid/tree-common.h:308-326) on a tree that holds
    fields: each pass applies boundary (all) and refine_cell (point, all),
    so the user scalars, a and b get refine_bilinear in the new leaves (2:1
    cascades included); tree and fields equal the oracle's."""
    print("refine carries fields", dim, base, maxlevel, flush=True)
    g = bmg.Tree(dim, 1.0, (0.0, 0.0, 0.0))
    o = OracleTree(dim, 1.0, (0.0, 0.0, 0.0))
    for t in (g, o):
        t.refine(lambda x, y, z, lev: lev < base)
        _, xyz = t.leaves()
        p = xyz[:, :dim]
        t.set_a(np.sin(3 * p[:, 0]) + p[:, 1] ** 2)
        t.set_b(np.cos(2 * p.sum(1)))
        t.set_s(_bumps(p, [np.full(dim, 0.4)], 0.1), k=0)
        t.set_s(p[:, 0] - 2 * p[:, 1], k=2)
    for rnd, (c, r) in enumerate(((0.3, 0.15), (0.62, 0.1))):
        cond = lambda x, y, z, lev: ((x - c) ** 2 + (y - c) ** 2 < r * r) & (lev < maxlevel - rnd)
        for t in (g, o):
            t.refine(cond)
        _same_tree(g, o)
        for k in (0, 2):
            assert np.array_equal(g.get_s(k=k), o.get_s(k)), k
        assert np.array_equal(g.get_a(), o.get_a())
        assert np.array_equal(g.get_b(), o.get_b())
