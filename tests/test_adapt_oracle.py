"""The tree oracle's adapt_wavelet against the reference (no GPU needed).

- halo.ref (src/test/halo.c): a 32^2 Gaussian coarsened by
  ``while (adapt_wavelet ({h}, {1e-2}, 5).nc);``, then every cell of the
  prolongation caches of levels 0..depth-1 printed as ``x y level
  cell.neighbors``.  The fixture pins the coarsening pass, decrement_neighbors
  and the caches of the adapted tree, line for line.
- tree_check (grid/tree-common.h:747-792) after rounds of refinement and
  coarsening: every cell's neighbour count equals the number of refined cells
  in its 3^dim ring, cells without refined neighbours have no children, leaves
  are interior, local cells are active or prolongation, every active cell
  descends from active non-leaf cells.
- A linear field carried through refinement and coarsening stays exact:
  bilinear prolongation and the children's average reproduce linear data.
"""
import os

import numpy as np
import pytest

from oracle.tree import GH, OracleTree

HERE = os.path.dirname(__file__)


def _halo_lines(t):
    out = []
    for l in range(t.depth()):
        nb = t.level_meta(l)[2]
        D = 1.0 / (1 << l)
        for i, j, _ in t.cache("prolongation", l):
            x = (i - GH + 0.5) * D - 0.5
            y = (j - GH + 0.5) * D - 0.5
            out.append("%g %g %d %d halo" % (x, y, l, nb[i, j]))
    return out


def test_halo_ref():
    t = OracleTree(2, 1.0, (-0.5, -0.5))
    t.refine(lambda x, y, z, lev: lev < 5)
    _, xyz = t.leaves()
    R0 = 0.1
    t.set_s(np.exp(-(xyz[:, 0] ** 2 + xyz[:, 1] ** 2) / (R0 * R0)))
    rounds = 0
    while t.adapt_wavelet(1e-2, 5)[1]:
        rounds += 1
        assert rounds < 20
    with open(os.path.join(HERE, "golden", "ref", "halo.ref")) as f:
        ref = [ln for ln in f.read().split("\n") if ln]
    assert _halo_lines(t) == ref


def _ring_sum(a, dim):
    """Sum over the 3^dim ring (the cell included) of a padded array."""
    p = np.pad(a.astype(np.int64), 1)
    s = np.zeros(a.shape, np.int64)
    n = a.shape[0]
    offs = [(x, y, z) for x in (-1, 0, 1) for y in (-1, 0, 1) for z in ((-1, 0, 1) if dim == 3 else (0,))]
    for o in offs:
        sl = tuple(slice(1 + o[d], 1 + o[d] + n) for d in range(dim))
        s += p[sl]
    return s


def tree_check(t):
    dim = t.dim
    meta = [t.level_meta(l) for l in range(t.depth() + 1)]
    for l, (alloc, flags, nb, pid) in enumerate(meta):
        al = alloc.astype(bool)
        leaf = (flags & 2).astype(bool) & al
        active = (flags & 1).astype(bool) & al
        assert (pid[leaf] >= 0).all(), "boundary leaf"
        refined = al & ~leaf & (nb > 0) & (pid >= 0)
        prolong = al & ~leaf & (nb == 0) & (pid >= 0)
        local = al & (pid == 0)
        assert (active[local] | prolong[local]).all(), "local cell neither active nor prolongation"
        cnt = _ring_sum(refined, dim)
        assert (nb[al] == cnt[al]).all(), "neighbour count at level %d" % l
        # adapt_wavelet's user flags are all cleared again
        assert not (flags[al] & (8 | 16 | 32 | 64)).any()
        if l < t.depth():
            ch_alloc = meta[l + 1][0].astype(bool)
            # a cell's children: (2i-GH+a, 2j-GH+b) -> parent of child c is (c+GH)//2
            idx = np.indices(ch_alloc.shape).reshape(dim, -1).T
            par = (idx + GH) // 2
            has = np.zeros(alloc.shape, bool)
            has[tuple(par[ch_alloc.reshape(-1)].T)] = True
            assert not (has & (nb == 0)).any(), "children without refined neighbours"
            # active children descend from active non-leaf parents
            ch_active = (meta[l + 1][1] & 1).astype(bool) & ch_alloc
            pa = par[ch_active.reshape(-1)]
            assert (active & ~leaf)[tuple(pa.T)].all()
    # leaves tile the box
    lev, _ = t.leaves()
    assert abs(np.sum(0.5 ** (dim * lev.astype(float))) - 1.0) < 1e-12


def _gauss_field(xyz, centres, w):
    v = np.zeros(len(xyz))
    for c in centres:
        v += np.exp(-((xyz - c) ** 2).sum(1) / (w * w))
    return v


@pytest.mark.parametrize("dim,base,maxlevel", [(2, 4, 8), (3, 3, 6)])
def test_adapt_rounds_keep_tree_consistent(dim, base, maxlevel):
    rng = np.random.default_rng(5 + dim)
    t = OracleTree(dim, 1.0, (0.0, 0.0, 0.0))
    t.refine(lambda x, y, z, lev: lev < base)
    totf = totc = 0
    for rnd in range(4):
        # a moving bump: refinement near it, coarsening where it left
        c = rng.uniform(0.25, 0.75, (2, 3))[:, :dim]
        _, xyz = t.leaves()
        t.set_s(_gauss_field(xyz[:, :dim], c, 0.08))
        nf, nc = t.adapt_wavelet(5e-3, maxlevel, 2)
        totf += nf
        totc += nc
        tree_check(t)
        assert t.depth() <= maxlevel
    assert totf > 0 and totc > 0


@pytest.mark.parametrize("dim", [2, 3])
def test_linear_field_carried_exactly(dim):
    """a = 1 + 2x - 3y (+ 0.5z) is reproduced exactly by refine_bilinear and
    by restriction_average, so after adapt_wavelet every leaf of a at least
    two parent cells away from the walls equals the linear function at its
    centre (the symmetry ghosts near the walls are not linear)."""
    coef = np.array([2.0, -3.0, 0.5])[:dim]
    t = OracleTree(dim, 1.0, (0.0, 0.0, 0.0))
    t.refine(lambda x, y, z, lev: lev < (4 if dim == 2 else 3))
    totf = totc = 0
    for rnd in range(3):
        _, xyz = t.leaves()
        t.set_a(1.0 + xyz[:, :dim] @ coef)
        t.set_s(_gauss_field(xyz[:, :dim], [np.full(dim, 0.3 + 0.2 * rnd)], 0.07))
        nf, nc = t.adapt_wavelet(2e-3, 7 if dim == 2 else 5)
        totf += nf
        totc += nc
        lev, xyz = t.leaves()
        a = t.get_a()
        h = 0.5 ** lev.astype(float)
        ok = np.all((xyz[:, :dim] > 4 * h[:, None]) & (xyz[:, :dim] < 1 - 4 * h[:, None]), axis=1)
        exact = 1.0 + xyz[:, :dim] @ coef
        assert ok.sum() > 100
        np.testing.assert_allclose(a[ok], exact[ok], rtol=0, atol=1e-12)
    assert totf > 0 and totc > 0


def test_coarsen_ref():
    """test/coarsen.c: init_grid (64), refine (level < 8), unrefine (level >
    5), tree_check (): coarsen.ref's last line is `depth: 6 leaves: 4096`
    (its `mem` figures are the reference's allocator and not compared)."""
    t = OracleTree(2, 1.0, (0.0, 0.0, 0.0))
    t.refine(lambda x, y, z, lev: lev < 6)
    assert t.depth() == 6
    t.refine(lambda x, y, z, lev: lev < 8)
    assert t.depth() == 8
    nc = t.unrefine(lambda x, y, z, lev: lev > 5)
    tree_check(t)
    with open(os.path.join(HERE, "golden", "ref", "coarsen.ref")) as f:
        last = [ln for ln in f.read().split("\n") if ln][-1]
    assert last.startswith("depth: %d leaves: %d " % (t.depth(), t.nleaves()))
    assert nc == 4 ** 6 + 4 ** 7


@pytest.mark.parametrize("dim", [2, 3])
def test_unrefine_region_keeps_tree_consistent(dim):
    """unrefine of a region of an adaptive tree: the 2:1 rule refuses some
    cells, the tree stays consistent and a linear field stays exact."""
    t = OracleTree(dim, 1.0, (0.0, 0.0, 0.0))
    t.refine(lambda x, y, z, lev: (lev < 3) | ((lev < (7 if dim == 2 else 5)) & (x < 0.6)))
    _, xyz = t.leaves()
    coef = np.array([1.5, -2.0, 0.75])[:dim]
    t.set_a(0.5 + xyz[:, :dim] @ coef)
    nc = t.unrefine(lambda x, y, z, lev: (x > 0.3) & (lev > 2))
    assert nc > 0
    tree_check(t)
    lev, xyz = t.leaves()
    a = t.get_a()
    h = 0.5 ** lev.astype(float)
    ok = np.all((xyz[:, :dim] > 4 * h[:, None]) & (xyz[:, :dim] < 1 - 4 * h[:, None]), axis=1)
    np.testing.assert_allclose(a[ok], 0.5 + xyz[ok][:, :dim] @ coef, rtol=0, atol=1e-12)


# -- adapt_wavelet with a list of scalars (struct Adapt's slist and max[])

def _leafset(t):
    lev, xyz = t.leaves()
    return lev, xyz


def _level_at(t, pts):
    """Level of the leaf of t containing each point."""
    lev, xyz = t.leaves()
    dim = t.dim
    keys = {}
    for l, x in zip(lev, xyz[:, :dim]):
        keys[(int(l),) + tuple(np.floor(x * (1 << int(l))).astype(int))] = int(l)
    out = np.zeros(len(pts), int)
    for n, p in enumerate(pts[:, :dim]):
        for l in range(int(lev.max()), -1, -1):
            k = (l,) + tuple(np.floor(p * (1 << l)).astype(int))
            if k in keys:
                out[n] = l
                break
        else:
            raise AssertionError("no leaf holds %s" % p)
    return out


def _two_bumps_tree(dim, base, fields, plant=0):
    """A tree refined to `base` with the user scalars `fields` (k -> function
    of the leaf centres) set."""
    t = OracleTree(dim, 1.0, (0.0, 0.0, 0.0), plant=plant)
    t.refine(lambda x, y, z, lev: lev < base)
    _, xyz = t.leaves()
    for k, f in fields.items():
        t.set_s(f(xyz[:, :dim]), k)
    return t


def _union_levels(dim, plant=0):
    """Levels of one call with {s0, s1} and of {s0} and {s1} alone (from
    the same tree, coarsening off) at every leaf centre of the three."""
    base, maxlevel = (4, 7) if dim == 2 else (3, 5)
    c0, c1 = np.full(dim, 0.35), np.full(dim, 0.62)
    f0 = lambda p: _gauss_field(p, [c0], 0.07)
    f1 = lambda p: _gauss_field(p, [c1], 0.05)
    both = _two_bumps_tree(dim, base, {0: f0, 1: f1}, plant)
    only0 = _two_bumps_tree(dim, base, {0: f0, 1: f1})
    only1 = _two_bumps_tree(dim, base, {0: f0, 1: f1})
    mx = 4e-3
    nf = both.adapt_wavelet([mx, mx], maxlevel, maxlevel + 1, scalars=[0, 1])[0]
    n0 = only0.adapt_wavelet(mx, maxlevel, maxlevel + 1, scalars=[0])[0]
    n1 = only1.adapt_wavelet(mx, maxlevel, maxlevel + 1, scalars=[1])[0]
    assert n0 > 0 and n1 > 0
    if not plant:
        assert nf >= max(n0, n1)
        tree_check(both)
    pts = np.concatenate([t.leaves()[1] for t in (both, only0, only1)])
    return _level_at(both, pts), _level_at(only0, pts), _level_at(only1, pts)


@pytest.mark.parametrize("dim", [2, 3])
def test_adapt_list_refines_the_union(dim):
    """With coarsening off (minlevel above maxlevel), one call with {s0, s1}
    leaves every point at the finer of the levels {s0} and {s1} reach alone:
    each scalar can only add the too_coarse flag."""
    lb, l0, l1 = _union_levels(dim)
    assert np.array_equal(lb, np.maximum(l0, l1))


@pytest.mark.parametrize("plant", [11, 12])
def test_adapt_list_union_catches_planted_mistakes(plant):
    """Flags cleared before each scalar (only the last one decides), or only
    the first scalar estimated: the union check fails."""
    lb, l0, l1 = _union_levels(2, plant)
    assert not np.array_equal(lb, np.maximum(l0, l1))


@pytest.mark.parametrize("dim", [2, 3])
def test_adapt_list_order_and_null_scalar(dim):
    """Refinement and coarsening together: {s0, s1} and {s1, s0} give the
    same tree and fields (refine when any scalar asks, just-fine when any
    says so, coarsen only when all allow it), and adding a scalar that is
    zero everywhere (its estimate always allows coarsening) changes
    nothing."""
    base, maxlevel = (5, 7) if dim == 2 else (3, 5)
    c0, c1 = np.full(dim, 0.3), np.full(dim, 0.66)
    f0 = lambda p: _gauss_field(p, [c0], 0.08)
    f1 = lambda p: _gauss_field(p, [c1], 0.06)
    zero = lambda p: np.zeros(len(p))
    a = _two_bumps_tree(dim, base, {0: f0, 1: f1, 2: zero})
    b = _two_bumps_tree(dim, base, {0: f0, 1: f1, 2: zero})
    c = _two_bumps_tree(dim, base, {0: f0, 1: f1, 2: zero})
    d = _two_bumps_tree(dim, base, {0: f0, 1: f1, 2: zero})
    for rnd in range(2):
        sa = a.adapt_wavelet([3e-3, 2e-3], maxlevel, 2, scalars=[0, 1])
        sb = b.adapt_wavelet([2e-3, 3e-3], maxlevel, 2, scalars=[1, 0])
        sc = c.adapt_wavelet([3e-3, 1e-3, 2e-3], maxlevel, 2, scalars=[0, 2, 1])
        assert sa == sb == sc and sa[0] > 0
        for t in (b, c):
            assert np.array_equal(t.leaves()[0], a.leaves()[0])
            assert np.array_equal(t.leaves()[1], a.leaves()[1])
            for k in range(3):
                assert np.array_equal(t.get_s(k), a.get_s(k))
    # {s0} alone (s1 ignored) gives another tree
    for rnd in range(2):
        d.adapt_wavelet(3e-3, maxlevel, 2, scalars=[0])
    assert d.nleaves() != a.nleaves()
    tree_check(a)


@pytest.mark.parametrize("dim", [2, 3])
def test_every_user_scalar_follows_the_mesh(dim):
    """A user scalar that is not estimated is carried like a: a linear
    function stays exact away from the walls (refine_bilinear and the
    children's average reproduce linear data) in user scalars 1..3."""
    coef = np.array([2.0, -3.0, 0.5])[:dim]
    t = OracleTree(dim, 1.0, (0.0, 0.0, 0.0))
    t.refine(lambda x, y, z, lev: lev < (4 if dim == 2 else 3))
    for rnd in range(3):
        _, xyz = t.leaves()
        p = xyz[:, :dim]
        if rnd == 0:
            for k in (1, 2, 3):
                t.set_s(k + p @ coef * k, k)
        t.set_s(_gauss_field(p, [np.full(dim, 0.3 + 0.2 * rnd)], 0.07), 0)
        nf, nc = t.adapt_wavelet(2e-3, 7 if dim == 2 else 5)
        lev, xyz = t.leaves()
        h = 0.5 ** lev.astype(float)
        ok = np.all((xyz[:, :dim] > 4 * h[:, None]) & (xyz[:, :dim] < 1 - 4 * h[:, None]), axis=1)
        assert ok.sum() > 100
        for k in (1, 2, 3):
            np.testing.assert_allclose(t.get_s(k)[ok], (k + xyz[:, :dim] @ coef * k)[ok], rtol=0, atol=1e-12)
