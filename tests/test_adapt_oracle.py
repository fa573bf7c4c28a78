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
