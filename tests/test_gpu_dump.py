"""dump() / restore() of the GPU tree (bmg_tree_dump / bmg_tree_restore,
output.h:1043-1390) against the oracle's writer and reader
(oracle/dump.py, pinned by test_dump_oracle.py): the GPU tree's file is the
oracle's byte for byte; the GPU tree restored from the oracle's file has the
same caches and leaf values, and solves the same poisson() problem bit for
bit."""
import numpy as np
import pytest

import basilisk_amd as bmg
from oracle import dump
from oracle.tree import T_A, T_B, OracleTree, t_s

pytestmark = pytest.mark.gpu

ORIGIN = (-1.0, 0.5, 0.25)


def _refine(dim, full):
    deep = 6 if dim == 2 else 4

    def cond(x, y, z, lev):
        if full:
            return lev < deep
        r = np.sqrt((x + 0.4) ** 2 + (y - 1.1) ** 2 + ((z - 0.8) ** 2 if dim == 3 else 0.0))
        return (lev < 2) | ((lev < deep + 1) & (np.abs(r - 0.35) < 0.1))
    return cond


def _fill(t, dim, seed):
    rng = np.random.default_rng(seed)
    _, xyz = t.leaves()
    t.set_a(np.sin(3 * xyz[:, 0]) + xyz[:, 1] ** 2)
    t.set_b(rng.standard_normal(len(xyz)))
    t.set_s(xyz[:, 0] * xyz[:, 1])


def _same_tree(g, o):
    assert g.depth() == o.depth() and g.nleaves() == o.nleaves()
    la, xa = g.leaves()
    lb, xb = o.leaves()
    assert np.array_equal(la, lb) and np.array_equal(xa, xb)
    for which in ("active", "prolongation", "boundary"):
        for l in range(o.depth() + 1):
            assert np.array_equal(g.cache(which, l), o.cache(which, l)), (which, l)


@pytest.mark.parametrize("dim,full", [(2, False), (2, True), (3, False), (3, True)])
def test_dump_equals_oracle_file(dim, full, tmp_path):
    print("case dump", dim, full, flush=True)
    g = bmg.Tree(dim, 2.0, ORIGIN)
    o = OracleTree(dim, 2.0, ORIGIN)
    for t in (g, o):
        t.refine(_refine(dim, full))
        _fill(t, dim, 4)
    path = tmp_path / "g.dump"
    g.dump(path, names={"a": "p", "b": "rhs", "s": "f"}, t=3.25, i=42)
    data = path.read_bytes()
    ref = dump.dump_bytes(o, [("p", T_A), ("rhs", T_B), ("f", t_s(dim))], t=3.25, i=42)
    assert len(data) == len(ref)
    assert data == ref


@pytest.mark.parametrize("dim", [2, 3])
def test_restore_oracle_file_then_solve(dim, tmp_path):
    print("case restore", dim, flush=True)
    o = OracleTree(dim, 2.0, ORIGIN)
    o.refine(_refine(dim, False))
    _fill(o, dim, 9)
    path = tmp_path / "o.dump"
    path.write_bytes(dump.dump_bytes(o, [("p", T_A), ("rhs", T_B), ("f", t_s(dim))], t=0.5, i=3))
    g = bmg.Tree(dim)  # any box: restore takes the file's
    t, i = g.restore(path, names={"a": "p", "b": "rhs", "s": "f"})
    assert (t, i) == (0.5, 3)
    _same_tree(g, o)
    assert np.array_equal(g.get_a(), o.get_a())
    assert np.array_equal(g.get_b(), o.get_b())
    assert np.array_equal(g.get_s(), o.get_s())
    # the restored tree solves: poisson() bit for bit with the oracle
    g.set_smoother("rbgs")
    o.set_smoother("rbgs")
    for tr in (g, o):
        tr.set_alpha_const(1.0)
        tr.set_lambda(-1.0)
        tr.set_niter(30)
    sg = g.poisson(tolerance=1e-7)
    so = o.poisson(tolerance=1e-7)
    for k in ("i", "resb", "resa", "nrelax"):
        assert sg[k] == so[k], k
    assert np.array_equal(g.get_a(), o.get_a())
    # and its own dump is the oracle's file again
    g2 = tmp_path / "g2.dump"
    _fill(o, dim, 9)
    _fill(g, dim, 9)
    g.dump(g2, names={"a": "p", "b": "rhs", "s": "f"}, t=0.5, i=3)
    assert g2.read_bytes() == path.read_bytes()


def test_restore_errors_and_missing_names(tmp_path):
    print("case restore errors", flush=True)
    o = OracleTree(2, 1.0, (0.0, 0.0, 0.0))
    o.refine(lambda x, y, z, lev: lev < 3)
    _fill(o, 2, 1)
    path = tmp_path / "x.dump"
    data = dump.dump_bytes(o, [("p", T_A)])
    path.write_bytes(data)
    g = bmg.Tree(2)
    g.restore(path, names={"a": "p", "b": "missing"})
    assert np.array_equal(g.get_a(), o.get_a())
    assert not g.get_b().any()
    bad = tmp_path / "bad.dump"
    bad.write_bytes(data[:-5])
    with pytest.raises(bmg.BMGError):
        g.restore(bad)
    with pytest.raises(bmg.BMGError):
        g.restore(tmp_path / "nope.dump")


@pytest.mark.parametrize("dim,depth", [(2, 6), (3, 4), (1, 7)])
def test_grid_dump_equals_oracle_and_restores(dim, depth, tmp_path):
    """The multigrid grid's dump (every level in pre-order, coarse cells the
    restriction of the finest level) equals the oracle's writer byte for
    byte; restoring it into a fresh grid gives the finest level back and the
    file's origin and box size."""
    print("case grid dump", dim, depth, flush=True)
    rng = np.random.default_rng(dim + depth)
    n = 1 << depth
    a = rng.standard_normal((n,) * dim)
    b = rng.standard_normal((n,) * dim)
    g = bmg.Grid(dim, depth, L0=2.0, origin=(-1.0, 0.5, 0.25))
    g.set_a(a)
    g.set_b(b)
    path = tmp_path / "grid.dump"
    g.dump(path, names={"a": "p", "b": "rhs"}, t=1.25, i=9)
    ref = dump.grid_dump_bytes([("p", a), ("rhs", b)], L0=2.0, origin=(-1.0, 0.5, 0.25), t=1.25, i=9)
    assert path.read_bytes() == ref
    h = bmg.Grid(dim, depth)
    assert h.restore(path, names={"a": "p", "b": "rhs"}) == (1.25, 9)
    assert np.array_equal(h.get_a(), a) and np.array_equal(h.get_b(), b)
    h.dump(tmp_path / "again.dump", names={"a": "p", "b": "rhs"}, t=1.25, i=9)
    assert (tmp_path / "again.dump").read_bytes() == ref
    with pytest.raises(bmg.BMGError):
        bmg.Grid(dim, depth - 1).restore(path)


@pytest.mark.parametrize("dim", [2, 3])
def test_root_only_tree_dump(dim, tmp_path):
    """A tree of one leaf: one record, size 1, flags leaf."""
    print("case dump root", dim, flush=True)
    g = bmg.Tree(dim)
    o = OracleTree(dim)
    for t in (g, o):
        t.set_a(np.array([2.5]))
        t.set_b(np.array([-1.0]))
        t.set_s(np.array([0.25]))
    path = tmp_path / "root.dump"
    g.dump(path, names={"a": "a", "b": "b", "s": "s"})
    assert path.read_bytes() == dump.dump_bytes(o, [("a", T_A), ("b", T_B), ("s", t_s(dim))])
    h = bmg.Tree(dim)
    h.refine(lambda x, y, z, lev: lev < 2)
    h.restore(path, names={"a": "a", "s": "s"})
    assert h.nleaves() == 1 and h.depth() == 0
    assert h.get_a()[0] == 2.5 and h.get_s()[0] == 0.25
