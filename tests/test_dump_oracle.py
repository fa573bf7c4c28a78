"""dump() / restore() of the tree oracle (oracle/dump.py, output.h:1043-1390)
without a GPU.  The reference ships no dump files, so the format is pinned by
its own invariants: struct DumpHeader is 56 bytes with the host C layout (the
reference writes it with one fwrite); the records are the active cells in
pre-order, the leaf flag set on exactly the leaves, the subtree size of the
root equal to the record count and every parent's size 1 + its children's;
a parent's values are the average of its children's; restore() rebuilds the
same tree, caches and leaf values."""
import ctypes
import struct

import numpy as np
import pytest

from oracle import dump
from oracle.tree import T_A, T_B, OracleTree, t_s


def _tree(dim, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.3, 0.7, 3)
    deep = 6 if dim == 2 else 5

    def cond(x, y, z, lev):
        r = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + ((z - c[2]) ** 2 if dim == 3 else 0.0))
        return (lev < 2) | ((lev < deep) & (np.abs(r - 0.2) < 0.05))

    t = OracleTree(dim, 2.0, (-1.0, 0.5, 0.25))
    t.refine(cond)
    _, xyz = t.leaves()
    t.set_a(np.sin(3 * xyz[:, 0]) + xyz[:, 1])
    t.set_b(rng.standard_normal(len(xyz)))
    t.set_s(xyz[:, 0] * xyz[:, 1])
    return t


def test_header_layout():
    assert ctypes.sizeof(dump.DumpHeader) == 56
    assert [f[0] for f in dump.DumpHeader._fields_] == ["t", "len", "i", "depth", "npe", "version", "n"]
    assert dump.DumpHeader.len.offset == 8 and dump.DumpHeader.n.offset == 32


@pytest.mark.parametrize("dim", [2, 3])
def test_records_invariants(dim):
    t = _tree(dim, dim)
    data = dump.dump_bytes(t, [("a", T_A), ("b", T_B), ("s", t_s(dim))], t=2.5, i=11)
    h, names, origin, recs = dump.read_dump(data, dim)
    assert (h.t, h.len, h.i, h.depth, h.npe, h.version) == (2.5, 4, 11, t.depth(), 1, 170901)
    assert names == ["size", "a", "b", "s"]
    assert origin == (-1.0, 0.5, 0.25, 2.0)
    nactive = sum(len(t.cache("active", l)) for l in range(t.depth() + 1))
    assert len(recs) == nactive
    assert sum(1 for f, _ in recs if f == 2) == t.nleaves()
    assert all(f in (0, 2) for f, _ in recs)
    assert recs[0][1][0] == len(recs)
    # every parent: size 1 + children's, values the children's average
    nch = 1 << dim

    def check(q):
        f, v = recs[q]
        if f == 2:
            assert v[0] == 1.0
            return q + 1
        r, size, sums = q + 1, 1.0, [0.0, 0.0, 0.0]
        for _ in range(nch):
            size += recs[r][1][0]
            for m in range(3):
                sums[m] = sums[m] + recs[r][1][1 + m]
            r = check(r)
        assert v[0] == size
        for m in range(3):
            assert v[1 + m] == sums[m] / nch
        return r

    assert check(0) == len(recs)
    # the leaves' values in leaf order
    leafvals = [v[1] for f, v in recs if f == 2]
    assert np.array_equal(np.array(leafvals), t.get_a())


@pytest.mark.parametrize("dim", [2, 3])
def test_restore_round_trip(dim):
    t = _tree(dim, 10 + dim)
    data = dump.dump_bytes(t, [("a", T_A), ("s", t_s(dim))])
    r, h, names, fields = dump.restore_tree(data, dim)
    la, xa = t.leaves()
    lb, xb = r.leaves()
    assert np.array_equal(la, lb) and np.array_equal(xa, xb)
    for which in ("active", "prolongation", "boundary"):
        for l in range(t.depth() + 1):
            assert np.array_equal(t.cache(which, l), r.cache(which, l))
    assert np.array_equal(fields["a"], t.get_a())
    assert np.array_equal(fields["s"], t.get_s())
    # the dump of the restored tree is the same file
    r.set_a(fields["a"])
    r.set_s(fields["s"])
    assert dump.dump_bytes(r, [("a", T_A), ("s", t_s(dim))]) == data


def test_truncated_file_refused():
    t = _tree(2, 5)
    data = dump.dump_bytes(t, [("a", T_A)])
    with pytest.raises(ValueError):
        dump.read_dump(data[:-3], 2)
    h = struct.unpack_from("=i", data, 28)[0]
    assert h == 170901


@pytest.mark.parametrize("dim,depth", [(1, 5), (2, 4), (3, 3)])
def test_grid_records_invariants(dim, depth):
    """The multigrid dump: one record per cell of every level, leaves
    exactly the finest level, subtree sizes 1 + 2^dim x the child's, each
    coarse value the average of its children, the finest values in C order
    of a pre-order walk."""
    n = 1 << depth
    a = np.random.default_rng(dim).standard_normal((n,) * dim)
    data = dump.grid_dump_bytes([("a", a)], L0=3.0, origin=(1.0, 2.0, 3.0), t=0.5, i=2)
    h, names, origin, recs = dump.read_dump(data, dim)
    assert (h.depth, h.len, h.version, h.i) == (depth, 2, 170901, 2)
    assert origin == (1.0, 2.0, 3.0, 3.0)
    assert len(recs) == sum((1 << (dim * l)) for l in range(depth + 1))
    assert sum(1 for f, _ in recs if f == 2) == n ** dim
    assert recs[0][1][0] == len(recs)
    nch = 1 << dim

    def check(q, l):
        f, v = recs[q]
        if l == depth:
            assert f == 2 and v[0] == 1.0
            return q + 1, [v[1]]
        r, leaves, s = q + 1, [], 0.0
        for _ in range(nch):
            s = s + recs[r][1][1]
            r, lv = check(r, l + 1)
            leaves += lv
        assert v[1] == s / nch
        return r, leaves

    end, leaves = check(0, 0)
    assert end == len(recs)
    # the pre-order leaf sequence is a permutation of the finest level
    assert sorted(leaves) == sorted(a.reshape(-1).tolist())
