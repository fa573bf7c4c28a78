"""CPU oracle of dump() / restore() on trees (output.h:1043-1390, the
non-MPI branch, file version 170901).  TEST INFRASTRUCTURE ONLY: only tests/
may import this module; the product library never calls it.

The file, in the host's byte order:

  struct DumpHeader {double t; long len; int i, depth, npe, version;
                     coord n;}                     (output.h:1049-1053;
                                                     coord = 3 doubles)
  len x {unsigned length; char name[length]}      the names of `list`, the
                                                     subtree "size" first
                                                     (output.h:1067-1086)
  double o[4] = {X0, Y0, Z0, L0}
  then foreach_cell() in pre-order, stopping at leaves (so every active
  cell, non-leaf cells included): unsigned flags (leaf = 2 if a leaf, else
  0, grid/tree.h:31-34), then one double per scalar of `list`: the subtree
  size (1 for a leaf, 1 + the children's sizes otherwise: subtree_size
  (size, false), grid/multigrid-common.h:422-445), then the fields.

restore() (output.h:1221-1390) rebuilds the tree from init_grid(1) by the
same traversal, refining a leaf whose record is not a leaf.
"""
import ctypes
import struct

import numpy as np

DUMP_VERSION = 170901
LEAF = 2


class DumpHeader(ctypes.Structure):
    """struct DumpHeader with the host C layout."""
    _fields_ = [("t", ctypes.c_double), ("len", ctypes.c_long), ("i", ctypes.c_int), ("depth", ctypes.c_int),
                ("npe", ctypes.c_int), ("version", ctypes.c_int), ("n", ctypes.c_double * 3)]


def _preorder(tree):
    """(level, (i, j, k), is_leaf) of every active cell of an OracleTree in
    foreach_cell() order."""
    from oracle.tree import GH
    dim = tree.dim
    meta = [tree.level_meta(l) for l in range(tree.depth() + 1)]
    out = []

    def visit(l, c):
        alloc, flags = meta[l][0], meta[l][1]
        idx = tuple(c[:dim])
        if not alloc[idx]:
            return
        leaf = bool(flags[idx] & 2)
        out.append((l, tuple(c), leaf))
        if leaf or l >= tree.depth():
            return
        for a in range(2):
            for b in range(2):
                for e in range(2 if dim == 3 else 1):
                    visit(l + 1, (2 * c[0] - GH + a, 2 * c[1] - GH + b, 2 * c[2] - GH + e if dim == 3 else 0))

    visit(0, (GH, GH, GH if dim == 3 else 0))
    return out


def _sizes(cells, dim):
    """Subtree sizes of the cells in pre-order: 1 for a leaf, 1 + the sizes
    of its 2^dim children (which follow it) otherwise."""
    size = np.zeros(len(cells))

    def parse(q):
        if cells[q][2]:
            size[q] = 1.0
            return q + 1
        s, r = 1.0, q + 1
        for _ in range(1 << dim):
            nxt = parse(r)
            s += size[r]
            r = nxt
        size[q] = s
        return r

    if parse(0) != len(cells):
        raise ValueError("pre-order sequence is not one tree")
    return size


def _restricted(cells, leafval, dim):
    """Values of the cells in pre-order: a leaf's own value, a non-leaf
    cell's restriction_average of its children (summed from 0 in
    foreach_child order, grid/multigrid-common.h:26-32), as the tree
    boundary leaves them before a dump."""
    val = np.zeros(len(cells))

    def parse(q):
        if cells[q][2]:
            val[q] = leafval[q]
            return q + 1
        s, r = 0.0, q + 1
        for _ in range(1 << dim):
            nxt = parse(r)
            s = s + val[r]
            r = nxt
        val[q] = s / float(1 << dim)
        return r

    parse(0)
    return val


def dump_bytes(tree, fields, t=0.0, i=0):
    """The dump of an OracleTree: `fields` is a list of (name, field index);
    leaves write their values, non-leaf cells the restriction of them."""
    cells = _preorder(tree)
    size = _sizes(cells, tree.dim)
    names = ["size"] + [nm for nm, _ in fields]
    h = DumpHeader(t, len(names), i, tree.depth(), 1, DUMP_VERSION, (0.0, 0.0, 0.0))
    out = bytearray(bytes(h))
    for nm in names:
        b = nm.encode()
        out += struct.pack("=I", len(b)) + b
    o = list(tree.origin[:3]) + [tree.L0]
    out += struct.pack("=4d", *o)
    levels = {}
    for _, f in fields:
        levels[f] = [np.ctypeslib.as_array(tree.lib.ot_level_field(tree.t, l, f),
                                           shape=((1 << l) + 2 * 2,) * tree.dim) for l in range(tree.depth() + 1)]
    cols = []
    for _, f in fields:
        raw = np.array([levels[f][l][tuple(c[:tree.dim])] for (l, c, _) in cells])
        cols.append(_restricted(cells, raw, tree.dim))
    for q, (l, c, leaf) in enumerate(cells):
        out += struct.pack("=I", LEAF if leaf else 0)
        vals = [size[q]] + [col[q] for col in cols]
        out += struct.pack("=%dd" % len(vals), *vals)
    return bytes(out)


def read_dump(data, dim):
    """Parse a dump: (header, names, origin (X0, Y0, Z0, L0), records as a
    list of (flags, values))."""
    hs = ctypes.sizeof(DumpHeader)
    h = DumpHeader.from_buffer_copy(data[:hs])
    p = hs
    names = []
    for _ in range(h.len):
        (n,) = struct.unpack_from("=I", data, p)
        p += 4
        names.append(data[p:p + n].decode())
        p += n
    origin = struct.unpack_from("=4d", data, p)
    p += 32
    rec = 4 + 8 * h.len
    nrec = (len(data) - p) // rec
    if (len(data) - p) % rec:
        raise ValueError("truncated dump")
    records = []
    for q in range(nrec):
        (flags,) = struct.unpack_from("=I", data, p)
        vals = struct.unpack_from("=%dd" % h.len, data, p + 4)
        records.append((flags, vals))
        p += rec
    return h, names, origin, records


def restore_tree(data, dim):
    """restore(): an OracleTree rebuilt from the records, and the leaf
    values (leaf order) of every dumped field by name."""
    from oracle.tree import GH, OracleTree
    h, names, origin, records = read_dump(data, dim)
    t = OracleTree(dim, origin[3], origin[:3])
    pos = [0]
    cellvals = {}

    def visit(l, c):
        meta = t.level_meta(l) if l <= t.depth() else None
        if meta is None or not meta[0][tuple(c[:dim])]:
            return
        flags, vals = records[pos[0]]
        pos[0] += 1
        cellvals[(l,) + tuple(c)] = vals
        if not (flags & LEAF) and (meta[1][tuple(c[:dim])] & 2):
            t.refine_cell_at(l, c)
        meta = t.level_meta(l)
        if meta[1][tuple(c[:dim])] & 2 or l >= t.depth():
            return
        for a in range(2):
            for b in range(2):
                for e in range(2 if dim == 3 else 1):
                    visit(l + 1, (2 * c[0] - GH + a, 2 * c[1] - GH + b, 2 * c[2] - GH + e if dim == 3 else 0))

    visit(0, (GH, GH, GH if dim == 3 else 0))
    if pos[0] != len(records):
        raise ValueError("dump holds %d records, the tree %d cells" % (len(records), pos[0]))
    lev, xyz = t.leaves()
    fields = {}
    keys = []
    for n in range(len(lev)):
        l = int(lev[n])
        D = origin[3] / (1 << l)
        c = [int(round((xyz[n, d] - origin[d]) / D - 0.5)) + GH if d < dim else 0 for d in range(3)]
        keys.append((l,) + tuple(c))
    for m, nm in enumerate(names[1:], start=1):
        fields[nm] = np.array([cellvals[k][m] for k in keys])
    return t, h, names, fields


def grid_dump_bytes(fields, L0=1.0, origin=(0.0, 0.0, 0.0), t=0.0, i=0):
    """The dump of a uniform multigrid grid (output.h:1088-1118 on
    grid/multigrid.h): `fields` is a list of (name, finest-level array of
    shape (n,)*dim).  foreach_cell() visits every cell of every level in
    pre-order; the leaves are the finest level; a coarse cell holds the
    restriction_average of its children (as multigrid_restriction leaves
    it)."""
    a0 = np.asarray(fields[0][1])
    dim, n = a0.ndim, a0.shape[0]
    depth = n.bit_length() - 1
    levels = []
    for _, f in fields:
        lv = [None] * (depth + 1)
        lv[depth] = np.asarray(f, np.float64)
        for l in range(depth - 1, -1, -1):
            s = np.zeros((1 << l,) * dim)
            for ch in np.ndindex(*(2,) * dim):
                s = s + lv[l + 1][tuple(slice(o, None, 2) for o in ch)]
            lv[l] = s / float(1 << dim)
        levels.append(lv)
    size = [0.0] * (depth + 1)
    size[depth] = 1.0
    for l in range(depth - 1, -1, -1):
        size[l] = 1.0 + (1 << dim) * size[l + 1]
    names = ["size"] + [nm for nm, _ in fields]
    h = DumpHeader(t, len(names), i, depth, 1, DUMP_VERSION, (0.0, 0.0, 0.0))
    out = bytearray(bytes(h))
    for nm in names:
        b = nm.encode()
        out += struct.pack("=I", len(b)) + b
    out += struct.pack("=4d", *(list(origin[:3]) + [L0]))
    fmt = "=I%dd" % len(names)

    def visit(l, idx):
        vals = [size[l]] + [lv[l][idx] for lv in levels]
        out.extend(struct.pack(fmt, LEAF if l == depth else 0, *vals))
        if l == depth:
            return
        for ch in np.ndindex(*(2,) * dim):
            visit(l + 1, tuple(2 * idx[d] + ch[d] for d in range(dim)))

    visit(0, (0,) * dim)
    return bytes(out)
