"""bmg_poisson_view: poisson() on the reference's own multigrid storage.

The test lays out a buffer exactly as grid/multigrid.h:16-73 does -- one
array per level of (2^l + 4)^dim cell records of `datasize` bytes, field s at
byte offset 8*s.i, two ghost layers -- and checks that the view entry point
solves the same problem as the SoA entry points, bit for bit, writing only
the interior cells of `a` (and `res`).
"""
import ctypes

import numpy as np
import pytest

import basilisk_amd as bmg
from tests import cases

pytestmark = pytest.mark.gpu


class View(ctypes.Structure):
    _fields_ = [("dimension", ctypes.c_int), ("depth", ctypes.c_int),
                ("d", ctypes.POINTER(ctypes.c_void_p)), ("datasize", ctypes.c_size_t),
                ("nvarmax", ctypes.c_int), ("constant", ctypes.POINTER(ctypes.c_double)),
                ("L0", ctypes.c_double), ("X0", ctypes.c_double), ("Y0", ctypes.c_double),
                ("Z0", ctypes.c_double)]


class Args(ctypes.Structure):
    _fields_ = [("a", ctypes.c_int), ("b", ctypes.c_int), ("alpha", ctypes.c_int * 3),
                ("lambda_", ctypes.c_int), ("tolerance", ctypes.c_double),
                ("nrelax", ctypes.c_int), ("minlevel", ctypes.c_int), ("res", ctypes.c_int)]


class ViewBC(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int * 6), ("sigma", ctypes.c_int * 6),
                ("c", ctypes.POINTER(ctypes.c_double) * 6)]


NVAR, NVARMAX = 9, 64
A, B, RES, AX, AY, AZ, LAM = 0, 1, 2, 3, 4, 5, 6


def interior(levels, l, dim, h):
    n = 1 << l
    arr = levels[l].reshape((n + 4,) * dim + (NVAR,))
    return arr[(slice(2, n + 2),) * dim + (h,)]


def make_problem(dim, depth, seed=31):
    """Reference storage with a, b, alpha, lambda and two unrelated fields
    (handles 7 and 8, random, ghosts included), Dirichlet left/right."""
    rng = np.random.default_rng(seed)
    n = 1 << depth
    levels = [rng.standard_normal((((1 << l) + 4) ** dim) * NVAR) for l in range(depth + 1)]
    fin = levels[depth].reshape((n + 4,) * dim + (NVAR,))
    fin[..., :7] = 0.0
    b = rng.standard_normal((n,) * dim)
    interior(levels, depth, dim, B)[...] = b
    alpha = []
    for c in range(dim):
        shape = [n] * dim
        shape[c] += 1
        ac = np.exp(0.3 * rng.standard_normal(shape))
        alpha.append(ac)
        fin[tuple(slice(2, 2 + e) for e in shape) + (AX + c,)] = ac
    lam = -rng.uniform(0.5, 2.0, (n,) * dim)
    interior(levels, depth, dim, LAM)[...] = lam
    vals = [rng.standard_normal(n ** (dim - 1)) for _ in range(2)]
    return levels, b, alpha, lam, vals


def call_view(levels, dim, depth, vals, tol):
    bc = ViewBC()
    keep = []
    for side in range(2 * dim):
        bc.kind[side] = 4  # BMG_BC_AFFINE
        if side < 2:
            c = np.ascontiguousarray(2.0 * vals[side])
            keep.append(c)
            bc.sigma[side] = -1
            bc.c[side] = c.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        else:
            bc.sigma[side] = 1
    ptrs = (ctypes.c_void_p * (depth + 1))(*[lv.ctypes.data for lv in levels])
    const = np.zeros(4)
    view = View(dim, depth, ptrs, NVAR * 8, NVARMAX, const.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                1.0, 0.0, 0.0, 0.0)
    args = Args(A, B, (ctypes.c_int * 3)(AX, AY, AZ if dim == 3 else 0), LAM, tol, 0, 0, RES)
    stats = bmg.MGStats()
    lib = bmg.library()
    rc = lib.bmg_poisson_view(ctypes.byref(view), ctypes.byref(args), ctypes.byref(bc), 1e-3, 100, 1,
                              ctypes.byref(stats))
    assert rc == 0, lib.bmg_last_error()
    return stats


@pytest.mark.parametrize("smoother", ["gs", "rbgs"])
@pytest.mark.parametrize("dim,depth", [(2, 7), (3, 5)])
def test_view_against_oracle(dim, depth, smoother):
    """bmg_view_configure(FP64, smoother): the view's solve equals the CPU
    oracle's poisson() bit for bit -- with the reference's own GS, its
    every iterate -- and every byte of the record arrays other than the
    interior of a and res is unchanged."""
    from oracle.oracle import OracleGrid
    lib = bmg.library()
    print("case view vs oracle", dim, depth, smoother, flush=True)
    assert lib.bmg_view_configure(0, bmg.SMOOTHERS[smoother]) == 0
    try:
        levels, b, alpha, lam, vals = make_problem(dim, depth, seed=5)
        before = [lv.copy() for lv in levels]
        stats = call_view(levels, dim, depth, vals, 1e-9)
    finally:
        assert lib.bmg_view_configure(0, bmg.SMOOTHERS["rbgs"]) == 0
    o = OracleGrid(dim, depth)
    o.set_smoother(smoother)
    o.set_b(b)
    o.set_alpha(alpha)
    o.set_lambda(lam)
    # the oracle takes BC values over the ghost-inclusive tangential extent
    n = 1 << depth
    for side, v in ((bmg.RIGHT, vals[0]), (bmg.LEFT, vals[1])):
        full = np.zeros((n + 4,) * (dim - 1))
        full[(slice(2, n + 2),) * (dim - 1)] = v.reshape((n,) * (dim - 1))
        o.set_bc(side, bmg.BC_DIRICHLET, full.reshape(-1))
    o.set_a(np.zeros(o.shape()))
    so = o.poisson(tolerance=1e-9)
    assert stats.i == so["i"] and stats.resa == so["resa"] and stats.nrelax == so["nrelax"]
    assert np.array_equal(interior(levels, depth, dim, A), o.get_a())
    for l in range(depth + 1):
        if l < depth:
            assert np.array_equal(levels[l], before[l])
            continue
        now = levels[l].reshape((n + 4,) * dim + (NVAR,)).copy()
        old = before[l].reshape((n + 4,) * dim + (NVAR,)).copy()
        for h in (A, RES):
            now[(slice(2, n + 2),) * dim + (h,)] = 0.0
            old[(slice(2, n + 2),) * dim + (h,)] = 0.0
        assert np.array_equal(now, old)


def test_view_mixed_precision():
    """bmg_view_configure(MIXED, RBGS): the view runs the mixed solver (the
    same iterates as the SoA grid in mixed precision) and reaches the
    tolerance within a cycle of FP64."""
    lib = bmg.library()
    dim, depth = 3, 5
    assert lib.bmg_view_configure(1, bmg.SMOOTHERS["rbgs"]) == 0
    try:
        levels, b, alpha, lam, vals = make_problem(dim, depth, seed=9)
        stats = call_view(levels, dim, depth, vals, 1e-9)
    finally:
        assert lib.bmg_view_configure(0, bmg.SMOOTHERS["rbgs"]) == 0
    g = bmg.Grid(dim, depth, precision="mixed")
    g.set_b(b)
    g.set_alpha(alpha)
    g.set_lambda(lam)
    g.set_bc(bmg.RIGHT, bmg.BC_DIRICHLET, vals[0])
    g.set_bc(bmg.LEFT, bmg.BC_DIRICHLET, vals[1])
    g.set_a(np.zeros(g.shape()))
    s = g.poisson(tolerance=1e-9)
    assert stats.i == s["i"] and stats.resa == s["resa"] and stats.resa <= 1e-9
    assert np.array_equal(interior(levels, depth, dim, A), g.get_a())


def test_view_configure_rejects_unknown():
    lib = bmg.library()
    assert lib.bmg_view_configure(2, 0) == 1
    assert lib.bmg_view_configure(0, 9) == 1


@pytest.mark.parametrize("dim,depth", [(2, 7), (3, 5)])
def test_view_matches_soa(dim, depth):
    rng = np.random.default_rng(31)
    n = 1 << depth
    levels = [np.zeros((((1 << l) + 4) ** dim) * NVAR) for l in range(depth + 1)]
    # reference storage: fields a, b, alpha.{x,y,z}, lambda at the finest level
    b = rng.standard_normal((n,) * dim)
    interior(levels, depth, dim, B)[...] = b
    alpha = []
    for c in range(dim):
        shape = [n] * dim
        shape[c] += 1
        ac = np.exp(0.3 * rng.standard_normal(shape))
        alpha.append(ac)
        m = n + 4
        arr = levels[depth].reshape((m,) * dim + (NVAR,))
        sl = tuple(slice(2, 2 + s) for s in shape)
        arr[sl + (AX + c,)] = ac
    lam = -rng.uniform(0.5, 2.0, (n,) * dim)
    interior(levels, depth, dim, LAM)[...] = lam
    # Dirichlet on left/right (c = 2v, sigma = -1), symmetry elsewhere
    vals = [rng.standard_normal(n ** (dim - 1)) for _ in range(2)]
    bc = ViewBC()
    keep = []
    for side in range(2 * dim):
        bc.kind[side] = 4  # BMG_BC_AFFINE
        if side < 2:
            c = np.ascontiguousarray(2.0 * vals[side])
            keep.append(c)
            bc.sigma[side] = -1
            bc.c[side] = c.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        else:
            bc.sigma[side] = 1
    ptrs = (ctypes.c_void_p * (depth + 1))(*[lv.ctypes.data for lv in levels])
    const = np.zeros(4)
    view = View(dim, depth, ptrs, NVAR * 8, NVARMAX, const.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                1.0, 0.0, 0.0, 0.0)
    args = Args(A, B, (ctypes.c_int * 3)(AX, AY, AZ if dim == 3 else 0), LAM, 1e-8, 0, 0, RES)
    stats = bmg.MGStats()
    lib = bmg.library()
    rc = lib.bmg_poisson_view(ctypes.byref(view), ctypes.byref(args), ctypes.byref(bc), 1e-3, 100, 1,
                              ctypes.byref(stats))
    assert rc == 0, lib.bmg_last_error()
    # the same problem through the SoA entry points
    g = bmg.Grid(dim, depth)
    g.set_b(b)
    g.set_alpha(alpha)
    g.set_lambda(lam)
    g.set_bc(bmg.RIGHT, bmg.BC_DIRICHLET, vals[0])
    g.set_bc(bmg.LEFT, bmg.BC_DIRICHLET, vals[1])
    g.set_a(np.zeros(g.shape()))
    s = g.poisson(tolerance=1e-8)
    assert stats.i == s["i"] and stats.resa == s["resa"] and stats.nrelax == s["nrelax"]
    assert np.array_equal(interior(levels, depth, dim, A), g.get_a())
    assert np.array_equal(interior(levels, depth, dim, RES), g.get_res())
    # ghost cells of the host storage are left alone
    full = levels[depth].reshape((n + 4,) * dim + (NVAR,))
    assert np.all(full[(0,) + (slice(None),) * (dim - 1) + (A,)] == 0.0)
