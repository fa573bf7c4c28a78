"""bmg_poisson_noalloc: poisson() whose device arrays all come from the
library's static scratch (no hipMalloc), called through the C ABI.  Its
solutions and mgstats equal bmg_poisson's on the same inputs bit for bit, on
two calls in a row (the second smaller, over the first call's leftovers),
in FP64 and in mixed precision; a problem larger than the scratch is
refused before any kernel runs."""
import ctypes

import numpy as np
import pytest
import torch

import basilisk_amd as bmg
from tests import cases

pytestmark = pytest.mark.gpu

DP = ctypes.POINTER(ctypes.c_double)


class Problem(ctypes.Structure):
    _fields_ = [("dimension", ctypes.c_int), ("depth", ctypes.c_int),
                ("L0", ctypes.c_double), ("X0", ctypes.c_double), ("Y0", ctypes.c_double),
                ("Z0", ctypes.c_double), ("precision", ctypes.c_int), ("smoother", ctypes.c_int),
                ("bc_kind", ctypes.c_int * 6), ("bc_sigma", ctypes.c_int * 6), ("bc_c", DP * 6),
                ("a", ctypes.c_void_p), ("b", ctypes.c_void_p), ("alpha", ctypes.c_void_p * 3),
                ("alpha_c", ctypes.c_double), ("lambda_", ctypes.c_void_p), ("lambda_c", ctypes.c_double),
                ("TOLERANCE", ctypes.c_double), ("NITERMAX", ctypes.c_int), ("NITERMIN", ctypes.c_int),
                ("stream", ctypes.c_void_p)]


def lib():
    L = bmg.library()
    L.bmg_poisson_noalloc.restype = ctypes.c_int
    L.bmg_poisson_noalloc.argtypes = [ctypes.POINTER(Problem), ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(bmg.MGStats)]
    return L


def random_inputs(dim, depth, seed, periodic_x):
    rng = np.random.default_rng(seed)
    n = 1 << depth
    b = rng.standard_normal((n,) * dim)
    if periodic_x:
        b -= b.mean()
    alpha = []
    for c in range(dim):
        shape = [n] * dim
        shape[c] += 1
        alpha.append(np.exp(0.4 * rng.standard_normal(shape)))
    lam = -rng.uniform(0.5, 2.0, (n,) * dim)
    dvals = rng.standard_normal(n ** (dim - 1))
    return b, alpha, lam, dvals


def solve_noalloc(L, dim, depth, b, alpha, lam, dvals, periodic_x, precision, keep):
    dev = torch.device("cuda:0")
    n = 1 << depth
    a = torch.zeros((n,) * dim, dtype=torch.float64, device=dev)
    bt = torch.from_numpy(b).to(dev)
    at = [torch.from_numpy(x).to(dev) for x in alpha]
    lt = torch.from_numpy(lam).to(dev)
    torch.cuda.synchronize()
    q = Problem()
    q.dimension, q.depth, q.L0 = dim, depth, 1.0
    q.precision = bmg.PRECISIONS[precision]
    q.smoother = bmg.SMOOTHERS["rbgs"]
    c = np.ascontiguousarray(2.0 * dvals)
    keep.append(c)
    for side in range(2 * dim):
        if periodic_x and side < 2:
            q.bc_kind[side] = bmg.BC_PERIODIC
        else:
            q.bc_kind[side] = bmg.BC_AFFINE
            q.bc_sigma[side] = 1
    if not periodic_x:
        q.bc_sigma[0] = -1  # Dirichlet on the right: ghost = 2 v - s
        q.bc_c[0] = c.ctypes.data_as(DP)
    q.a, q.b = a.data_ptr(), bt.data_ptr()
    for k in range(dim):
        q.alpha[k] = at[k].data_ptr()
    q.lambda_ = lt.data_ptr()
    q.TOLERANCE, q.NITERMAX, q.NITERMIN = 1e-3, 100, 1
    st = bmg.MGStats()
    rc = L.bmg_poisson_noalloc(ctypes.byref(q), 1e-9, 0, 0, ctypes.byref(st))
    assert rc == 0, L.bmg_last_error()
    return a.cpu().numpy(), st.as_dict()


def solve_grid(dim, depth, b, alpha, lam, dvals, periodic_x, precision):
    g = bmg.Grid(dim, depth, precision=precision)
    g.set_b(b)
    g.set_alpha(alpha)
    g.set_lambda(lam)
    if periodic_x:
        g.set_periodic(0)
    else:
        g.set_bc(bmg.RIGHT, bmg.BC_DIRICHLET, dvals)
    g.set_a(np.zeros(g.shape()))
    st = g.poisson(tolerance=1e-9)
    a = g.get_a()
    g.close()
    return a, st


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_matches_poisson_twice(precision):
    L = lib()
    keep = []
    for k, (dim, depth, px) in enumerate([(3, 6, False), (2, 7, True)]):
        print("case noalloc", precision, dim, depth, "periodic" if px else "walls", flush=True)
        inp = random_inputs(dim, depth, 100 + k, px)
        a1, s1 = solve_noalloc(L, dim, depth, *inp, px, precision, keep)
        a2, s2 = solve_grid(dim, depth, *inp, px, precision)
        for key in ("i", "resb", "resa", "nrelax", "minlevel"):
            assert s1[key] == s2[key], (key, s1, s2)
        assert s1["resa"] <= 1e-9
        assert np.array_equal(a1, a2)


def test_noalloc_refuses_oversize_then_works():
    L = lib()
    keep = []
    print("case noalloc oversize", flush=True)
    n = 256
    q = Problem()
    q.dimension, q.depth, q.L0 = 3, 8, 1.0
    for side in range(6):
        q.bc_kind[side] = bmg.BC_AFFINE
        q.bc_sigma[side] = 1
    a = torch.zeros(n ** 3, dtype=torch.float64, device="cuda:0")
    b = torch.zeros(n ** 3, dtype=torch.float64, device="cuda:0")
    q.a, q.b = a.data_ptr(), b.data_ptr()
    q.alpha_c, q.TOLERANCE, q.NITERMAX, q.NITERMIN = 1.0, 1e-3, 100, 1
    st = bmg.MGStats()
    assert L.bmg_poisson_noalloc(ctypes.byref(q), 1e-9, 0, 0, ctypes.byref(st)) == 1
    assert b"scratch" in L.bmg_last_error()
    inp = random_inputs(3, 5, 7, False)
    a1, s1 = solve_noalloc(L, 3, 5, *inp, False, "fp64", keep)
    a2, s2 = solve_grid(3, 5, *inp, False, "fp64")
    assert s1["i"] == s2["i"] and s1["resa"] == s2["resa"]
    assert np.array_equal(a1, a2)
