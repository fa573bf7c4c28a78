"""Problem set-ups shared by the oracle and GPU tests.

Each function restates the set-up of one of the reference's solver tests
(`src/test/*.c`) or of a synthetic workload, and drives any grid object that
offers the methods used below (the oracle's ``OracleGrid`` and the GPU
library's ``basilisk_amd.Grid`` both do).
"""
import os

import numpy as np

PI = 3.14159265358979  # common.h:48 -- Basilisk's pi, not M_PI

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref")

BC_SYMMETRY, BC_DIRICHLET, BC_NEUMANN, BC_PERIODIC = range(4)


def read_ref(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return [line.rstrip("\n") for line in f if line.strip()]


def mesh(g, level=None):
    c = list(g.centers(level)) + [np.zeros(1)] * (3 - g.dim)
    return np.meshgrid(*c, indexing="ij")


def poisson_c(g, ncycles=13, nrelax=4):
    """test/poisson.c: Dirichlet cos(3 pi x) cos(3 pi y) cos(3 pi z) on
    [-0.5,0.5]^d, 13 mg_cycles with minlevel 0.  Returns the lines the
    program writes to stderr."""
    dim = g.dim

    def sol(x, y, z):
        return np.cos(3. * PI * x) * np.cos(3. * PI * y) * np.cos(3. * PI * z)

    g.set_bc_function(BC_DIRICHLET, sol)
    X, Y, Z = mesh(g)
    b = -9. * dim * PI * PI * np.cos(3. * PI * X) * np.cos(3. * PI * Y) * np.cos(3. * PI * Z)
    g.set_b(b.reshape(g.shape()))
    g.set_a(np.zeros(g.shape()))
    g.residual()
    out = []
    for i in range(ncycles):
        g.mg_cycle(nrelax, 0)
        out.append((i, g.residual()))
    err = np.abs(g.get_a() - sol(X, Y, Z).reshape(g.shape())).max()
    return out, err


def variable_c(g, ncycles=13, nrelax=4):
    """test/variable.c: alpha = x + y + 2 on faces, Dirichlet sin sin."""

    def sol(x, y, z):
        return np.sin(3. * PI * x) * np.sin(3. * PI * y)

    g.set_bc_function(BC_DIRICHLET, sol)
    X, Y, _ = mesh(g)
    X, Y = X[..., 0], Y[..., 0]
    b = ((-18. * (PI * PI)) * np.sin(3. * PI * X) * np.sin(3. * PI * Y) * (X + Y + 2.)
         + (3. * PI) * np.cos(3. * PI * X) * np.sin(3. * PI * Y)
         + (3. * PI) * np.sin(3. * PI * X) * np.cos(3. * PI * Y))
    g.set_b(b)
    g.set_a(np.zeros(g.shape()))
    n, D, X0 = g.n(), g.delta(), g.origin[0]
    xf = ((-0.5 + np.arange(n + 1)) + 0.5) * D + X0
    yc = (np.arange(n) + 0.5) * D + g.origin[1]
    g.set_alpha([(xf[:, None] + yc[None, :]) + 2., (yc[:, None] + xf[None, :]) + 2.])
    g.restriction_alpha_lambda()
    g.residual()
    out = []
    for i in range(ncycles):
        g.mg_cycle(nrelax, 0)
        out.append((i, g.residual()))
    err = np.abs(g.get_a() - sol(X, Y, 0)).max()
    return out, err


def periodic_c(g):
    """test/periodic.c for one n: poisson() with TOLERANCE 1e-10 on a
    periodic box; returns (max error after mean removal, mgstats)."""
    dim = g.dim
    for d in range(dim):
        g.set_periodic(d)
    X, Y, Z = mesh(g)
    b = (4. * dim * (PI * PI)) * np.sin(2. * PI * X) * np.cos(2. * PI * Y) * np.cos(2. * PI * Z)
    g.set_b(b.reshape(g.shape()))
    g.set_a(np.zeros(g.shape()))
    g.set_tolerance(1e-10)
    st = g.poisson()
    e = g.get_a() + (np.sin(2. * PI * X) * np.cos(2. * PI * Y) * np.cos(2. * PI * Z)).reshape(g.shape())
    dv = g.delta() ** dim
    e = e - (dv * e).sum() / (dv * e.size)  # statsf(e).sum / statsf(e).volume
    return e.max(), st


def poisson1d_c(g):
    """test/poisson1D.c for one depth: poisson() with TOLERANCE 1e-4."""

    def sol(x, y, z):
        return np.sin(3. * PI * x)

    g.set_bc_function(BC_DIRICHLET, sol)
    (X,) = g.centers()
    g.set_b(-((3. * PI) * (3. * PI)) * np.sin(3. * PI * X))
    g.set_a(0 * X)
    g.set_tolerance(1e-4)
    st = g.poisson()
    return np.abs(g.get_a() - np.sin(3. * PI * X)).max(), st


def random_problem(g, seed, lam=0.0, alpha=True, bcs=None):
    """A seeded variable-coefficient Poisson-Helmholtz problem with the given
    BC kinds for `a` (list of 2*dim kinds; Dirichlet/Neumann values random)."""
    rng = np.random.default_rng(seed)
    dim = g.dim
    if bcs is None:
        bcs = [BC_SYMMETRY] * (2 * dim)
    for b, kind in enumerate(bcs):
        if kind == BC_PERIODIC:
            if b % 2 == 0:
                g.set_periodic(b // 2)
            continue
        if kind in (BC_DIRICHLET, BC_NEUMANN):
            m = g.n() + 4
            g.set_bc(b, kind, rng.standard_normal(m ** (dim - 1)))
        else:
            g.set_bc(b, kind, None)
    g.set_b(rng.standard_normal(g.shape()))
    g.set_a(rng.standard_normal(g.shape()))
    if alpha:
        g.set_alpha([np.exp(0.5 * rng.standard_normal(g.face_shape(c))) for c in range(dim)])
    if not np.isscalar(lam) or lam != 0.0:
        g.set_lambda(lam)
    return rng
