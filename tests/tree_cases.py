"""Adaptive-tree problem set-ups shared by the oracle and GPU tests.

Each function restates one of the reference's tree solver tests
(`src/test/circle.c`, `circle-variable.c`) or a synthetic workload, and drives
any tree object with the methods used below (the oracle's ``OracleTree`` and
the GPU library's ``basilisk_amd.Tree``).
"""
import numpy as np

from tests.cases import PI, BC_DIRICHLET


def circle_refine(depth):
    """The refinement of test/circle.c:
    refine (level < depth - 2 || level <= depth*(1. - sqrt(x*x + y*y + z*z)))."""
    def cond(x, y, z, lev):
        return (lev < depth - 2) | (lev <= depth * (1. - np.sqrt(x * x + y * y + z * z)))
    return cond


def circle_setup(t, depth, variable=False):
    """Refines `t` (built on [-0.5,0.5]^d) and sets b, a, the Dirichlet BCs
    and alpha of test/circle.c or test/circle-variable.c.  Returns the exact
    solution as a function of (x, y, z)."""
    dim = t.dim
    t.refine(circle_refine(depth))
    _, xyz = t.leaves()
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    if variable:
        def sol(x, y, z):
            return np.sin(3. * PI * x) * np.sin(3. * PI * y)
        b = (-18. * PI ** 2 * np.sin(3. * PI * x) * np.sin(3. * PI * y) * (x + y + 2.)
             + 3. * PI * np.cos(3. * PI * x) * np.sin(3. * PI * y)
             + 3. * PI * np.sin(3. * PI * x) * np.cos(3. * PI * y))
    else:
        def sol(x, y, z):
            return np.cos(3. * PI * x) * np.cos(3. * PI * y) * np.cos(3. * PI * z)
        b = -9. * dim * PI * PI * np.cos(3. * PI * x) * np.cos(3. * PI * y) * np.cos(3. * PI * z)
    t.set_b(b)
    t.set_a(np.zeros_like(b))
    for side in range(2 * dim):
        t.set_bc_function(side, BC_DIRICHLET, sol)
    t.set_lambda(0.0)
    if variable:
        t.set_alpha_function(lambda x, y, z, d: x + y + 2.)
        t.restriction_alpha_lambda()
    return sol


def circle_run(t, depth, niter, nrelax=4):
    """The cycle loop of test/circle.c: residual, then niter x {mg_cycle
    (minlevel 0, maxlevel depth); residual}.  Returns the max residuals."""
    out = []
    t.residual()
    for _ in range(niter):
        t.mg_cycle(nrelax, 0, depth)
        out.append(t.residual())
    return out


def circle_lines(t, depth, variable=False):
    """The stderr lines test/circle.c (or circle-variable.c) prints."""
    sol = circle_setup(t, depth, variable)
    res = circle_run(t, t.depth(), 13 if variable else 15)
    _, xyz = t.leaves()
    err = np.abs(t.get_a() - sol(xyz[:, 0], xyz[:, 1], xyz[:, 2])).max()
    return ["residual %d %d %.1g" % (depth, i, r) for i, r in enumerate(res)] + \
        ["max error %d %g" % (depth, err)]


def spheres_tree(t, depth, minlevel, nspheres, seed, rmin, rmax, band=2.0):
    """Refines around the interfaces of seeded spheres (circles in 2D) down to
    `depth` within `band` finest cells of each interface, at least `minlevel`
    everywhere (the configurations of the adaptive benchmarks), on the unit
    box [0,1]^dim (origin 0).  Returns the centres and radii."""
    rng = np.random.default_rng(seed)
    dim = t.dim
    c = rng.uniform(0.1, 0.9, (nspheres, dim))
    r = rng.uniform(rmin, rmax, nspheres)
    h = 1.0 / (1 << depth)

    def cond(x, y, z, lev):
        p = np.stack([x, y, z][:dim], axis=1)
        want = lev < minlevel
        size = 1.0 / (1 << lev.astype(np.int64))
        near = np.zeros(len(x), bool)
        for s in range(nspheres):
            dist = np.abs(np.sqrt(((p - c[s]) ** 2).sum(1)) - r[s])
            near |= dist < band * h + size * np.sqrt(dim)
        return want | ((lev < depth) & near)
    t.refine(cond)
    return c, r


def spheres_problem(t, depth, seed, nspheres=6, minlevel=3, ratio=1000.0, lam=True):
    """A variable-density Poisson-Helmholtz problem on a tree refined around
    seeded spheres: alpha = 1/rho on faces (rho = ratio inside the spheres,
    1 outside, arithmetic face average of the two sides), optional negative
    lambda field, Dirichlet left / Neumann right / symmetry elsewhere, and a
    seeded N(0,1) right-hand side at the leaves."""
    c, r = spheres_tree(t, depth, minlevel, nspheres, seed, 0.05, 0.15)
    dim = t.dim
    rng = np.random.default_rng(seed + 100)

    def rho(p):
        inside = np.zeros(len(p), bool)
        for s in range(len(r)):
            inside |= ((p - c[s]) ** 2).sum(1) < r[s] ** 2
        return np.where(inside, ratio, 1.0)

    _, xyz = t.leaves()
    n = len(xyz)
    t.set_b(rng.standard_normal(n))
    t.set_a(np.zeros(n))
    if lam:
        t.set_lambda(-rng.uniform(0.5, 2.0, n))
    else:
        t.set_lambda(0.0)
    fxyz, fd = t.face_centres()
    f = t.faces()
    nper = np.zeros(len(f), np.int64)
    for d in range(dim):
        nper += (f[:, 4] >> d) & 1
    h = np.repeat(1.0 / (1 << f[:, 0].astype(np.int64)), nper)
    p = fxyz[:, :dim]
    left = p.copy()
    right = p.copy()
    left[np.arange(len(p)), fd] -= h / 2
    right[np.arange(len(p)), fd] += h / 2
    alpha = 2.0 / (rho(left) + rho(right))
    t.set_alpha_faces(alpha)
    cb = t.bc_face_centres(1)
    t.set_bc(1, BC_DIRICHLET, np.sin(3 * cb[:, 1]))
    cb = t.bc_face_centres(0)
    t.set_bc(0, 2, np.cos(2 * cb[:, 1]))
    return c, r
