"""Seeded synthetic inputs of the benchmark workloads.

Every generator is deterministic (numpy's PCG64 with the given seed), so the
GPU solver and the CPU oracle are fed identical bytes.  Arrays follow the
library's layout: cell fields n^dim in C order (i slowest), face component c
with n+1 entries along c.
"""
import numpy as np


def _centres(n, L0=1.0, X0=0.0):
    return (np.arange(n) + 0.5) * (L0 / n) + X0


def face_average_alpha(rho):
    """alpha on faces = 1/((rho[] + rho[-1])/2), the arithmetic face
    average of two-phase.h:96-99; wall faces see the interior rho on both
    sides.  Returns the dim face components."""
    dim = rho.ndim
    comps = []
    for c in range(dim):
        lo = np.take(rho, [0], axis=c)
        hi = np.take(rho, [rho.shape[c] - 1], axis=c)
        left = np.concatenate([lo, rho], axis=c)   # rho[-1] for faces 0..n
        right = np.concatenate([rho, hi], axis=c)  # rho[] for faces 0..n
        comps.append(1.0 / ((right + left) / 2.0))
    return comps


def divergence_rhs(u, delta):
    """b = sum over dimensions of (u.x[1] - u.x[]) / Delta, the divergence
    of project(), poisson.h:482-488 with dt = 1."""
    b = None
    for c, uc in enumerate(u):
        n = uc.shape[c] - 1
        d = (np.take(uc, range(1, n + 1), axis=c) - np.take(uc, range(0, n), axis=c))
        b = d if b is None else b + d
    return b / delta


def config1(depth=8, seed=1):
    """2D multigrid, level 8 (256^2) on the unit square, alpha = 1,
    lambda = 0, symmetry BCs; b = 16 Gaussians (centres U[0.1,0.9]^2, widths
    U[0.02,0.05], amplitudes N(0,1)) minus its mean; tolerance 1e-6."""
    rng = np.random.default_rng(seed)
    n = 1 << depth
    x = _centres(n)
    X, Y = np.meshgrid(x, x, indexing="ij")
    b = np.zeros((n, n))
    c = rng.uniform(0.1, 0.9, (16, 2))
    w = rng.uniform(0.02, 0.05, 16)
    A = rng.standard_normal(16)
    for k in range(16):
        b += A[k] * np.exp(-((X - c[k, 0]) ** 2 + (Y - c[k, 1]) ** 2) / w[k] ** 2)
    b -= b.mean()
    return {"dim": 2, "depth": depth, "b": b, "alpha": None, "tolerance": 1e-6, "nrelax": 4}


def config2(depth=8, seed=1, nspheres=24, ratio=1000.0):
    """3D variable-coefficient projection Poisson problem at level `depth`
    (256^3 by default): rho = ratio inside `nspheres` seeded spheres
    (radius r ~ U[0.03,0.1], centre ~ U[r,1-r]^3), 1 outside; alpha = 1/rho
    arithmetically averaged on faces; b = divergence of a seeded N(0,1) face
    velocity with zero wall-normal velocity (so sum(b) = 0); symmetry
    (Neumann) walls; tolerance 1e-9."""
    rng = np.random.default_rng(seed)
    n = 1 << depth
    delta = 1.0 / n
    r = rng.uniform(0.03, 0.1, nspheres)
    cen = rng.uniform(r[:, None], 1.0 - r[:, None], (nspheres, 3))
    x = _centres(n).astype(np.float64)
    inside = np.zeros((n, n, n), dtype=bool)
    X = x[:, None, None]
    Y = x[None, :, None]
    Z = x[None, None, :]
    for s in range(nspheres):
        # bounding box of the sphere keeps this cheap at 256^3
        lo = np.maximum(((cen[s] - r[s]) * n).astype(int) - 1, 0)
        hi = np.minimum(((cen[s] + r[s]) * n).astype(int) + 2, n)
        sl = tuple(slice(lo[d], hi[d]) for d in range(3))
        d2 = ((X[sl[0]] - cen[s, 0]) ** 2 + (Y[:, sl[1]] - cen[s, 1]) ** 2 + (Z[:, :, sl[2]] - cen[s, 2]) ** 2)
        inside[sl] |= d2 < r[s] ** 2
    rho = np.where(inside, ratio, 1.0)
    alpha = face_average_alpha(rho)
    u = []
    for c in range(3):
        shape = [n, n, n]
        shape[c] = n + 1
        uc = rng.standard_normal(shape)
        idx = [slice(None)] * 3
        idx[c] = 0
        uc[tuple(idx)] = 0.0
        idx[c] = n
        uc[tuple(idx)] = 0.0
        u.append(uc)
    b = divergence_rhs(u, delta)
    return {"dim": 3, "depth": depth, "b": b, "alpha": alpha, "tolerance": 1e-9, "nrelax": 4,
            "rho_inside_fraction": float(inside.mean()), "u": u, "inside": inside}


def config2_viscous(prob, seed=2, mu_in=1.0, mu_out=0.01, dt=1e-3):
    """The implicit viscous step of the config-2 two-phase field: cell
    velocity N(0,1) per component, rho = 1000 in the spheres and 1 outside,
    mu = mu_in / mu_out arithmetically averaged on faces (walls: the cell's
    value), dt = 1e-3."""
    inside = prob["inside"]
    n = inside.shape[0]
    rng = np.random.default_rng(seed)
    muc = np.where(inside, mu_in, mu_out)
    mu = []
    for c in range(3):
        shape = [n, n, n]
        shape[c] = n + 1
        f = np.empty(shape)
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[c], hi[c] = slice(0, n - 1), slice(1, n)
        mid = [slice(None)] * 3
        mid[c] = slice(1, n)
        f[tuple(mid)] = (muc[tuple(lo)] + muc[tuple(hi)]) / 2.0
        e0, e1, c0, c1 = [slice(None)] * 3, [slice(None)] * 3, [slice(None)] * 3, [slice(None)] * 3
        e0[c], e1[c], c0[c], c1[c] = 0, n, 0, n - 1
        f[tuple(e0)] = muc[tuple(c0)]
        f[tuple(e1)] = muc[tuple(c1)]
        mu.append(f)
    u = [rng.standard_normal((n, n, n)) for _ in range(3)]
    rho = np.where(inside, 1000.0, 1.0)
    return {"u": u, "mu": mu, "rho": rho, "dt": dt}


def config4(depth=10, seed=4):
    """2D adjoint problem: log-normal face alpha (sigma 0.5) and a seeded
    32-Gaussian source on homogeneous Dirichlet walls (well posed, so J is
    defined without a free constant).  The target phi_t of
    J = 0.5 ||phi - phi_t||^2 is the solution for the same source with
    alpha = 1."""
    rng = np.random.default_rng(seed)
    n = 1 << depth
    alpha = [np.exp(0.5 * rng.standard_normal((n + 1, n) if c == 0 else (n, n + 1))) for c in range(2)]
    x = _centres(n)
    X, Y = np.meshgrid(x, x, indexing="ij")
    src = np.zeros((n, n))
    c = rng.uniform(0.1, 0.9, (32, 2))
    w = rng.uniform(0.02, 0.05, 32)
    A = rng.standard_normal(32)
    for k in range(32):
        src += A[k] * np.exp(-((X - c[k, 0]) ** 2 + (Y - c[k, 1]) ** 2) / w[k] ** 2)
    src -= src.mean()
    return {"dim": 2, "depth": depth, "alpha": alpha, "source": src, "tolerance": 1e-9,
            "bc": {side: (1, None) for side in range(4)}}


# --- adaptive trees (configs 3 and 5) --------------------------------------

def shell_refine(dim, depth, minlevel, centres, radii, band=2.0):
    """refine() condition of the adaptive workloads: every cell up to
    `minlevel`, and down to `depth` every cell whose centre lies within
    band finest cells plus its own half-diagonal of a sphere's surface.
    The cells of each level are enumerated per sphere over its bounding box,
    so one pass costs O(shell cells), not O(leaves x spheres)."""
    h = 1.0 / (1 << depth)
    marks = {}

    def level_keys(lev):
        if lev in marks:
            return marks[lev]
        n = 1 << lev
        size = 1.0 / n
        tol = band * h + size * np.sqrt(dim)
        keys = []
        for c, r in zip(centres, radii):
            lo = np.maximum(np.floor((c - r - tol) * n - 0.5).astype(np.int64), 0)
            hi = np.minimum(np.ceil((c + r + tol) * n - 0.5).astype(np.int64), n - 1)
            axes = [np.arange(lo[d], hi[d] + 1) for d in range(dim)]
            grids = np.meshgrid(*axes, indexing="ij")
            d2 = np.zeros(grids[0].shape)
            for d in range(dim):
                d2 += ((grids[d] + 0.5) * size - c[d]) ** 2
            sel = np.abs(np.sqrt(d2) - r) < tol
            key = np.zeros(int(sel.sum()), np.int64)
            for d in range(dim):
                key = key * n + grids[d][sel]
            keys.append(key)
        marks[lev] = np.unique(np.concatenate(keys)) if keys else np.zeros(0, np.int64)
        return marks[lev]

    def cond(x, y, z, lev):
        lev = np.asarray(lev)
        out = lev < minlevel
        p = [x, y, z][:dim]
        for l in np.unique(lev):
            if l >= depth or l < minlevel:
                continue
            sel = lev == l
            n = 1 << int(l)
            key = np.zeros(int(sel.sum()), np.int64)
            for d in range(dim):
                key = key * n + np.floor(p[d][sel] * n).astype(np.int64)
            k = level_keys(int(l))
            pos = np.searchsorted(k, key)
            pos[pos >= len(k)] = 0
            out[np.nonzero(sel)[0]] = (len(k) > 0) & (k[pos] == key)
        return out
    return cond


def inside_spheres(q, centres, radii, nbins=16):
    """Whether each point of q (n, dim) lies inside any sphere; the points
    are binned on an nbins^dim grid so each sphere tests only the points of
    the bins its bounding box covers."""
    dim = q.shape[1]
    b = np.clip((q * nbins).astype(np.int64), 0, nbins - 1)
    lin = np.zeros(len(q), np.int64)
    for d in range(dim):
        lin = lin * nbins + b[:, d]
    order = np.argsort(lin, kind="stable")
    bounds = np.searchsorted(lin[order], np.arange(nbins ** dim + 1))
    inside = np.zeros(len(q), bool)
    for c, r in zip(centres, radii):
        lo = np.clip(((c - r) * nbins).astype(np.int64), 0, nbins - 1)
        hi = np.clip(((c + r) * nbins).astype(np.int64), 0, nbins - 1)
        for cell in np.ndindex(*[hi[d] - lo[d] + 1 for d in range(dim)]):
            key = 0
            for d in range(dim):
                key = key * nbins + lo[d] + cell[d]
            idx = order[bounds[key]:bounds[key + 1]]
            if len(idx):
                inside[idx] |= ((q[idx] - c) ** 2).sum(1) < r * r
    return inside


def _tree_rho_alpha(t, centres, radii, ratio):
    """alpha = 1/rho on the tree's faces (face-cache order), arithmetic
    average of rho on the two sides of each face, rho = ratio inside the
    spheres and 1 outside."""
    dim = t.dim
    xyz, dd = t.face_centres()
    f = t.faces()
    nper = np.zeros(len(f), np.int64)
    for d in range(dim):
        nper += (f[:, 4] >> d) & 1
    h = np.repeat(1.0 / (1 << f[:, 0].astype(np.int64)), nper)
    p = xyz[:, :dim]

    def rho(q):
        return np.where(inside_spheres(q, centres, radii), ratio, 1.0)
    left = p.copy()
    right = p.copy()
    idx = np.arange(len(p))
    left[idx, dd] -= h / 2
    right[idx, dd] += h / 2
    return 1.0 / ((rho(left) + rho(right)) / 2.0)


def config3_tree(t, seed=3, depth=11, minlevel=5, ncircles=40, ratio=1000.0, band=3.0):
    """2D quadtree refined to `depth` around the interfaces of seeded
    circles (radii U[0.01, 0.05]) from level `minlevel`; alpha = 1/rho
    (density ratio 1000), lambda = 0; b = the 16 seeded Gaussians of config 1
    at the leaf centres, minus their volume-weighted mean (symmetry walls).
    `t` is an empty tree of the unit square (basilisk_amd.Tree or the
    oracle's)."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.1, 0.9, (ncircles, 2))
    radii = rng.uniform(0.01, 0.05, ncircles)
    t.refine(shell_refine(2, depth, minlevel, centres, radii, band))
    lev, xyz = t.leaves()
    g1 = np.random.default_rng(1)
    c = g1.uniform(0.1, 0.9, (16, 2))
    w = g1.uniform(0.02, 0.05, 16)
    A = g1.standard_normal(16)
    b = np.zeros(len(lev))
    for k in range(16):
        b += A[k] * np.exp(-((xyz[:, 0] - c[k, 0]) ** 2 + (xyz[:, 1] - c[k, 1]) ** 2) / w[k] ** 2)
    vol = (1.0 / (1 << lev.astype(np.int64))) ** 2
    b -= (b * vol).sum() / vol.sum()
    t.set_b(b)
    t.set_a(np.zeros(len(b)))
    t.set_lambda(0.0)
    t.set_alpha_faces(_tree_rho_alpha(t, centres, radii, ratio))
    return {"tolerance": 1e-9, "leaves": len(b)}


def config5_tree(t, seed=5, depth=10, minlevel=5, nspheres=200, ratio=1000.0, band=2.0):
    """3D octree refined to `depth` around seeded spheres (radii
    U[0.005, 0.03]) from level `minlevel`; alpha = 1/rho (density ratio
    1000); lambda = -1e-2 / h_min^2 (the Helmholtz shift of an implicit
    step); b seeded N(0,1) at the leaves; symmetry walls."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.1, 0.9, (nspheres, 3))
    radii = rng.uniform(0.005, 0.03, nspheres)
    t.refine(shell_refine(3, depth, minlevel, centres, radii, band))
    n = t.nleaves()
    t.set_b(rng.standard_normal(n))
    t.set_a(np.zeros(n))
    hmin = 1.0 / (1 << depth)
    t.set_lambda(-1e-2 / hmin ** 2)
    t.set_alpha_faces(_tree_rho_alpha(t, centres, radii, ratio))
    return {"tolerance": 1e-9, "leaves": n}
