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
            "rho_inside_fraction": float(inside.mean())}


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
