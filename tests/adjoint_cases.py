"""Adjoint (coefficient-gradient) set-ups shared by the CPU and GPU tests.

J(phi) = 0.5 * sum (phi - phi_t)^2 over the finest-level cells, with phi the
solution of div(alpha grad phi) + lambda phi = b.  The CPU adjoint here
(the oracle's solver plus the face formula in numpy) is the reference the GPU
gradients are checked against; the finite-difference tests pin it.
"""
import numpy as np

from tests.cases import BC_DIRICHLET, BC_NEUMANN, BC_SYMMETRY

RIGHT, LEFT = 0, 1


def problem(dim, depth, seed, lam=True):
    """Log-normal face alpha (sigma 0.5), optional negative lambda field,
    N(0,1) b, Dirichlet on the left side and Neumann on the right side
    (seeded boundary values), symmetry elsewhere, and a seeded target."""
    rng = np.random.default_rng(seed)
    n = 1 << depth
    alpha = []
    for c in range(dim):
        s = [n] * dim
        s[c] += 1
        alpha.append(np.exp(0.5 * rng.standard_normal(s)))
    lamf = -rng.uniform(0.2, 1.0, (n,) * dim) if lam else None
    b = rng.standard_normal((n,) * dim)
    phit = 0.01 * rng.standard_normal((n,) * dim)
    bcs = {LEFT: (BC_DIRICHLET, rng.standard_normal(n ** (dim - 1))),
           RIGHT: (BC_NEUMANN, rng.standard_normal(n ** (dim - 1)))}
    return {"dim": dim, "depth": depth, "alpha": alpha, "lam": lamf, "b": b, "phit": phit, "bcs": bcs}


def full_extent(dim, n, v):
    """n^(dim-1) boundary values -> the (n+4)^(dim-1) array the oracle takes."""
    if dim == 1:
        return np.asarray(v, dtype=float).reshape(1)
    out = np.zeros((n + 4,) * (dim - 1))
    out[(slice(2, n + 2),) * (dim - 1)] = np.asarray(v).reshape((n,) * (dim - 1))
    return out


def setup(g, prob, rhs, homogeneous=False):
    """Coefficients, BCs (homogeneous for the adjoint) and rhs on a grid with
    the OracleGrid / basilisk_amd.Grid interface."""
    dim, n = prob["dim"], 1 << prob["depth"]
    g.set_alpha(prob["alpha"])
    g.set_lambda(prob["lam"] if prob["lam"] is not None else 0.0)
    for side, (kind, v) in prob["bcs"].items():
        g.set_bc(side, kind, None if homogeneous else full_extent(dim, n, v))
    g.set_b(rhs)
    g.set_a(np.zeros(rhs.shape))


def face_gradient(prob, a, psi):
    """dJ/dalpha on every face and dJ/dlambda, in the GPU kernel's order of
    operations (interior: ((psi_r - psi_l)/D) ((a_r - a_l)/D); wall:
    -psi (((c + sigma a) - a)/D)/D with the BC as c + sigma a)."""
    dim, n = prob["dim"], 1 << prob["depth"]
    D = 1.0 * (1.0 / n)
    out = []
    for c in range(dim):
        s = [n] * dim
        s[c] += 1
        g = np.empty(s)
        lo = [slice(None)] * dim
        hi = [slice(None)] * dim
        mid = [slice(None)] * dim
        lo[c], hi[c], mid[c] = slice(0, n - 1), slice(1, n), slice(1, n)
        lo, hi, mid = tuple(lo), tuple(hi), tuple(mid)
        g[mid] = ((psi[hi] - psi[lo]) / D) * ((a[hi] - a[lo]) / D)
        for side, face, cell in ((2 * c + 1, 0, 0), (2 * c, n, n - 1)):
            kind, v = prob["bcs"].get(side, (BC_SYMMETRY, None))
            fi = [slice(None)] * dim
            ci = [slice(None)] * dim
            fi[c], ci[c] = face, cell
            ac, pc = a[tuple(ci)], psi[tuple(ci)]
            if kind == BC_DIRICHLET:
                cval, sigma = 2. * np.asarray(v).reshape(ac.shape), -1.0
            elif kind == BC_NEUMANN:
                cval, sigma = D * np.asarray(v).reshape(ac.shape), 1.0
            else:
                cval, sigma = 0.0, 1.0
            ghost = cval + sigma * ac
            g[tuple(fi)] = -pc * (((ghost - ac) / D) / D)
        out.append(g)
    glam = -psi * a
    return out, glam


def objective(phi, phit):
    return 0.5 * float(((phi - phit) ** 2).sum())
