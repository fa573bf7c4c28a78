"""The viscosity() oracle (oracle/viscosity.py) pinned by the mathematics
(no GPU needed).  The reference has no fixture for viscosity() alone (its
logs come from whole Navier-Stokes runs), so:

- residual_viscosity of a quadratic field with constant mu and rho is exact
  in the interior: the discrete operator equals u - dt/rho mu (lap u +
  grad div u) there (a wrong sign, index or factor in any term breaks it);
- with constant mu the operator, walls included, is symmetric;
- relax_viscosity and residual_viscosity describe the same operator: the
  Jacobi target satisfies w - u = Delta^2 res / diag;
- mg_solve converges to the direct sparse solution of the same operator,
  with the Jacobi relax and with the red-black Gauss-Seidel extension.
The per-cycle iterates have no reference output and are parity unpinned
beyond these checks.
"""
import numpy as np
import pytest
import scipy.sparse.linalg

from oracle.viscosity import Viscosity, default_sigma


def _cells(dim, n):
    return (np.indices((n,) * dim) + 0.5) / n


def _faces(rng, dim, n, lo=0.5, hi=2.0):
    return [rng.uniform(lo, hi, tuple(n + (1 if e == d else 0) for e in range(dim))) for d in range(dim)]


@pytest.mark.parametrize("dim,depth", [(2, 5), (3, 4)])
def test_quadratic_exact_in_interior(dim, depth):
    V = Viscosity(dim, depth)
    n = 1 << depth
    mu, rho, dt = 0.7, 1.3, 0.01
    V.set_coefficients(mu, rho)
    x = _cells(dim, n)
    rng = np.random.default_rng(dim)
    A = rng.standard_normal((dim, dim, dim))
    B = rng.standard_normal((dim, dim))
    u = [sum(A[c, a, b] * x[a] * x[b] for a in range(dim) for b in range(dim))
         + sum(B[c, a] * x[a] for a in range(dim)) for c in range(dim)]
    H = A + A.transpose(0, 2, 1)  # second derivatives
    r = [u[c] - dt / rho * mu * (sum(H[c, a, a] for a in range(dim)) + sum(H[a, a, c] for a in range(dim)))
         for c in range(dim)]
    res, m = V.residual(u, r, dt)
    inner = tuple(slice(2, n - 2) for _ in range(dim))
    assert max(np.abs(res[c][inner]).max() for c in range(dim)) < 1e-11
    assert m > 1.0  # the walls (mirror ghosts) are not polynomial


def _operator(V, dt):
    dim, n = V.dim, 1 << V.depth
    N = n ** dim * dim
    J = np.zeros((N, N))
    z = [np.zeros((n,) * dim) for _ in range(dim)]
    for k in range(N):
        uu = [np.zeros((n,) * dim) for _ in range(dim)]
        c, i = divmod(k, n ** dim)
        uu[c].reshape(-1)[i] = 1.0
        rr, _ = V.residual(uu, z, dt)
        J[:, k] = -np.concatenate([q.reshape(-1) for q in rr])
    return J


@pytest.mark.parametrize("dim,depth", [(2, 4), (3, 3)])
def test_constant_mu_operator_symmetric(dim, depth):
    V = Viscosity(dim, depth)
    V.set_coefficients(0.9, 1.0)
    J = _operator(V, 0.02)
    assert np.abs(J - J.T).max() == 0.0
    # and positive definite (lambda = 1 plus a viscous part)
    assert np.linalg.eigvalsh(J).min() > 0.5


@pytest.mark.parametrize("dim,depth", [(2, 5), (3, 4)])
def test_relax_matches_residual(dim, depth):
    V = Viscosity(dim, depth)
    n = 1 << depth
    rng = np.random.default_rng(7 + dim)
    V.set_coefficients(_faces(rng, dim, n), rng.uniform(0.5, 2.0, (n,) * dim))
    dt = 0.003
    u = [rng.standard_normal((n,) * dim) for _ in range(dim)]
    r = [rng.standard_normal((n,) * dim) for _ in range(dim)]
    res, _ = V.residual(u, r, dt)
    new = V.relax(u, r, depth, dt)  # (u + 2 w)/3
    D2 = (1.0 / n) ** 2
    for c in range(dim):
        w = (3.0 * new[c] - u[c]) / 2.0
        diag = D2 + dt / V.rho[depth] * V._diag(c, depth, n)
        np.testing.assert_allclose(w - u[c], D2 * res[c] / diag, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("dim,depth,sigma", [(2, 5, None), (2, 4, "noslip"), (3, 3, None)])
@pytest.mark.parametrize("smoother", ["jacobi", "rbgs"])
def test_solve_matches_direct(dim, depth, sigma, smoother):
    V = Viscosity(dim, depth)
    V.smoother = smoother
    n = 1 << depth
    rng = np.random.default_rng(3 * dim + depth)
    if sigma == "noslip":
        V.sigma = -np.ones((dim, 2 * dim), np.int64)
    V.set_coefficients(_faces(rng, dim, n), rng.uniform(0.5, 2.0, (n,) * dim))
    dt = 0.01
    V.TOLERANCE = 1e-12
    u0 = [rng.standard_normal((n,) * dim) for _ in range(dim)]
    u, s, hist = V.solve(u0, dt)
    assert s["resa"] <= 1e-12 and s["i"] < 100
    assert all(b < a for a, b in zip(hist, hist[1:]))
    J = _operator(V, dt)
    direct = scipy.sparse.linalg.spsolve(scipy.sparse.csr_matrix(J), np.concatenate([q.reshape(-1) for q in u0]))
    got = np.concatenate([q.reshape(-1) for q in u])
    assert np.abs(got - direct).max() < 1e-10 * max(1.0, np.abs(direct).max())


def test_default_sigma():
    s = default_sigma(3)
    assert s.tolist() == [[-1, -1, 1, 1, 1, 1], [1, 1, -1, -1, 1, 1], [1, 1, 1, 1, -1, -1]]
