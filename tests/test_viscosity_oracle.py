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
  with the reference's in-place Gauss-Seidel (the serial i-major loop), the
  Jacobi relax and the red-black Gauss-Seidel extension; the serial loop's
  per-cell arithmetic equals the vectorised stencil's bit for bit.
- on periodic dimensions (grid/multigrid.h:600-701) the residual commutes
  with a cyclic shift of the fields, converges at second order to the
  continuous operator in every cell, corners between a periodic and a wall
  direction included, and the solve still matches the direct solution.
The per-cycle iterates have no reference output and are parity unpinned
beyond these checks; so is the level-0 rule on periodic grids (every ghost
of the single root cell is the root value, multigrid.h:655-664), which only
shapes the coarsest correction and cancels out of the converged solution.
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


def _periodic_faces(mu, periodic):
    """Face n of a periodic dimension is face 0 (the same face)."""
    out = []
    for d, m in enumerate(mu):
        m = m.copy()
        if periodic[d]:
            sl0 = [slice(None)] * m.ndim
            sln = [slice(None)] * m.ndim
            sl0[d], sln[d] = 0, -1
            m[tuple(sln)] = m[tuple(sl0)]
        out.append(m)
    return out


@pytest.mark.parametrize("dim,depth,sigma,periodic", [
    (2, 5, None, None), (2, 4, "noslip", None), (3, 3, None, None),
    (2, 4, None, (True, False)), (2, 4, "noslip", (False, True)), (2, 4, None, (True, True)),
    (3, 3, None, (True, False, True)), (3, 3, None, (True, True, True))])
@pytest.mark.parametrize("smoother", ["jacobi", "rbgs", "gs"])
def test_solve_matches_direct(dim, depth, sigma, periodic, smoother):
    V = Viscosity(dim, depth)
    V.smoother = smoother
    n = 1 << depth
    rng = np.random.default_rng(3 * dim + depth)
    if sigma == "noslip":
        V.sigma = -np.ones((dim, 2 * dim), np.int64)
    if periodic:
        V.periodic = list(periodic)
    V.set_coefficients(_periodic_faces(_faces(rng, dim, n), V.periodic), rng.uniform(0.5, 2.0, (n,) * dim))
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


def _roll_faces(mu, d, k):
    """Shift face arrays by k cells along the periodic dimension d."""
    out = []
    for e, m in enumerate(mu):
        if e == d:
            sl = [slice(None)] * m.ndim
            sl[d] = slice(0, -1)
            r = np.roll(m[tuple(sl)], k, axis=d)
            out.append(np.concatenate([r, np.take(r, [0], axis=d)], axis=d))
        else:
            out.append(np.roll(m, k, axis=d))
    return out


@pytest.mark.parametrize("dim,depth,periodic", [(2, 4, (True, False)), (2, 4, (False, True)),
                                                (3, 3, (True, False, True)), (3, 3, (False, True, False))])
def test_periodic_residual_commutes_with_shift(dim, depth, periodic):
    n = 1 << depth
    rng = np.random.default_rng(11 * dim + depth)
    mu = _periodic_faces(_faces(rng, dim, n), periodic)
    rho = rng.uniform(0.5, 2.0, (n,) * dim)
    u = [rng.standard_normal((n,) * dim) for _ in range(dim)]
    r = [rng.standard_normal((n,) * dim) for _ in range(dim)]
    V = Viscosity(dim, depth)
    V.periodic = list(periodic)
    V.set_coefficients(mu, rho)
    res, m = V.residual(u, r, 0.01)
    for d in range(dim):
        if not periodic[d]:
            continue
        for k in (1, 3):
            W = Viscosity(dim, depth)
            W.periodic = list(periodic)
            W.set_coefficients(_roll_faces(mu, d, k), np.roll(rho, k, axis=d))
            rs, ms = W.residual([np.roll(x, k, axis=d) for x in u], [np.roll(x, k, axis=d) for x in r], 0.01)
            assert ms == m
            for c in range(dim):
                assert np.array_equal(rs[c], np.roll(res[c], k, axis=d)), (d, k, c)
    # a wall dimension treated as periodic does not commute: the test can fail
    W = Viscosity(dim, depth)
    W.set_coefficients(mu, rho)
    rw, _ = W.residual(u, r, 0.01)
    assert max(np.abs(rw[c] - res[c]).max() for c in range(dim)) > 1e-3


def _trig_field(dim, periodic, x):
    """A smooth velocity whose extension across every boundary matches the
    ghost rule: periodic along periodic dimensions; across a wall, odd for
    the normal component (sigma -1) and even for the tangential ones.
    u_c = prod_a g_ca(x_a); returns, per component, a function of the
    derivative orders (one per dimension) by the product rule."""
    tp = 2 * np.pi
    comps = []
    for c in range(dim):
        g = []
        for a in range(dim):
            y = x[a]
            if periodic[a]:
                k, ph = tp * (1 + (a + c) % 2), 0.3 * (c + 1) + 0.7 * a
                g.append([np.sin(k * y + ph), k * np.cos(k * y + ph), -k * k * np.sin(k * y + ph)])
            elif a == c:
                k = np.pi
                g.append([np.sin(k * y), k * np.cos(k * y), -k * k * np.sin(k * y)])
            else:
                k = tp
                g.append([np.cos(k * y), -k * np.sin(k * y), -k * k * np.cos(k * y)])

        def part(orders, g=g):
            p = 1.0
            for a in range(dim):
                p = p * g[a][orders[a]]
            return p
        comps.append(part)
    return comps


@pytest.mark.parametrize("dim,periodic", [(2, (True, False)), (2, (True, True)), (2, (False, True)),
                                          (3, (True, False, True)), (3, (True, True, True))])
def test_periodic_residual_second_order_everywhere(dim, periodic):
    """residual_viscosity of a smooth field against u - dt/rho mu
    (lap u + grad div u): the error falls by ~4 per refinement in every
    cell, so the wrapped ghosts and the wall/periodic corners are right."""
    mu, rho, dt = 0.7, 1.3, 0.01
    errs = []
    for depth in ((4, 5, 6) if dim == 2 else (4, 5)):
        n = 1 << depth
        x = _cells(dim, n)
        P = _trig_field(dim, periodic, x)
        u = [P[c]((0,) * dim) for c in range(dim)]

        def second(c, a, b):
            o = [0] * dim
            o[a] += 1
            o[b] += 1
            return P[c](tuple(o))

        lap_plus = [sum(second(c, a, a) for a in range(dim)) + sum(second(a, a, c) for a in range(dim))
                    for c in range(dim)]
        r = [u[c] - dt / rho * mu * lap_plus[c] for c in range(dim)]
        V = Viscosity(dim, depth)
        V.periodic = list(periodic)
        V.set_coefficients(mu, rho)
        _, m = V.residual(u, r, dt)
        errs.append(m)
    for a, b in zip(errs, errs[1:]):
        assert 3.5 < a / b < 4.5, errs


# -- AXI (viscosity.h:12-14): x the axis, y the radius, with the metric
# axi.h gives the caller's fields: mu on x faces times the cell's y, on y
# faces times the face's y; rho times the cell's y

def _axi_problem(depth, dt, plant=0):
    """u = sin(a x) y^2, v = cos(a x) y^3 (mu = 1) and the right-hand side
    r = u - dt div(2 D(u)) of the axisymmetric operator, worked out by hand:
      (div tau)_x = sin(a x) (4 - 4 a y^2 - 2 a^2 y^2)
      (div tau)_r = cos(a x) (16 y + 2 a y - a^2 y^3)
    (tau_xx = 2 a cos y^2, tau_rr = 6 cos y^2, tau_xr = sin (2 y - a y^3),
    tau_thetatheta = 2 v / y: the hoop term -tau_thetatheta / y)."""
    n = 1 << depth
    D = 1.0 / n
    V = Viscosity(2, depth, axi=True, plant=plant)
    xc = (np.arange(n) + 0.5) * D
    X, Y = np.meshgrid(xc, xc, indexing="ij")
    yf = np.arange(n + 1) * D
    mux = np.tile(xc[None, :], (n + 1, 1))    # x faces: the cell's y
    muy = np.tile(yf[None, :], (n, 1))        # y faces: the face's y
    V.set_coefficients([mux, muy], Y.copy())  # rho = cm = y
    a = 2.0
    u = np.sin(a * X) * Y ** 2
    v = np.cos(a * X) * Y ** 3
    divx = np.sin(a * X) * (4 - 4 * a * Y ** 2 - 2 * a * a * Y ** 2)
    divr = np.cos(a * X) * (16 * Y + 2 * a * Y - a * a * Y ** 3)
    r = [u - dt * divx, v - dt * divr]
    return V, [u, v], r, X, Y


@pytest.mark.parametrize("plant", [0, 1])
def test_axi_residual_second_order(plant):
    """The residual of the exact solution against the hand-derived
    axisymmetric operator falls at second order in the cells away from the
    walls and the axis; the AXI term with its sign flipped (plant 1) leaves
    an O(1) residual."""
    dt = 0.01
    errs = []
    for depth in (5, 6, 7):
        V, U, r, X, Y = _axi_problem(depth, dt, plant)
        res, _ = V.residual(U, r, dt)
        n = 1 << depth
        m = (X > 2.5 / n) & (X < 1 - 2.5 / n) & (Y > 0.25) & (Y < 1 - 2.5 / n)
        errs.append(max(np.abs(res[c][m]).max() for c in range(2)))
    if plant:
        assert errs[-1] > 1e-4
    else:
        assert errs[0] / errs[1] > 3.5 and errs[1] / errs[2] > 3.5, errs
        assert errs[-1] < 1e-5


@pytest.mark.parametrize("smoother", ["jacobi", "rbgs"])
def test_axi_relax_matches_residual_and_solve_matches_direct(smoother):
    """relax and residual share the AXI lambda (the Jacobi target satisfies
    w - u = Delta^2 res / diag with diag = Delta^2 lambda + dt/rho sum mu),
    and mg_solve converges to the direct solution of the same operator."""
    dt = 0.01
    V, U, r, X, Y = _axi_problem(5, dt)
    n = 32
    D2 = (1.0 / n) ** 2
    res, _ = V.residual(U, r, dt)
    V.smoother = "jacobi"
    new = V.relax(U, r, 5, dt)
    for c in range(2):
        w = (3.0 * new[c] - U[c]) / 2.0
        diag = D2 * V._lambda(c, 5, n, dt) + dt / V.rho[5] * V._diag(c, 5, n)
        np.testing.assert_allclose(w - U[c], D2 * res[c] / diag, rtol=1e-9, atol=1e-12)
    V.smoother = smoother
    V.TOLERANCE = 1e-12
    u, s, hist = V.solve(r, dt)
    assert s["resa"] <= 1e-12
    J = _operator(V, dt)
    direct = scipy.sparse.linalg.spsolve(scipy.sparse.csr_matrix(J), np.concatenate([q.reshape(-1) for q in r]))
    got = np.concatenate([q.reshape(-1) for q in u])
    assert np.abs(got - direct).max() < 1e-10 * max(1.0, np.abs(direct).max())


@pytest.mark.parametrize("dim,depth,axi", [(2, 4, False), (2, 4, True), (3, 3, False)])
def test_gs_cell_arithmetic_equals_the_vectorised_stencil(dim, depth, axi):
    """The serial Gauss-Seidel loop evaluates relax_viscosity one cell at a
    time; on a state it has not yet modified (the first cell) and on every
    cell of a fixed state its w equals the vectorised Jacobi target bit for
    bit, so the loop differs from the pinned stencil only in which values it
    reads."""
    rng = np.random.default_rng(dim + depth)
    n = 1 << depth
    V = Viscosity(dim, depth, axi=axi, Y0=0.3 if axi else 0.0)
    V.set_coefficients(_faces(rng, dim, n), rng.uniform(0.5, 2.0, (n,) * dim))
    dt = 0.01
    u = [rng.standard_normal((n,) * dim) for _ in range(dim)]
    r = [rng.standard_normal((n,) * dim) for _ in range(dim)]
    from oracle.viscosity import _pad
    P = [_pad(u[c], V.sigma[c], dim, V.periodic) for c in range(dim)]
    w = V._target(u, r, depth, dt)
    D2 = V.delta(depth) ** 2
    for at in np.ndindex(*(n,) * dim):
        dtr = V._dtr(depth, dt, at)
        for c in range(dim):
            wc = (dtr * V._terms(P, c, depth, n, False, at) + r[c][at] * D2) / \
                (D2 * V._lambda(c, depth, n, dt, at) + dtr * V._diag(c, depth, n, at))
            assert wc == w[c][at]
    # one GS sweep: the first cell in i-major order takes w, the last one
    # differs from it (it reads updated neighbours)
    V.smoother = "gs"
    new = V.relax(u, r, depth, dt)
    first, last = (0,) * dim, (n - 1,) * dim
    assert all(new[c][first] == w[c][first] for c in range(dim))
    assert any(new[c][last] != w[c][last] for c in range(dim))


@pytest.mark.parametrize("dim,depth", [(2, 5), (3, 4)])
def test_explicit_step_of_a_quadratic(dim, depth):
    """viscosity_explicit (viscosity.h:180-190) adds residual_viscosity
    (u, u): for a quadratic field with constant mu and rho, in the interior
    exactly dt/rho mu (lap u + grad div u)."""
    V = Viscosity(dim, depth)
    n = 1 << depth
    mu, rho, dt = 0.7, 1.3, 0.01
    V.set_coefficients(mu, rho)
    x = _cells(dim, n)
    rng = np.random.default_rng(10 + dim)
    A = rng.standard_normal((dim, dim, dim))
    u = [sum(A[c, a, b] * x[a] * x[b] for a in range(dim) for b in range(dim)) for c in range(dim)]
    H = A + A.transpose(0, 2, 1)
    new, st = V.explicit(u, dt)
    inner = tuple(slice(2, n - 2) for _ in range(dim))
    for c in range(dim):
        du = dt / rho * mu * (sum(H[c, a, a] for a in range(dim)) + sum(H[a, a, c] for a in range(dim)))
        assert np.abs((new[c] - u[c] - du)[inner]).max() < 1e-11
    assert st["resb"] > 0 and st["i"] == 0
