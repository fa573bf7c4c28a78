"""centered_gradient and correction of navier-stokes/centered.h in the
oracle (oracle.centered_gradient, oracle.correction), pinned by what the
mathematics fixes -- the reference has no fixture for them alone:

- hydrostatic balance: a constant face acceleration A balanced by the
  pressure p = A x / alpha with the walls' neumann_pressure conditions
  (centered.h:84-111) gives g = 0 exactly, walls included (the "exact
  balance" centered.h's comment asks of the discretisation);
- a linear pressure with matching Dirichlet walls gives g = -alpha grad p
  exactly in every cell, and a quadratic one -dp/dx to rounding;
- with no pressure, g is the average of the two face accelerations;
- correction adds dt g.
Each check fails on a planted mistake (sign of the pressure term, wrong
face averaged)."""
import numpy as np
import pytest

from oracle import oracle as orc
from oracle.oracle import OracleGrid
from tests import cases


def _full(n, dim, v):
    """a constant BC value over the oracle's ghost-inclusive extent"""
    return np.full((n + 4) ** (dim - 1), v)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("plant", [0, 1])
def test_hydrostatic_balance(dim, plant):
    depth, A, alpha = 4, 0.75, 0.5
    o = OracleGrid(dim, depth)
    n, D = o.n(), o.delta()
    X = o.centers()[0]
    p = np.broadcast_to((A / alpha * X).reshape((n,) + (1,) * (dim - 1)), (n,) * dim).copy()
    o.set_a(p)
    # neumann_pressure: p[right] = neumann(a.n*fm.n/alpha.n), p[left] = neumann(-a.n*fm.n/alpha.n)
    o.set_bc(0, cases.BC_NEUMANN, _full(n, dim, A / alpha))
    o.set_bc(1, cases.BC_NEUMANN, _full(n, dim, -A / alpha))
    alpha_f = [np.full(o.face_shape(c), alpha) for c in range(dim)]
    acc = [np.full(o.face_shape(c), A if c == 0 else 0.0) for c in range(dim)]
    g = orc.centered_gradient(o, alpha_f, acc, plant=plant)
    ok = all(np.array_equal(x, np.zeros_like(x)) for x in g)
    assert ok == (plant == 0)


@pytest.mark.parametrize("plant", [0, 1, 2])
def test_linear_and_quadratic_pressure(plant):
    dim, depth = 2, 5
    o = OracleGrid(dim, depth, origin=(-0.5, -0.5, 0.0))
    n = o.n()
    X, Y, _ = cases.mesh(o)
    X, Y = X[..., 0], Y[..., 0]
    # p = 3x - 2y with its own values at the walls: every face difference is exact
    o.set_bc_function(cases.BC_DIRICHLET, lambda x, y, z: 3.0 * x - 2.0 * y)
    o.set_a(3.0 * X - 2.0 * Y)
    g = orc.centered_gradient(o, plant=plant)
    lin_ok = np.allclose(g[0], -3.0, rtol=0, atol=1e-12) and np.allclose(g[1], 2.0, rtol=0, atol=1e-12)
    # p = x^2: interior cells give -(p(x+D) - p(x-D))/(2D) = -2x
    o.set_bc_function(cases.BC_DIRICHLET, lambda x, y, z: x * x)
    o.set_a(X * X)
    g = orc.centered_gradient(o, plant=plant)
    quad_ok = np.allclose(g[0][1:-1], -2.0 * X[1:-1], rtol=0, atol=1e-11) and np.allclose(g[1], 0.0, atol=1e-12)
    assert (lin_ok and quad_ok) == (plant == 0)


@pytest.mark.parametrize("plant", [0, 2])
def test_acceleration_only_and_correction(plant):
    dim, depth = 3, 3
    o = OracleGrid(dim, depth)
    o.set_a(np.zeros(o.shape()))
    rng = np.random.default_rng(4)
    acc = [rng.standard_normal(o.face_shape(c)) for c in range(dim)]
    g = orc.centered_gradient(o, acc=acc, plant=plant)
    ok = True
    for c in range(dim):
        lo = [slice(None)] * dim
        hi = [slice(None)] * dim
        lo[c], hi[c] = slice(0, -1), slice(1, None)
        ok &= np.array_equal(g[c], (acc[c][tuple(lo)] + acc[c][tuple(hi)]) / 2.0)
    assert ok == (plant == 0)
    if plant:
        return
    u = [rng.standard_normal(o.shape()) for _ in range(dim)]
    v = orc.correction(u, g, 0.25)
    for c in range(dim):
        assert np.array_equal(v[c], u[c] + 0.25 * g[c])
