"""project() in the oracle (poisson.h:469-509), pinned by what the
mathematics fixes (no reference fixture exercises it on its own):

- the face update and the divergence are exact discrete adjoints of the
  residual's flux form, so after projection the divergence of uf divided by
  dt equals the final residual cell by cell (lambda = 0), walls and
  inhomogeneous boundary values included;
- a face field that is already a discrete gradient, dt * grad q with the
  boundary conditions of p, is removed entirely (p = q + const, since
  div(alpha grad p) = div(uf)/dt, and uf -> 0 to the solver tolerance);
- dt = 0 and nrelax = 0 take the reference's defaults (1 and 4).
"""
import numpy as np
import pytest

from oracle.oracle import OracleGrid
from tests import cases


def _random_uf(o, rng, zero_walls=True):
    n, uf = o.n(), []
    for c in range(o.dim):
        u = rng.standard_normal(o.face_shape(c))
        if zero_walls:
            idx = [slice(None)] * o.dim
            idx[c] = 0
            u[tuple(idx)] = 0.0
            idx[c] = n
            u[tuple(idx)] = 0.0
        uf.append(np.ascontiguousarray(u))
    return uf


def _div(uf, delta):
    return sum(np.diff(u, axis=c) for c, u in enumerate(uf)) / delta


@pytest.mark.parametrize("dim,depth", [(2, 6), (3, 4)])
@pytest.mark.parametrize("bcs", ["symmetry", "dirichlet_neumann", "periodic"])
@pytest.mark.parametrize("dt", [1.0, 0.37])
def test_divergence_after_projection_is_the_residual(dim, depth, bcs, dt):
    o = OracleGrid(dim, depth)
    o.set_smoother("rbgs")
    rng = np.random.default_rng(3)
    kinds = {"symmetry": [cases.BC_SYMMETRY] * (2 * dim),
             "dirichlet_neumann": [cases.BC_DIRICHLET, cases.BC_NEUMANN] * dim,
             "periodic": [cases.BC_PERIODIC] * 2 + [cases.BC_SYMMETRY] * (2 * dim - 2)}[bcs]
    cases.random_problem(o, 5, bcs=kinds, alpha=False)
    # face coefficients; across a periodic direction the last face is the
    # first one
    al = [np.exp(0.5 * rng.standard_normal(o.face_shape(c))) for c in range(dim)]
    if bcs == "periodic":
        al[0][-1] = al[0][0]
    o.set_alpha(al)
    o.set_a(np.zeros(o.shape()))
    o.set_tolerance(1e-4)
    uf = _random_uf(o, rng, zero_walls=bcs != "dirichlet_neumann")
    st = o.project(uf, dt=dt)
    assert st["resa"] <= 1e-4 / dt ** 2
    res = o.get_res()
    d = _div(uf, o.delta()) / dt
    scale = np.abs(_div(_random_uf(o, np.random.default_rng(3), zero_walls=bcs != "dirichlet_neumann"),
                        o.delta())).max() / dt
    assert np.abs(d - res).max() <= 1e-12 * scale
    assert abs(np.abs(d).max() - st["resa"]) <= 1e-12 * scale


@pytest.mark.parametrize("dim,depth", [(2, 6), (3, 4)])
def test_gradient_field_is_removed(dim, depth):
    """uf = dt grad q on interior faces, 0 on the walls (Neumann p): the
    projection returns p = q + const and uf = 0 to the tolerance."""
    o = OracleGrid(dim, depth)
    o.set_smoother("rbgs")
    rng = np.random.default_rng(4)
    n, D, dt = o.n(), o.delta(), 0.5
    q = rng.standard_normal(o.shape())
    uf = []
    for c in range(dim):
        u = np.zeros(o.face_shape(c))
        inner = [slice(None)] * dim
        inner[c] = slice(1, n)
        u[tuple(inner)] = dt * np.diff(q, axis=c) / D
        uf.append(np.ascontiguousarray(u))
    o.set_a(np.zeros(o.shape()))
    o.set_tolerance(1e-9)
    st = o.project(uf, dt=dt)
    assert st["resa"] <= 1e-9 / dt ** 2
    p = o.get_a()
    assert np.abs((p - p.mean()) - (q - q.mean())).max() < 1e-8
    assert max(np.abs(u).max() for u in uf) < 1e-8


def test_defaults_dt_and_nrelax():
    """dt = 0 means 1 and nrelax = 0 means 4 (poisson.h:475-476)."""
    out = []
    for dt, nrelax in ((0.0, 0), (1.0, 4)):
        o = OracleGrid(2, 5)
        o.set_smoother("rbgs")
        o.set_a(np.zeros(o.shape()))
        uf = _random_uf(o, np.random.default_rng(6))
        st = o.project(uf, dt=dt, nrelax=nrelax)
        out.append((st, uf, o.get_a()))
    (s0, u0, a0), (s1, u1, a1) = out
    assert s0 == s1
    assert all(np.array_equal(x, y) for x, y in zip(u0, u1))
    assert np.array_equal(a0, a1)
