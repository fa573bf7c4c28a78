"""The adjoint of poisson() on adaptive trees in the oracle
(OracleTree.residual_transpose, mg_cycle_transpose, adjoint), pinned by what
the mathematics fixes -- the reference has no adjoint:

- r^T is the transpose of the leaf operator r(x) = -lambda x - div(alpha
  grad x) with homogeneous boundary conditions: <r x, y> = <x, r^T y> to
  rounding for random x, y (the forward r is the oracle's residual, pinned
  by circle*.ref);
- the transposed V-cycle is the transpose of the forward V-cycle:
  <M r, y> = <r, M^T y>;
- the tree operator is not symmetric (<r x, y> != <r y, x>), which is why
  the transposes are needed;
- the gradients of J = <w, a> + 0.5 |a|^2 with respect to b, alpha (every
  face value) and lambda agree with central differences of the oracle's own
  solves along seeded directions."""
import numpy as np
import pytest

from oracle.tree import OracleTree, T_RES
from tests import tree_cases


def problem(dim, depth, seed=11, zero_bc=False):
    t = OracleTree(dim, 1.0, (0.0, 0.0, 0.0))
    t.set_smoother("rbgs")
    tree_cases.spheres_problem(t, depth, seed, nspheres=3, minlevel=2, lam=True)
    if zero_bc:
        t.set_bc(0, 2, None)
        t.set_bc(1, 1, None)
    t.restriction_alpha_lambda()
    return t


def r_of(t, x):
    n = t.nleaves()
    a0, b0 = t.get_a(), t.get_b()
    t.set_a(x)
    t.set_b(np.zeros(n))
    t.residual()
    r = t.get_res()
    t.set_a(a0)
    t.set_b(b0)
    return r


@pytest.mark.parametrize("dim,depth", [(2, 6), (3, 4)])
def test_residual_transpose_identity(dim, depth):
    t = problem(dim, depth, zero_bc=True)
    rng = np.random.default_rng(1)
    n = t.nleaves()
    for _ in range(3):
        x, y = rng.standard_normal(n), rng.standard_normal(n)
        lhs, rhs = r_of(t, x) @ y, x @ t.residual_transpose(y)
        assert abs(lhs - rhs) <= 1e-13 * np.abs(r_of(t, x)).sum() * np.abs(y).max()
    # and the operator is not symmetric
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    assert abs(r_of(t, x) @ y - r_of(t, y) @ x) > 1e-3 * abs(r_of(t, x) @ y)


@pytest.mark.parametrize("dim,depth", [(2, 6), (3, 4)])
def test_cycle_transpose_identity(dim, depth):
    t = problem(dim, depth)
    rng = np.random.default_rng(2)
    n = t.nleaves()

    def M(r):
        a0 = t.get_a()
        t.set_a(np.zeros(n))
        t.set_leaf(T_RES, r)
        t.mg_cycle(4, 1, t.depth())
        out = t.get_a()
        t.set_a(a0)
        return out
    for _ in range(3):
        r, y = rng.standard_normal(n), rng.standard_normal(n)
        lhs, rhs = M(r) @ y, r @ t.mg_cycle_transpose(y)
        assert abs(lhs - rhs) <= 1e-13 * np.abs(M(r)).sum() * np.abs(y).max()


@pytest.mark.parametrize("dim,depth", [(2, 6), (3, 4)])
def test_gradients_match_central_differences(dim, depth):
    t = problem(dim, depth)
    rng = np.random.default_rng(3)
    n, nv = t.nleaves(), t.nfacevalues()
    b0 = rng.standard_normal(n)
    alpha0 = np.exp(0.3 * rng.standard_normal(nv))
    lam0 = -rng.uniform(0.5, 2.0, n)
    w = rng.standard_normal(n)

    def solve(b, alpha, lam):
        t.set_b(b)
        t.set_alpha_faces(alpha)
        t.set_lambda(lam)
        t.set_a(np.zeros(n))
        st = t.poisson(tolerance=2e-12)
        assert st["resa"] <= 2e-12
        return t.get_a()

    def J(a):
        return w @ a + 0.5 * a @ a

    t.set_niter(300, 1)
    a = solve(b0, alpha0, lam0)
    psi, ga, gl, st = t.adjoint(w + a, tolerance=1e-13, nitermax=100, lam=True)
    assert st["resa"] <= 1e-13 and st["i"] < 60
    eps = 1e-6
    for k in range(4):
        db, dal, dl = rng.standard_normal(n), rng.standard_normal(nv), rng.standard_normal(n)
        want = psi @ db + ga @ dal + gl @ dl
        jp = J(solve(b0 + eps * db, alpha0 + eps * dal, lam0 + eps * dl))
        jm = J(solve(b0 - eps * db, alpha0 - eps * dal, lam0 - eps * dl))
        fd = (jp - jm) / (2 * eps)
        print("direction %d: fd %.12g adjoint %.12g rel %.2e" % (k, fd, want, abs(fd - want) / abs(want)))
        assert abs(fd - want) <= 1e-6 * abs(want), (k, fd, want)
