"""The adjoint gradient formula, pinned on the CPU by central finite
differences of J through the oracle's solver (no GPU).

The reference has no adjoint; what fixes the formula is the mathematics:
for J(phi(b, alpha, lambda)), the directional derivative along a random
perturbation must equal <grad, delta> to the finite-difference accuracy.
"""
import numpy as np
import pytest

from oracle.oracle import OracleGrid
from tests import adjoint_cases as ac


def _solve(prob, rhs=None, homogeneous=False, alpha=None, lam=None, b=None):
    p = dict(prob)
    if alpha is not None:
        p["alpha"] = alpha
    if lam is not None:
        p["lam"] = lam
    g = OracleGrid(p["dim"], p["depth"])
    ac.setup(g, p, p["b"] if rhs is None and b is None else (b if b is not None else rhs), homogeneous)
    g.set_niter(200, 1)
    st = g.poisson(tolerance=1e-12)
    assert st["resa"] <= 1e-12
    return g.get_a()


def _gradients(prob):
    phi = _solve(prob)
    psi = _solve(prob, rhs=phi - prob["phit"], homogeneous=True)
    ga, gl = ac.face_gradient(prob, phi, psi)
    return phi, psi, ga, gl


@pytest.mark.parametrize("dim,depth", [(2, 5), (3, 3)])
def test_adjoint_matches_finite_differences(dim, depth):
    prob = ac.problem(dim, depth, seed=11)
    phi, psi, ga, gl = _gradients(prob)
    rng = np.random.default_rng(5)
    eps = 1e-5
    for _ in range(4):
        da = [rng.standard_normal(a.shape) for a in prob["alpha"]]
        dl = rng.standard_normal(prob["lam"].shape)
        db = rng.standard_normal(prob["b"].shape)
        want = sum(float((g * d).sum()) for g, d in zip(ga, da)) + float((gl * dl).sum()) + float((psi * db).sum())
        jp = ac.objective(_solve(prob, alpha=[a + eps * d for a, d in zip(prob["alpha"], da)],
                                 lam=prob["lam"] + eps * dl, b=prob["b"] + eps * db), prob["phit"])
        jm = ac.objective(_solve(prob, alpha=[a - eps * d for a, d in zip(prob["alpha"], da)],
                                 lam=prob["lam"] - eps * dl, b=prob["b"] - eps * db), prob["phit"])
        fd = (jp - jm) / (2 * eps)
        assert abs(fd - want) <= 1e-6 * abs(want), (fd, want)


def test_adjoint_planted_sign_error_fails():
    """A plausible mistake (the wall term with the opposite sign) is caught
    by the same finite-difference check."""
    prob = ac.problem(2, 4, seed=3)
    phi, psi, ga, gl = _gradients(prob)
    bad = [g.copy() for g in ga]
    bad[0][0] = -bad[0][0]
    rng = np.random.default_rng(1)
    da = [rng.standard_normal(a.shape) for a in prob["alpha"]]
    eps = 1e-5
    fd = (ac.objective(_solve(prob, alpha=[a + eps * d for a, d in zip(prob["alpha"], da)]), prob["phit"])
          - ac.objective(_solve(prob, alpha=[a - eps * d for a, d in zip(prob["alpha"], da)]), prob["phit"])) / (2 * eps)
    good = sum(float((g * d).sum()) for g, d in zip(ga, da))
    wrong = sum(float((g * d).sum()) for g, d in zip(bad, da))
    assert abs(fd - good) <= 1e-6 * abs(good)
    assert abs(fd - wrong) > 1e-4 * abs(good)
