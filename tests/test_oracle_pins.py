"""Mathematical pins of the oracle's smoothers, ghost rules and nrelax rule.

No reference log covers the weighted-Jacobi and red-black sweeps that the GPU
parity tests run through, nor a Neumann condition with a non-zero value, so
these are fixed here by what the mathematics says about them:

* one weighted-Jacobi sweep (poisson.h:293-294, 325-328) multiplies a
  Fourier mode of the constant-coefficient operator by the textbook
  amplification 1 - w (1 - c) with w = 2/3, where c is the mode's Jacobi
  symbol; sine modes pin the homogeneous Dirichlet ghost, cosine modes the
  symmetry ghost, periodic modes the periodic copy;
* one red-black sweep on a periodic grid (the two colours alternate across
  the seam) leaves c v on the first colour and c^2 v on the second: the
  first colour reads only old values of the other, the second only new
  values of the first, through a ghost fill in between;
* every smoother keeps the discrete solution of the flux-form operator
  L u = div(alpha grad u) + lambda u fixed (poisson.h:303-320 is a
  splitting of L): built here independently with numpy, variable alpha,
  a lambda field and mixed boundary kinds;
* a manufactured solution with non-zero Neumann data (common.h:959, the
  outward derivative) and variable alpha converges at second order;
* the nrelax rule (poisson.h:198-204) on hand-worked cases.

Each check is also run on an oracle built with one plausible mistake
(mg_oracle.c's ORC_PLANT switch) and must fail there.
"""
import numpy as np
import pytest

from oracle.oracle import F_DA, F_RES, OracleGrid, nrelax_update
from tests import cases

SYM, DIR, NEU, PER = cases.BC_SYMMETRY, cases.BC_DIRICHLET, cases.BC_NEUMANN, cases.BC_PERIODIC


# ---------------------------------------------------------------- helpers

def _grid(dim, depth, kinds, smoother, plant=0, alpha=1.0, lam=0.0):
    """An oracle grid on [0,1]^dim whose dimension d has boundary kind
    kinds[d] on both sides (homogeneous)."""
    g = OracleGrid(dim, depth, plant=plant)
    for d in range(dim):
        if kinds[d] == PER:
            g.set_periodic(d)
        else:
            g.set_bc(2 * d, kinds[d], None)
            g.set_bc(2 * d + 1, kinds[d], None)
    g.set_smoother(smoother)
    g.lib.orc_set_alpha_const(g.g, float(alpha))
    g.set_lambda(float(lam))
    return g


def _mode(n, kind, k):
    """1D eigenvector of the 3-point second difference for the homogeneous
    ghost rule of `kind`, and its symbol cos(theta) (v[i+1] + v[i-1] =
    2 cos(theta) v[i] with the ghost values included)."""
    i = np.arange(n) + 0.5
    if kind == PER:
        th = 2.0 * np.pi * k / n
        return np.cos(th * (i - 0.5) + 0.3), np.cos(th)
    th = np.pi * k / n
    if kind == DIR:
        return np.sin(th * i), np.cos(th)
    return np.cos(th * i), np.cos(th)  # symmetry / homogeneous Neumann


def _fourier(dim, n, kinds, ks):
    v = np.ones((1,) * 0)
    mu = 0.0
    for d in range(dim):
        m, c = _mode(n, kinds[d], ks[d])
        v = np.multiply.outer(v, m) if d else m
        mu += c
    return v, mu / dim


def _jacobi_symbol(dim, n, mu, alpha, lam, rho):
    """Jacobi target of a mode v with res = rho v: (2 dim alpha mu v -
    Delta^2 rho v) / (2 dim alpha - lambda Delta^2)."""
    D2 = (1.0 / n) ** 2
    return (2 * dim * alpha * mu - D2 * rho) / (2 * dim * alpha - lam * D2)


def _sweep(g, v, res):
    lev = g.depth
    g.set_cell(F_DA, v, lev)
    g.set_cell(F_RES, res, lev)
    g.boundary(F_DA, lev, True)
    g.relax(lev)
    return g.get_cell(F_DA, lev)


JACOBI_CASES = [
    (2, 4, (PER, PER), (1, 2)),
    (2, 4, (DIR, SYM), (2, 7)),
    (3, 3, (PER, DIR, SYM), (1, 3, 5)),
    (3, 3, (SYM, PER, PER), (6, 2, 1)),
    (1, 5, (DIR,), (9,)),
]


def _jacobi_error(dim, depth, kinds, ks, plant=0):
    n = 1 << depth
    alpha, lam, rho = 1.7, -3.0, 0.9
    g = _grid(dim, depth, kinds, "jacobi", plant, alpha, lam)
    v, mu = _fourier(dim, n, kinds, ks)
    c = _jacobi_symbol(dim, n, mu, alpha, lam, rho)
    got = _sweep(g, v, rho * v)
    want = (1.0 - 2.0 / 3.0 * (1.0 - c)) * v
    return np.abs(got - want).max() / np.abs(want).max()


@pytest.mark.parametrize("dim,depth,kinds,ks", JACOBI_CASES)
def test_jacobi_sweep_fourier_amplification(dim, depth, kinds, ks):
    assert _jacobi_error(dim, depth, kinds, ks) < 1e-13


@pytest.mark.parametrize("plant", [1, 6])
def test_jacobi_pin_catches_planted_mistakes(plant):
    """(2a+c)/3 instead of (a+2c)/3, and alpha faces paired with the wrong
    neighbours (visible only when the two faces differ: a mode whose res
    term makes the symbol differ from 1 is enough for plant 1; plant 6 needs
    variable alpha and is caught by the fixed-point test below)."""
    if plant == 1:
        assert max(_jacobi_error(*c, plant=plant) for c in JACOBI_CASES) > 1e-3
    else:
        assert _fixed_point_error(2, 4, (DIR, SYM), "jacobi", plant=plant) > 1e-3


RB_CASES = [
    (2, 4, (PER, PER), (1, 2)),
    (2, 3, (PER, PER), (1, 0)),
    (3, 3, (PER, PER, PER), (1, 0, 2)),
    (1, 5, (PER,), (3,)),
]


def _rb_error(dim, depth, kinds, ks, plant=0):
    n = 1 << depth
    alpha, lam = 0.8, -5.0
    g = _grid(dim, depth, kinds, "rbgs", plant, alpha, lam)
    v, mu = _fourier(dim, n, kinds, ks)
    c = _jacobi_symbol(dim, n, mu, alpha, lam, 0.0)
    assert abs(c) > 0.2  # a mode the colours tell apart
    idx = np.indices(v.shape).sum(axis=0)
    want = np.where(idx % 2 == 0, c * v, c * c * v)
    got = _sweep(g, v, np.zeros_like(v))
    return np.abs(got - want).max() / np.abs(want).max()


@pytest.mark.parametrize("dim,depth,kinds,ks", RB_CASES)
def test_red_black_sweep_two_colour_symbol(dim, depth, kinds, ks):
    assert _rb_error(dim, depth, kinds, ks) < 1e-13


@pytest.mark.parametrize("plant", [2, 3])
def test_red_black_pin_catches_planted_mistakes(plant):
    """Colours swapped, and no ghost fill between the colours (the seam
    cells then read the first colour's old values)."""
    assert max(_rb_error(*c, plant=plant) for c in RB_CASES) > 1e-3


# ------------------------------------------- fixed point of the operator L

def _ghosted(u, kinds, dim):
    """u with one ghost layer per side from the homogeneous rules."""
    p = u
    for d in range(dim):
        lo = np.take(p, [0], axis=d)
        hi = np.take(p, [p.shape[d] - 1], axis=d)
        if kinds[d] == PER:
            lo, hi = np.take(p, [p.shape[d] - 1], axis=d), np.take(p, [0], axis=d)
        elif kinds[d] == DIR:
            lo, hi = -lo, -hi
        p = np.concatenate([lo, p, hi], axis=d)
    return p


def _apply_L(u, alpha, lam, kinds, delta):
    """Flux-form L u = sum_d (a_{+}(u_{+} - u) - a_{-}(u - u_{-}))/Delta^2
    + lambda u, alpha[d] on the n+1 faces along d."""
    dim = u.ndim
    p = _ghosted(u, kinds, dim)
    inner = tuple(slice(1, -1) for _ in range(dim))
    out = lam * u
    for d in range(dim):
        up = [slice(1, -1)] * dim
        um = [slice(1, -1)] * dim
        up[d], um[d] = slice(2, None), slice(0, -2)
        n = u.shape[d]
        ahi = np.take(alpha[d], range(1, n + 1), axis=d)
        alo = np.take(alpha[d], range(0, n), axis=d)
        out = out + (ahi * (p[tuple(up)] - p[inner]) - alo * (p[inner] - p[tuple(um)])) / delta ** 2
    return out


def _fixed_point_error(dim, depth, kinds, smoother, plant=0):
    n = 1 << depth
    rng = np.random.default_rng(7)
    g = _grid(dim, depth, kinds, smoother, plant)
    alpha = [np.exp(0.5 * rng.standard_normal(g.face_shape(c))) for c in range(dim)]
    for d in range(dim):
        if kinds[d] == PER:  # the seam face is one face seen from both sides
            sl = [slice(None)] * dim
            sl[d] = -1
            s0 = [slice(None)] * dim
            s0[d] = 0
            alpha[d][tuple(sl)] = alpha[d][tuple(s0)]
    lam = -np.exp(rng.standard_normal((n,) * dim))
    g.set_alpha(alpha)
    g.set_lambda(lam)
    u = rng.standard_normal((n,) * dim)
    res = _apply_L(u, alpha, lam, kinds, 1.0 / n)
    got = _sweep(g, u, res)
    return np.abs(got - u).max() / np.abs(u).max()


@pytest.mark.parametrize("smoother", ["gs", "jacobi", "rbgs"])
@pytest.mark.parametrize("dim,depth,kinds", [(2, 4, (DIR, SYM)), (3, 3, (SYM, PER, DIR)), (2, 3, (PER, PER))])
def test_relax_keeps_discrete_solution_fixed(smoother, dim, depth, kinds):
    assert _fixed_point_error(dim, depth, kinds, smoother) < 1e-12


# ------------------------------- manufactured solution, Neumann with values

def _manufactured(dim):
    p, q, r, t, u, s = 1.3, 0.4, 0.7, 2.1, 0.2, 0.5
    lam = -1.0

    def exact(x, y, z):
        v = np.cos(p * x + q) * np.exp(r * y)
        return v * np.cos(t * z + u) if dim == 3 else v

    def grad(x, y, z):
        cz = np.cos(t * z + u) if dim == 3 else 1.0
        gx = -p * np.sin(p * x + q) * np.exp(r * y) * cz
        gy = r * np.cos(p * x + q) * np.exp(r * y) * cz
        gz = -t * np.cos(p * x + q) * np.exp(r * y) * np.sin(t * z + u) if dim == 3 else 0.0 * x
        return gx, gy, gz

    def alpha(x, y, z):
        return 1.0 + s * x * y

    def rhs(x, y, z):
        lap = (-p * p + r * r - (t * t if dim == 3 else 0.0)) * exact(x, y, z)
        gx, gy, _ = grad(x, y, z)
        return alpha(x, y, z) * lap + s * y * gx + s * x * gy + lam * exact(x, y, z)
    return exact, grad, alpha, rhs, lam


def _neumann_error(dim, depth, plant=0):
    exact, grad, alpha, rhs, lam = _manufactured(dim)
    g = OracleGrid(dim, depth, plant=plant)
    for b in range(2 * dim):
        d, sign = b // 2, (1.0 if b % 2 == 0 else -1.0)
        g.set_bc_function(NEU, lambda x, y, z, d=d, sign=sign: sign * grad(x, y, z)[d], directions=[b])
    X, Y, Z = cases.mesh(g)
    shape = g.shape()
    g.set_b(rhs(X, Y, Z).reshape(shape))
    g.set_a(np.zeros(shape))
    n, D = g.n(), g.delta()
    comps = []
    for c in range(dim):
        axes = []
        for e in range(3):
            if e >= dim:
                axes.append(np.zeros(1))
            elif e == c:
                axes.append(np.arange(n + 1) * D)
            else:
                axes.append((np.arange(n) + 0.5) * D)
        F = np.meshgrid(*axes, indexing="ij")
        comps.append(alpha(*F).reshape(g.face_shape(c)))
    g.set_alpha(comps)
    g.set_lambda(lam)
    g.set_smoother("gs")
    g.set_niter(200, 1)
    st = g.poisson(tolerance=1e-9)
    assert st["resa"] <= 1e-9
    return np.abs(g.get_a() - exact(X, Y, Z).reshape(shape)).max()


@pytest.mark.parametrize("dim,depths", [(2, (4, 5, 6, 7)), (3, (3, 4, 5))])
def test_neumann_values_second_order(dim, depths):
    errs = [_neumann_error(dim, d) for d in depths]
    rates = [errs[i] / errs[i + 1] for i in range(len(errs) - 1)]
    assert all(3.3 < r < 4.7 for r in rates), (errs, rates)


def test_neumann_pin_catches_flipped_sign():
    errs = [_neumann_error(2, d, plant=4) for d in (4, 5, 6)]
    assert errs[-1] > 1e-2 and not (errs[0] / errs[1] > 3.3), errs


# ------------------------------------------------------------- residual

def _residual_error(dim, depth, plant=0):
    """residual() = b - L a with inhomogeneous Dirichlet values; compared
    with the numpy flux form with ghosts 2v - s."""
    n = 1 << depth
    rng = np.random.default_rng(3)
    g = OracleGrid(dim, depth, plant=plant)
    vals = []
    for b in range(2 * dim):
        v = rng.standard_normal((n + 4) ** (dim - 1))
        g.set_bc(b, DIR, v)
        vals.append(v.reshape((n + 4,) * (dim - 1))[(slice(2, -2),) * (dim - 1)] if dim > 1 else v)
    alpha = [np.exp(0.5 * rng.standard_normal(g.face_shape(c))) for c in range(dim)]
    lam = rng.standard_normal((n,) * dim)
    a = rng.standard_normal((n,) * dim)
    bb = rng.standard_normal((n,) * dim)
    g.set_alpha(alpha)
    g.set_lambda(lam)
    g.set_a(a)
    g.set_b(bb)
    maxres = g.residual()
    got = g.get_res()
    # the homogeneous L plus the inhomogeneous ghost contribution 2v on each wall
    La = _apply_L(a, alpha, lam, (DIR,) * dim, 1.0 / n)
    for b in range(2 * dim):
        d, hi = b // 2, b % 2 == 0
        face = n if hi else 0
        sl = [slice(None)] * dim
        sl[d] = [n - 1 if hi else 0]
        af = np.take(alpha[d], [face], axis=d)
        vb = vals[b] if dim == 1 else np.expand_dims(vals[b], d)
        La[tuple(sl)] += af * 2.0 * vb * n * n
    want = bb - La
    assert maxres == np.abs(got).max()
    return np.abs(got - want).max() / np.abs(want).max()


@pytest.mark.parametrize("dim,depth", [(1, 5), (2, 4), (3, 3)])
def test_residual_is_b_minus_flux_form(dim, depth):
    assert _residual_error(dim, depth) < 1e-12


def test_residual_pin_catches_lambda_sign():
    assert _residual_error(2, 4, plant=7) > 1e-6


# --------------------------------------------------------- nrelax rule

# (resb, resa, tolerance, nrelax) -> nrelax after the cycle, worked out from
# poisson.h:198-204: only while resa > tolerance; +1 below a reduction of
# 1.2 (up to 100), -1 above 10 (down to 2); both comparisons strict.
NRELAX_CASES = [
    ((1.0, 0.9, 0.0, 4), 5),
    ((1.0, 2.0, 0.0, 4), 5),      # growing residual
    ((1.2, 1.0, 0.0, 4), 4),      # ratio exactly 1.2
    ((5.0, 1.0, 0.0, 4), 4),
    ((10.0, 1.0, 0.0, 4), 4),     # ratio exactly 10
    ((11.0, 1.0, 0.0, 4), 3),
    ((11.0, 1.0, 0.0, 3), 2),
    ((11.0, 1.0, 0.0, 2), 2),     # floor
    ((1.0, 0.9, 0.0, 99), 100),
    ((1.0, 0.9, 0.0, 100), 100),  # cap
    ((1.0, 0.9, 0.95, 4), 4),     # converged: unchanged
    ((100.0, 1.0, 1.0, 4), 4),    # resa == tolerance counts as converged
    ((100.0, 1.0, 0.5, 4), 3),
]


@pytest.mark.parametrize("args,want", NRELAX_CASES)
def test_nrelax_rule(args, want):
    assert nrelax_update(*args) == want


def test_nrelax_scripted_sequence():
    """A scripted residual history, as mg_solve would see it."""
    seq = [1.0, 0.95, 0.9, 0.5, 0.04, 0.003, 0.0029, 1e-4]
    nr, got = 4, []
    for resb, resa in zip(seq, seq[1:]):
        nr = nrelax_update(resb, resa, 2e-4, nr)
        got.append(nr)
    assert got == [5, 6, 6, 5, 4, 5, 5]


def test_nrelax_pin_catches_swapped_thresholds():
    assert any(nrelax_update(*a, plant=5) != w for a, w in NRELAX_CASES)
