"""The adjoint of poisson() on adaptive trees on the GPU (bmg_tree_adjoint,
bmg_tree_residual_transpose, bmg_tree_cycle_transpose) against the tree
oracle, and the differentiable tree solve against central differences.

The transposed stages gather their terms in the forward order, so the
transposed operator, the transposed V-cycle, the adjoint state and the
gradients equal the oracle's bit for bit (FP64)."""
import numpy as np
import pytest
import torch

import basilisk_amd as bmg
from basilisk_amd import configs
from oracle.tree import OracleTree
from tests import tree_cases

pytestmark = pytest.mark.gpu


def pair(dim, depth, seed=11):
    g = bmg.Tree(dim, 1.0, (0.0, 0.0, 0.0), smoother="rbgs")
    o = OracleTree(dim, 1.0, (0.0, 0.0, 0.0))
    o.set_smoother("rbgs")
    for t in (g, o):
        tree_cases.spheres_problem(t, depth, seed, nspheres=3, minlevel=2, lam=True)
    return g, o


@pytest.mark.parametrize("dim,depth", [(2, 7), (3, 5)])
def test_transposed_operators_bitexact(dim, depth):
    print("case tree transposes", dim, depth, flush=True)
    g, o = pair(dim, depth)
    for t in (g, o):
        t.restriction_alpha_lambda()
    rng = np.random.default_rng(4)
    n = g.nleaves()
    for _ in range(2):
        y = rng.standard_normal(n)
        assert np.array_equal(g.residual_transpose(y), o.residual_transpose(y))
        assert np.array_equal(g.cycle_transpose(y, 4, 1), o.mg_cycle_transpose(y, 4, 1))


@pytest.mark.parametrize("dim,depth", [(2, 7), (3, 5)])
def test_adjoint_bitexact(dim, depth):
    """psi, dJ/dalpha and dJ/dlambda of the GPU equal the oracle's."""
    print("case tree adjoint", dim, depth, flush=True)
    g, o = pair(dim, depth)
    for t in (g, o):
        t.set_tolerance(1e-10)
        st = t.poisson(tolerance=1e-10)
    assert np.array_equal(g.get_a(), o.get_a())
    w = np.random.default_rng(5).standard_normal(g.nleaves())
    pg, ag, lg, sg = g.adjoint(w, tolerance=1e-10, nitermax=100, lam=True)
    po, ao, lo, so = o.adjoint(w, tolerance=1e-10, nitermax=100, lam=True)
    assert sg["i"] == so["i"] and sg["resa"] == so["resa"] and sg["resa"] <= 1e-10, (sg, so)
    assert np.array_equal(pg, po)
    assert np.array_equal(ag, ao)
    assert np.array_equal(lg, lo)
    # the same from device tensors
    pd, ad, ld, sd = g.adjoint(torch.from_numpy(w).cuda(), tolerance=1e-10, nitermax=100, lam=True)
    assert np.array_equal(pd.cpu().numpy(), po) and np.array_equal(ad.cpu().numpy(), ao)
    assert np.array_equal(ld.cpu().numpy(), lo)


def test_config5_octree_gradient_finite_differences():
    """The config-5 generator (octree around 200 spheres, density ratio
    1000, Helmholtz shift -1e-2/h_min^2) at depth 6: J = 0.5 |a - a_t|^2
    from the mixed-precision solver through TreePoissonSolver with forward
    recomputation; dJ/db and dJ/dalpha along 8 seeded directions against
    the central difference (J(x + eps d) - J(x - eps d)) / (2 eps),
    eps = 1e-6, of FP64 solves, each to 1e-6 of its own derivative.

    J is about 1e4 and its central difference about 1e-7, below J's last
    bit, so the difference is not formed by subtracting two values of J:
    with a_- the solution at x - eps d and delta = a_+ - a_-, which solves
    L(alpha_+) delta = 2 eps (db - div(dalpha grad a_-)) (the operator is
    linear in b and alpha; symmetry walls), J_+ - J_- equals
    delta . (a_- + delta / 2 - a_t) exactly.  Every step is a forward
    operation: the tree residual with alpha = dalpha, b = 0, lambda = 0
    gives -div(dalpha grad a_-), and delta comes from a forward solve to a
    tolerance relative to its right-hand side."""
    from basilisk_amd.autograd import TreePoissonSolver
    print("case config5 tree adjoint fd", flush=True)
    dev = torch.device("cuda:0")
    tm = bmg.Tree(3, 1.0, (0.0, 0.0, 0.0), precision="mixed", smoother="rbgs")
    configs.config5_tree(tm, depth=6)
    n, nv = tm.nleaves(), tm.nfacevalues()
    rng = np.random.default_rng(9)
    b0 = torch.from_numpy(rng.standard_normal(n)).to(dev)
    alpha0 = torch.from_numpy(np.exp(0.3 * rng.standard_normal(nv))).to(dev)
    lam = -1e-2 * (1 << 6) ** 2
    at = torch.from_numpy(rng.standard_normal(n)).to(dev)
    solver = TreePoissonSolver(tm, tolerance=1e-10, recompute=True)
    b = b0.clone().requires_grad_(True)
    alpha = alpha0.clone().requires_grad_(True)
    a = solver(b, alpha, lam)
    J = 0.5 * ((a - at) ** 2).sum()
    J.backward()
    assert solver.last_adjoint_stats["resa"] <= 1e-10
    gb, galpha = b.grad, alpha.grad
    tf = bmg.Tree(3, 1.0, (0.0, 0.0, 0.0), precision="fp64", smoother="rbgs")
    configs.config5_tree(tf, depth=6)
    tf.set_niter(200, 1)
    zero = torch.zeros(n, dtype=torch.float64, device=dev)

    def solve(rhs, aa, tol):
        tf.set_alpha_faces(aa.cpu().numpy())
        tf.set_lambda(lam)
        tf.set_b(rhs)
        tf.set_a(zero)
        st = tf.poisson(tolerance=tol)
        assert st["resa"] <= tol, st
        out = torch.empty(n, dtype=torch.float64, device=dev)
        tf.get_a(out)
        return out

    gen = torch.Generator(device="cpu").manual_seed(7)
    eps = 1e-6
    for k in range(8):
        db = torch.randn(n, generator=gen, dtype=torch.float64).to(dev)
        da = torch.randn(nv, generator=gen, dtype=torch.float64).to(dev)
        want = float((gb * db).sum()) + float((galpha * da).sum())
        am = solve(b0 - eps * db, alpha0 - eps * da, 1e-12)
        tf.set_alpha_faces(da.cpu().numpy())
        tf.set_lambda(0.0)
        tf.set_b(zero)
        tf.set_a(am)
        tf.residual()
        r = torch.from_numpy(tf.get_res()).to(dev)
        rhs = 2 * eps * (db + r)
        delta = solve(rhs, alpha0 + eps * da, 1e-12 * float(rhs.abs().max()))
        fd = float((delta * (am + 0.5 * delta - at)).sum()) / (2 * eps)
        rel = abs(fd - want) / abs(want)
        print("direction %d: fd %.12g adjoint %.12g rel %.2e" % (k, fd, want, rel), flush=True)
        assert rel <= 1e-6
