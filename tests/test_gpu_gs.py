"""The reference's default smoother on the GPU: in-place Gauss-Seidel in the
i-major order of foreach_level (poisson.h:296, 303-320;
grid/multigrid.h:191-215), run hyperplane by hyperplane (i+j+k = s).

It performs the serial loop's operations in the same order, so every
iterate, residual and mgstats field equals the oracle's serial GS bit for
bit, and test/poisson.cref is reproduced line for line on the GPU.
"""
import numpy as np
import pytest

import basilisk_amd as bmg
from oracle.oracle import OracleGrid
from tests import cases
from tests.test_gpu_parity import BC_MIXES, assert_stats_equal

pytestmark = pytest.mark.gpu


def pair(dim, depth, origin=(0.0, 0.0, 0.0), precision="fp64"):
    g = bmg.Grid(dim, depth, origin=origin, smoother="gs", precision=precision)
    o = OracleGrid(dim, depth, origin=origin)
    o.set_smoother("gs")
    return g, o


def test_poisson_cref_on_gpu():
    """test/poisson.cref: 2D 512^2 multigrid, 13 mg_cycles of the default
    relax with minlevel 0, Dirichlet -- every printed line, and every
    residual and value bit for bit with the serial oracle."""
    print("case poisson.cref gs", flush=True)
    g, o = pair(2, 9, origin=(-0.5, -0.5, -0.5))
    out_g, err_g = cases.poisson_c(g)
    got = ["%d %.2g" % (i, r) for i, r in out_g] + ["# max error %g" % err_g]
    assert got == cases.read_ref("poisson.cref")
    out_o, err_o = cases.poisson_c(o)
    assert [r for _, r in out_g] == [r for _, r in out_o]
    assert np.array_equal(g.get_a(), o.get_a())


@pytest.mark.parametrize("dim,depth", [(1, 9), (2, 7), (3, 5), (3, 7)])
@pytest.mark.parametrize("bcs", ["mixed", "periodic_x", "periodic_all"])
def test_gs_random_cycles_bitexact(dim, depth, bcs):
    """Variable alpha, a lambda field and every boundary kind: five cycles of
    the hyperplane GS equal the oracle's serial sweeps.  128^3 runs the
    one-launch-per-hyperplane kernel, the other levels the one-workgroup
    kernel."""
    print("case gs random", dim, depth, bcs, flush=True)
    g, o = pair(dim, depth)
    kinds = BC_MIXES[bcs][: 2 * dim]
    lam = np.random.default_rng(11).uniform(-2.0, -0.5, (1 << depth,) * dim)
    for grid in (g, o):
        cases.random_problem(grid, 4321 + dim, lam=lam, bcs=kinds)
        grid.restriction_alpha_lambda()
    assert g.residual() == o.residual()
    for it in range(3 if depth >= 7 and dim == 3 else 5):
        g.mg_cycle(3 - it % 2, 0)
        o.mg_cycle(3 - it % 2, 0)
        rg, ro = g.residual(), o.residual()
        assert rg == ro, (it, rg, ro)
    assert np.array_equal(g.get_a(), o.get_a())
    assert np.array_equal(g.get_res(), o.get_res())


@pytest.mark.parametrize("dim,depth", [(2, 8), (3, 5)])
def test_gs_one_launch_per_hyperplane(dim, depth, monkeypatch):
    """BMG_GS_BLOCK_MAX=1 sends every level above the coarse block to the
    one-launch-per-hyperplane kernel (in 2D it otherwise runs only from
    8192^2 up)."""
    monkeypatch.setenv("BMG_GS_BLOCK_MAX", "1")
    print("case gs plane kernel", dim, depth, flush=True)
    g, o = pair(dim, depth)
    lam = np.random.default_rng(5).uniform(-2.0, -0.5, (1 << depth,) * dim)
    for grid in (g, o):
        cases.random_problem(grid, 99, lam=lam, bcs=BC_MIXES["periodic_x"][: 2 * dim])
        grid.restriction_alpha_lambda()
    assert g.residual() == o.residual()
    for it in range(3):
        g.mg_cycle(4, 0)
        o.mg_cycle(4, 0)
        assert g.residual() == o.residual(), it
    assert np.array_equal(g.get_a(), o.get_a())


@pytest.mark.parametrize("dim", [2, 3])
def test_gs_poisson_stats_bitexact(dim):
    """poisson() with the default relax: iteration count, nrelax adaptation
    and every residual equal the oracle's."""
    depth = 6 if dim == 3 else 8
    print("case gs poisson", dim, flush=True)
    g, o = pair(dim, depth)
    lam = -np.random.default_rng(3).uniform(0.0, 5.0, (1 << depth,) * dim)
    for grid in (g, o):
        cases.random_problem(grid, 99, lam=lam, bcs=[cases.BC_DIRICHLET] * (2 * dim))
        grid.set_a(np.zeros(grid.shape()))
        grid.set_tolerance(1e-9)
    sg = g.poisson(nrelax=2)
    so = o.poisson(nrelax=2)
    assert_stats_equal(sg, so, g.get_b())
    assert np.array_equal(g.get_a(), o.get_a())


def test_gs_config2_bitexact():
    """The density-ratio-1000 sphere problem (64^3) through poisson() with
    the reference's default relax: the solve equals the oracle's serial GS
    bit for bit; the mixed-precision GS reaches the tolerance within one
    cycle of it and stays within 1e-9 of its solution."""
    from basilisk_amd import configs
    prob = configs.config2(depth=6, seed=1)
    print("case gs config2 64^3", flush=True)
    g, o = pair(3, 6)
    gm = bmg.Grid(3, 6, smoother="gs", precision="mixed")
    for grid in (g, o, gm):
        grid.set_b(prob["b"])
        grid.set_alpha(prob["alpha"])
        grid.set_a(np.zeros(prob["b"].shape))
    sg = g.poisson(tolerance=prob["tolerance"])
    so = o.poisson(tolerance=prob["tolerance"])
    sm = gm.poisson(tolerance=prob["tolerance"])
    assert_stats_equal(sg, so, prob["b"])
    assert np.array_equal(g.get_a(), o.get_a())
    assert sm["resa"] <= prob["tolerance"] and abs(sm["i"] - so["i"]) <= 1
    ref = o.get_a() - o.get_a().mean()
    am = gm.get_a() - gm.get_a().mean()
    assert np.abs(am - ref).max() / np.abs(ref).max() < 1e-9


def test_gs_refused_where_the_order_differs():
    """A full tree's GS runs in Morton order, not the multigrid grid's
    i-major order: refused, not silently replaced.  (viscosity() runs GS
    in the i-major order: tests/test_gpu_viscosity.py.)"""
    g = bmg.Grid(2, 4, smoother="gs")
    with pytest.raises(bmg.BMGError):
        g.set_kind("tree")
    g.set_smoother("rbgs")
    g.set_kind("tree")
    with pytest.raises(bmg.BMGError):
        g.set_smoother("gs")
