"""GPU parity: libbmg (HIP, gfx950) against the CPU oracle.

The FP64 path with the Jacobi and red-black smoothers performs the same
operations in the same order as the oracle, so iterates, residuals and
mgstats are compared bit for bit.  The mixed-precision path is compared
against the FP64 oracle with the tolerances stated in each test.
"""
import numpy as np
import pytest

import basilisk_amd as bmg
from oracle.oracle import OracleGrid
from tests import cases

pytestmark = pytest.mark.gpu


def assert_stats_equal(sg, so, b):
    """mgstats equal except `sum`, which the GPU adds in a different
    (deterministic, tree) order than the oracle's serial loop: the two sums
    agree within the rounding bound 1e-14 * sum(|b|)."""
    for k in ("i", "resb", "resa", "nrelax", "minlevel"):
        assert sg[k] == so[k], (k, sg, so)
    assert abs(sg["sum"] - so["sum"]) <= 1e-14 * np.abs(b).sum()


def pair(dim, depth, smoother, origin=(0.0, 0.0, 0.0), precision="fp64"):
    g = bmg.Grid(dim, depth, origin=origin, smoother=smoother, precision=precision)
    o = OracleGrid(dim, depth, origin=origin)
    o.set_smoother(smoother)
    return g, o


@pytest.mark.parametrize("smoother", ["jacobi", "rbgs"])
@pytest.mark.parametrize("dim,depth", [(2, 9), (3, 6), (1, 10)])
def test_poisson_c_bitexact(smoother, dim, depth):
    """test/poisson.c set-up (Dirichlet, 13 mg_cycles, minlevel 0)."""
    print("case poisson_c", smoother, dim, depth, flush=True)
    g, o = pair(dim, depth, smoother, origin=(-0.5, -0.5, -0.5))
    out_g, err_g = cases.poisson_c(g)
    out_o, err_o = cases.poisson_c(o)
    assert [r for _, r in out_g] == [r for _, r in out_o]
    assert err_g == err_o
    assert np.array_equal(g.get_a(), o.get_a())
    assert np.array_equal(g.get_res(), o.get_res())


@pytest.mark.parametrize("smoother", ["jacobi", "rbgs"])
def test_variable_c_bitexact(smoother):
    """test/variable.c set-up: variable face coefficients restricted on the GPU."""
    print("case variable_c", smoother, flush=True)
    g, o = pair(2, 9, smoother, origin=(-0.5, -0.5, 0.0))
    out_g, err_g = cases.variable_c(g)
    out_o, err_o = cases.variable_c(o)
    assert [r for _, r in out_g] == [r for _, r in out_o]
    for lev in range(9):
        for c in range(2):
            assert np.array_equal(g.get_alpha(c, lev), o.get_face(c, lev)), (lev, c)
    assert np.array_equal(g.get_a(), o.get_a())


@pytest.mark.parametrize("dim", [2, 3])
def test_periodic_ref_on_gpu(dim):
    """test/periodic.c through poisson() on the GPU: bit-exact with the
    oracle's red-black solve, and the printed errors equal the reference
    log (the converged solution does not depend on the smoother)."""
    ref = cases.read_ref("periodic.ref" if dim == 2 else "periodic.3D.ref")
    lines = []
    for n in (8, 16, 32, 64):
        print("case periodic", dim, n, flush=True)
        depth = int(np.log2(n))
        g, o = pair(dim, depth, "rbgs")
        e_g, st_g = cases.periodic_c(g)
        e_o, st_o = cases.periodic_c(o)
        assert_stats_equal(st_g, st_o, g.get_b())
        assert e_g == e_o
        lines.append("%d %g" % (n, e_g))
    assert lines == ref


def test_poisson1d_ref_on_gpu():
    ref = cases.read_ref("poisson1D.ref")
    lines = []
    for depth in range(5, 9):
        g, o = pair(1, depth, "jacobi", origin=(-0.5, 0.0, 0.0))
        e_g, st_g = cases.poisson1d_c(g)
        e_o, st_o = cases.poisson1d_c(o)
        assert_stats_equal(st_g, st_o, g.get_b())
        assert e_g == e_o
        lines.append("%d %g" % (depth, e_g))
    # Jacobi stops at a different iterate than the reference's GS at
    # TOLERANCE 1e-4; the discretisation error agrees to 4 digits
    for a, b in zip(lines, ref):
        assert abs(float(a.split()[1]) / float(b.split()[1]) - 1) < 1e-4


BC_MIXES = {
    "dirichlet": [cases.BC_DIRICHLET] * 6,
    "neumann": [cases.BC_NEUMANN] * 6,
    "mixed": [cases.BC_DIRICHLET, cases.BC_NEUMANN, cases.BC_SYMMETRY, cases.BC_DIRICHLET,
              cases.BC_NEUMANN, cases.BC_SYMMETRY],
    "periodic_x": [cases.BC_PERIODIC, cases.BC_PERIODIC, cases.BC_DIRICHLET, cases.BC_SYMMETRY,
                   cases.BC_NEUMANN, cases.BC_DIRICHLET],
    "periodic_all": [cases.BC_PERIODIC] * 6,
}


@pytest.mark.parametrize("bcs", sorted(BC_MIXES))
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("smoother", ["jacobi", "rbgs"])
def test_random_problem_cycles_bitexact(bcs, dim, smoother):
    """Random variable-coefficient Helmholtz problems with every boundary
    kind: relax, prolongation (edge and corner ghosts), residual and five
    mg_cycles with minlevel 0 equal the oracle bit for bit."""
    depth = 5 if dim == 3 else 7
    print("case random", bcs, dim, smoother, flush=True)
    g, o = pair(dim, depth, smoother)
    kinds = BC_MIXES[bcs][: 2 * dim]
    rng_seed = 1234 + dim
    lam = np.random.default_rng(7).uniform(-2.0, -0.5, (1 << depth,) * dim)
    for grid in (g, o):
        cases.random_problem(grid, rng_seed, lam=lam, bcs=kinds)
        grid.restriction_alpha_lambda()
    for lev in range(depth + 1):
        for c in range(dim):
            assert np.array_equal(g.get_alpha(c, lev), o.get_face(c, lev)), (lev, c)
    assert g.residual() == o.residual()
    assert np.array_equal(g.get_res(), o.get_res())
    for it in range(5):
        g.mg_cycle(3, 0)
        o.mg_cycle(3, 0)
        rg, ro = g.residual(), o.residual()
        assert rg == ro, (it, rg, ro)
    assert np.array_equal(g.get_a(), o.get_a())


@pytest.mark.parametrize("kernel", ["default", "inplace", "stream", "stream2", "stream3", "pair", "fused"])
@pytest.mark.parametrize("bcs", ["mixed", "periodic_all", "periodic_x"])
@pytest.mark.parametrize("depth", [7, 8])
def test_large_levels_bitexact(kernel, bcs, depth, monkeypatch):
    """Every red-black kernel of the large levels -- the two-launch in-place
    one ("inplace", the default below 256^3 and with periodic directions),
    the single-pass streaming ones (plane ring in LDS, halo
    recomputation; "stream2" with colour-mapped threads and prefetched
    coefficients; "stream3" with per-column coefficient vectors, walls
    only, so the periodic cases run the default kernel), the in-place one
    over vector pairs ("pair") and the
    one-launch fused one (red neighbours recomputed on the fly) -- and the
    fused residual/restriction kernel give the
    oracle's two-colour iterates bit for bit, periodic wrap included."""
    if kernel != "default":
        monkeypatch.setenv("BMG_RB_KERNEL", kernel)
    print("case large levels", kernel, bcs, depth, flush=True)
    g, o = pair(3, depth, "rbgs")
    kinds = BC_MIXES[bcs]
    lam = np.random.default_rng(8).uniform(-2.0, -0.5, (1 << depth,) * 3)
    for grid in (g, o):
        cases.random_problem(grid, 77, lam=lam, bcs=kinds)
        grid.restriction_alpha_lambda()
    assert g.residual() == o.residual()
    for it in range(2 if depth == 8 else 3):
        g.mg_cycle(3 - it % 2, 0)
        o.mg_cycle(3 - it % 2, 0)
        assert g.residual() == o.residual(), it
    assert np.array_equal(g.get_a(), o.get_a())
    assert np.array_equal(g.get_res(), o.get_res())


@pytest.mark.parametrize("kernel", ["default", "pair"])
@pytest.mark.parametrize("bcs", ["dirichlet", "periodic"])
def test_large_levels_2d_bitexact(kernel, bcs, monkeypatch):
    """2D levels above 4096 cells with the in-place kernels (one cell per
    thread, or pairs of cells with vector loads) against the oracle."""
    if kernel != "default":
        monkeypatch.setenv("BMG_RB_KERNEL", kernel)
    print("case large levels 2d", kernel, bcs, flush=True)
    g, o = pair(2, 10, "rbgs")
    lam = np.random.default_rng(4).uniform(-2.0, -0.5, (1 << 10,) * 2)
    for grid in (g, o):
        cases.random_problem(grid, 41, lam=lam, bcs=[cases.BC_DIRICHLET, cases.BC_NEUMANN] * 2)
        if bcs == "periodic":
            grid.set_periodic(1)
        grid.restriction_alpha_lambda()
    assert g.residual() == o.residual()
    for it in range(2):
        g.mg_cycle(2, 0)
        o.mg_cycle(2, 0)
        assert g.residual() == o.residual(), it
    assert np.array_equal(g.get_a(), o.get_a())


@pytest.mark.parametrize("dim", [2, 3])
def test_mg_solve_stats_bitexact(dim):
    """poisson() with a variable alpha and a lambda field: the iteration
    count, the nrelax adaptation and every residual equal the oracle's."""
    depth = 6 if dim == 3 else 8
    g, o = pair(dim, depth, "rbgs")
    lam = -np.random.default_rng(3).uniform(0.0, 5.0, (1 << depth,) * dim)
    for grid in (g, o):
        cases.random_problem(grid, 99, lam=lam, bcs=[cases.BC_DIRICHLET] * (2 * dim))
        grid.set_a(np.zeros(grid.shape()))
        grid.set_tolerance(1e-9)
    sg = g.poisson(nrelax=2)
    so = o.poisson(nrelax=2)
    assert sg["i"] == so["i"] and sg["nrelax"] == so["nrelax"] and sg["minlevel"] == so["minlevel"]
    assert sg["resb"] == so["resb"] and sg["resa"] == so["resa"]
    assert abs(sg["sum"] - so["sum"]) <= 1e-12 * max(1.0, abs(so["sum"]) + np.abs(g.get_b()).sum())
    assert np.array_equal(g.get_a(), o.get_a())


def test_level_zero_and_tiny_grids():
    """depth 0 and 1, periodic level-0 ghost rule (grid/multigrid.h:655-664)."""
    for depth in (0, 1, 2):
        for periodic in (False, True):
            g, o = pair(2, depth, "jacobi")
            for grid in (g, o):
                cases.random_problem(grid, 5, bcs=[cases.BC_DIRICHLET] * 4)
                if periodic:
                    grid.set_periodic(0)
                grid.restriction_alpha_lambda()
                grid.residual()
            for it in range(3):
                g.mg_cycle(2, 0)
                o.mg_cycle(2, 0)
                assert g.residual() == o.residual()
            assert np.array_equal(g.get_a(), o.get_a())


@pytest.mark.parametrize("dim", [2, 3])
def test_mixed_precision_converges_to_fp64_solution(dim):
    """FP32 smoothing inside FP64 defect correction reaches the same
    tolerance as FP64 in at most one more cycle, and its solution is within
    1e-9 relative max-norm of the FP64 oracle's.  The tolerance sits just
    above the FP64 rounding floor of each problem (about 5e-10 for 512^2 with
    O(1) data, 1e-12 for 64^3)."""
    depth = 6 if dim == 3 else 9
    tol = 1e-10 if dim == 3 else 2e-9
    g64 = bmg.Grid(dim, depth, smoother="rbgs")
    gmx = bmg.Grid(dim, depth, smoother="rbgs", precision="mixed")
    o = OracleGrid(dim, depth)
    o.set_smoother("rbgs")
    stats = []
    for grid in (g64, gmx, o):
        cases.random_problem(grid, 11, bcs=[cases.BC_DIRICHLET] * (2 * dim))
        grid.set_a(np.zeros(grid.shape()))
        stats.append(grid.poisson(tolerance=tol))
    s64, smx, so = stats
    assert s64["i"] == so["i"]
    assert smx["resa"] <= tol
    assert abs(smx["i"] - s64["i"]) <= 1
    ref = o.get_a()
    scale = np.abs(ref).max()
    assert np.abs(g64.get_a() - ref).max() == 0.0
    assert np.abs(gmx.get_a() - ref).max() <= 1e-9 * scale


@pytest.mark.parametrize("kernel", ["inplace", "pair"])
def test_mixed_precision_kernels_agree(kernel, monkeypatch):
    """The FP32 smoother of the mixed path at 256^3: the default kernel (the
    single-pass column-streamed sweep) and the two-colour in-place kernels
    perform the same FP32 operations, so whole mixed-precision solves agree
    bit for bit, statistics included."""
    print("case mixed kernels", kernel, flush=True)
    lam = np.random.default_rng(12).uniform(-2.0, -0.5, (256,) * 3)
    out = []
    for k in ("default", kernel):
        if k != "default":
            monkeypatch.setenv("BMG_RB_KERNEL", k)
        g = bmg.Grid(3, 8, smoother="rbgs", precision="mixed")
        cases.random_problem(g, 21, lam=lam, bcs=BC_MIXES["mixed"])
        g.set_a(np.zeros(g.shape()))
        st = g.poisson(tolerance=1e-9)
        out.append((st, g.get_a()))
    (s0, a0), (s1, a1) = out
    for key in ("i", "resb", "resa", "nrelax", "minlevel"):
        assert s0[key] == s1[key], key
    assert np.array_equal(a0, a1)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fused_correction_residual_agrees(precision, monkeypatch):
    """At 256^3 mg_solve leaves each cycle's correction a += da to the
    streaming residual kernel, which also restricts the residual: the solve
    equals, bit for bit, the one with the separate correction, residual and
    restriction kernels (BMG_FUSED_RESIDUAL=0)."""
    print("case fused residual", precision, flush=True)
    lam = np.random.default_rng(13).uniform(-2.0, -0.5, (256,) * 3)
    out = []
    for fused in ("1", "0"):
        monkeypatch.setenv("BMG_FUSED_RESIDUAL", fused)
        g = bmg.Grid(3, 8, smoother="rbgs", precision=precision)
        cases.random_problem(g, 23, lam=lam, bcs=BC_MIXES["mixed"])
        g.set_a(np.zeros(g.shape()))
        st = g.poisson(tolerance=1e-9)
        out.append((st, g.get_a(), g.get_res()))
    (s0, a0, r0), (s1, a1, r1) = out
    for key in ("i", "resb", "resa", "nrelax", "minlevel"):
        assert s0[key] == s1[key], key
    assert np.array_equal(a0, a1)
    assert np.array_equal(r0, r1)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("depth", [7, 8])
def test_cycle_graph_replays_equal_plain_launches(precision, depth, monkeypatch):
    """The cycle runs as a captured HIP graph, replayed for every cycle with
    the same (nrelax, minlevel, maxlevel): each replay, with eager residual
    kernels and synchronizations in between, equals the plain launches bit
    for bit, and so does a whole solve."""
    print("case graphs", precision, depth, flush=True)
    lam = np.random.default_rng(14).uniform(-2.0, -0.5, (1 << depth,) * 3)
    runs = []
    for graphs in ("1", "0"):
        monkeypatch.setenv("BMG_GRAPHS", graphs)
        g = bmg.Grid(3, depth, smoother="rbgs", precision=precision)
        cases.random_problem(g, 31, lam=lam, bcs=BC_MIXES["mixed"])
        g.restriction_alpha_lambda()
        res = [g.residual()]
        for it in range(4):
            g.mg_cycle(3, 1)
            res.append(g.residual())
        a = g.get_a()
        g.set_a(np.zeros(g.shape()))
        st = g.poisson(tolerance=1e-9)
        runs.append((res, a, st, g.get_a()))
    (r0, a0, s0, f0), (r1, a1, s1, f1) = runs
    assert r0 == r1
    assert np.array_equal(a0, a1)
    for key in ("i", "resb", "resa", "nrelax", "minlevel"):
        assert s0[key] == s1[key], key
    assert np.array_equal(f0, f1)


def test_nan_is_reported():
    g = bmg.Grid(2, 5)
    b = np.zeros(g.shape())
    b[3, 4] = np.nan
    g.set_b(b)
    with pytest.raises(bmg.BMGError) as e:
        g.residual()
    assert e.value.code == 5


def test_unknown_smoother_refused():
    g = bmg.Grid(2, 3)
    with pytest.raises(bmg.BMGError) as e:
        bmg._check(g.lib.bmg_set_smoother(g._g, 7))
    assert e.value.code == 1


@pytest.mark.parametrize("smoother", ["jacobi", "rbgs"])
@pytest.mark.parametrize("dim,depth", [(2, 9), (3, 6)])
def test_full_tree_bitexact(smoother, dim, depth):
    """test/poisson.c on a full quadtree / octree (Dirichlet, so the tree's
    corner ghost rule enters the prolongation): GPU equals the tree-mode
    oracle bit for bit."""
    print("case full tree", smoother, dim, depth, flush=True)
    g = bmg.Grid(dim, depth, origin=(-0.5, -0.5, -0.5), smoother=smoother, kind="tree")
    o = OracleGrid(dim, depth, origin=(-0.5, -0.5, -0.5))
    o.set_smoother(smoother)
    o.set_tree()
    out_g, err_g = cases.poisson_c(g)
    out_o, err_o = cases.poisson_c(o)
    assert [r for _, r in out_g] == [r for _, r in out_o]
    assert np.array_equal(g.get_a(), o.get_a())
