"""The GPU adaptive-tree path against the tree oracle.

FP64 with the order-independent smoothers (weighted Jacobi and red-black GS)
is compared bit for bit: every residual norm, the residual field, the
solution and the solver statistics.  The red-black GPU solve also reproduces
the final-error lines of the reference's own logs (circle.ref,
circle-variable.ref, circle.3D.ref), which depend only on the discretisation
once the cycles have converged.
"""
import numpy as np
import pytest

import basilisk_amd as bmg
from oracle.tree import OracleTree
from tests import cases, tree_cases

pytestmark = pytest.mark.gpu

ORIGIN = (-0.5, -0.5, -0.5)


def _pair(dim, smoother, precision="fp64", origin=ORIGIN):
    g = bmg.Tree(dim, 1.0, origin, precision=precision, smoother=smoother)
    o = OracleTree(dim, 1.0, origin)
    o.set_smoother(smoother)
    return g, o


@pytest.mark.parametrize("smoother", ["jacobi", "rbgs"])
@pytest.mark.parametrize("dim,depth,variable", [(2, 7, False), (2, 7, True), (3, 5, False), (3, 5, True)])
def test_circle_cycles_bitexact(dim, depth, variable, smoother):
    print("circle", dim, depth, variable, smoother, flush=True)
    g, o = _pair(dim, smoother)
    tree_cases.circle_setup(g, depth, variable)
    tree_cases.circle_setup(o, depth, variable)
    assert g.residual() == o.residual()
    assert np.array_equal(g.get_res(), o.get_res())
    for _ in range(4):
        g.mg_cycle(4, 0, depth)
        o.mg_cycle(4, 0, depth)
        assert g.residual() == o.residual()
    assert np.array_equal(g.get_a(), o.get_a())
    assert np.array_equal(g.get_res(), o.get_res())


@pytest.mark.parametrize("dim,depth,variable,ref", [
    (2, 7, False, "circle.ref"), (2, 9, False, "circle.ref"), (2, 8, True, "circle-variable.ref"),
    (3, 5, False, "circle.3D.ref"), (3, 6, False, "circle.3D.ref")])
def test_circle_ref_final_error(dim, depth, variable, ref):
    """GPU red-black solve of test/circle.c: the converged error line equals
    the reference log's."""
    print("circle ref", dim, depth, variable, flush=True)
    g = bmg.Tree(dim, 1.0, ORIGIN, smoother="rbgs")
    lines = tree_cases.circle_lines(g, depth, variable)
    want = [l for l in cases.read_ref(ref) if l.startswith("max error %d " % depth)]
    assert lines[-1] == want[0]


def test_full_tree_bitexact():
    """A uniformly refined tree takes the reference's tree_is_full() paths."""
    print("full tree", flush=True)
    g, o = _pair(2, "rbgs")
    for t in (g, o):
        t.refine(lambda x, y, z, lev: lev < 6)
    assert g.nleaves() == 4 ** 6
    for t in (g, o):
        _, xyz = t.leaves()
        t.set_b(np.cos(3 * xyz[:, 0]) * np.sin(2 * xyz[:, 1]))
        t.set_a(np.zeros(len(xyz)))
        t.set_bc_function(bmg.LEFT, bmg.BC_DIRICHLET, lambda x, y, z: np.sin(4 * y))
    assert g.residual() == o.residual()
    for _ in range(3):
        g.mg_cycle(4, 0, 6)
        o.mg_cycle(4, 0, 6)
        assert g.residual() == o.residual()
    assert np.array_equal(g.get_a(), o.get_a())


@pytest.mark.parametrize("dim,depth", [(2, 8), (3, 6)])
@pytest.mark.parametrize("smoother", ["jacobi", "rbgs"])
def test_poisson_spheres_bitexact(dim, depth, smoother):
    """poisson() with alpha = 1/rho faces, a lambda field and mixed BCs on a
    tree refined around spheres: statistics and solution equal the oracle's."""
    print("spheres", dim, depth, smoother, flush=True)
    g, o = _pair(dim, smoother, origin=(0.0, 0.0, 0.0))
    tree_cases.spheres_problem(g, depth, 7)
    tree_cases.spheres_problem(o, depth, 7)
    for t in (g, o):
        t.set_niter(30)
    sg = g.poisson(tolerance=1e-6)
    so = o.poisson(tolerance=1e-6)
    for k in ("i", "resb", "resa", "nrelax", "minlevel"):
        assert sg[k] == so[k], k
    assert abs(sg["sum"] - so["sum"]) <= 1e-12 * max(1.0, abs(so["sum"])) * g.nleaves()
    assert np.array_equal(g.get_a(), o.get_a())


@pytest.mark.parametrize("dim,depth", [(2, 8), (3, 6)])
def test_mixed_precision_tree(dim, depth):
    """FP32 smoothing in FP64 defect correction reaches the same tolerance as
    FP64 and the same solution to 1e-9 (relative max norm)."""
    print("mixed", dim, depth, flush=True)
    ref = bmg.Tree(dim, 1.0, (0.0, 0.0, 0.0), precision="fp64")
    mix = bmg.Tree(dim, 1.0, (0.0, 0.0, 0.0), precision="mixed")
    for t in (ref, mix):
        tree_cases.spheres_problem(t, depth, 3)
        t.set_niter(60)
    tol = 1e-9
    s64 = ref.poisson(tolerance=tol)
    s32 = mix.poisson(tolerance=tol)
    assert s64["resa"] <= tol and s32["resa"] <= tol
    assert s32["i"] <= s64["i"] + 2
    a64, a32 = ref.get_a(), mix.get_a()
    assert np.abs(a32 - a64).max() <= 1e-9 * np.abs(a64).max()


def test_tree_errors():
    g = bmg.Tree(2)
    with pytest.raises(bmg.BMGError):
        g.set_smoother("gs")
    with pytest.raises(ValueError):
        g.set_a(np.zeros(5))
    g.refine(lambda x, y, z, lev: lev < 3)
    v = np.arange(g.nleaves(), dtype=float)
    g.set_b(v)
    assert np.array_equal(g.get_b(), v)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replays_match_eager(monkeypatch, precision):
    """The cycle and the residual run as captured HIP graphs by default;
    every replay (with eager host work in between) equals the eager launch
    sequence bit for bit, also after the graph key (nrelax, levels) changes."""
    print("graphs", precision, flush=True)
    trees = []
    for flag in ("1", "0"):
        monkeypatch.setenv("BMG_GRAPHS", flag)
        g = bmg.Tree(2, 1.0, ORIGIN, smoother="rbgs", precision=precision)
        tree_cases.circle_setup(g, 7, True)
        trees.append(g)
    ga, ge = trees
    assert ga.residual() == ge.residual()
    for i in range(5):
        ga.mg_cycle(4, 0, 7)
        ge.mg_cycle(4, 0, 7)
        assert ga.residual() == ge.residual()
        assert np.array_equal(ga.get_a(), ge.get_a())
    for nrelax, minlevel in ((3, 1), (4, 0), (2, 2)):
        ga.mg_cycle(nrelax, minlevel, 7)
        ge.mg_cycle(nrelax, minlevel, 7)
        assert ga.residual() == ge.residual()
    sa = ga.poisson(tolerance=1e-8)
    se = ge.poisson(tolerance=1e-8)
    assert sa["i"] == se["i"] and sa["resa"] == se["resa"]
    assert np.array_equal(ga.get_a(), ge.get_a())


TINY = {
    "root": lambda x, y, z, lev: np.zeros(len(lev), bool),
    "one_level": lambda x, y, z, lev: lev < 1,
    "corner": lambda x, y, z, lev: (lev < 1) | ((lev < 3) & (x < 0.3) & (y < 0.3)),
}


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("shape", list(TINY))
def test_tiny_trees_bitexact(dim, shape):
    """Degenerate trees: the root alone (depth 0), one full level, and one
    refined corner three levels deep next to unrefined cells of level 1."""
    print("tiny", dim, shape, flush=True)
    g, o = _pair(dim, "rbgs", origin=(0.0, 0.0, 0.0))
    stats = []
    for t in (g, o):
        t.refine(TINY[shape])
        _, xyz = t.leaves()
        t.set_b(np.sin(5 * xyz[:, 0]) + xyz[:, 1])
        t.set_a(np.zeros(len(xyz)))
        t.set_bc_function(bmg.LEFT, bmg.BC_DIRICHLET, lambda x, y, z: x + y)
        stats.append(t.poisson(tolerance=1e-10))
    for k in ("i", "resb", "resa", "nrelax"):
        assert stats[0][k] == stats[1][k], k
    assert np.array_equal(g.get_a(), o.get_a())


@pytest.mark.parametrize("dim,depth", [(2, 7), (3, 5)])
@pytest.mark.parametrize("dt", [1.0, 0.35])
def test_tree_project_bitexact(dim, depth, dt):
    """project() on an adaptive tree (bmg_tree_project): divergence after
    halo_face, poisson(), face update after the tree boundary of p -- the
    face velocities, the pressure and the statistics equal the tree oracle's
    bit for bit, from a host array and from a device tensor."""
    import torch
    from oracle.tree import OracleTree
    print("case tree project", dim, depth, dt, flush=True)
    g = bmg.Tree(dim, 1.0, (0.0, 0.0, 0.0), smoother="rbgs")
    o = OracleTree(dim, 1.0, (0.0, 0.0, 0.0))
    o.set_smoother("rbgs")
    for t in (g, o):
        tree_cases.spheres_problem(t, depth, 21, nspheres=4, minlevel=3, lam=False)
        t.set_tolerance(1e-7)
    nv = g.nfacevalues()
    assert nv == o.nfacevalues()
    u0 = np.random.default_rng(6).standard_normal(nv)
    ug, uo = u0.copy(), u0.copy()
    sg = g.project(ug, dt=dt)
    so = o.project(uo, dt=dt)
    for k in ("i", "resb", "resa", "nrelax", "minlevel"):
        assert sg[k] == so[k], (k, sg, so)
    assert np.array_equal(g.get_a(), o.get_a())
    assert np.array_equal(ug, uo)
    # the same from a device tensor, after a second projection of the result
    ud = torch.from_numpy(ug.copy()).cuda()
    sg2 = g.project(ud, dt=dt)
    so2 = o.project(uo, dt=dt)
    assert sg2["i"] == so2["i"] and sg2["resa"] == so2["resa"]
    assert np.array_equal(ud.cpu().numpy(), uo)
