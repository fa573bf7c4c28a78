"""The oracle's embedded-boundary branch (oracle/embed_oracle.h and the
EMBED paths of mg_oracle.c: fractions, fragment geometry, embed_flux,
dirichlet_gradient, embedded face gradients, bilinear_embed) against the
reference's own logs of test/neumann.c and test/dirichlet.c
(tests/golden/ref/neumann.ref, dirichlet.ref): for N = 32 .. 512 the
residual of the exact solution (full cells, all cells) and, after
poisson() to 1e-6 with the reference's GS from minlevel 4, the six error
norms print identically.  The Neumann log's cycle counts and final nrelax
are reproduced too; the Dirichlet log's cycle counts at N = 32, 64, 128
come from the reference's default quadtree (whose prolongation is the plain
bilinear), while this is the multigrid grid, and differ.  Planted mistakes
in the embedded flux, the face gradient and the interpolation fail."""
import numpy as np
import pytest

from oracle.oracle import OracleGrid
from tests import cases
from tests import embed_cases as ec


def _run(N, dirichlet, smoother="gs", plant=0, tol=1e-6):
    p = ec.problem(N, dirichlet)
    o = OracleGrid(2, p["depth"], L0=1.0, origin=(-0.5, -0.5, 0.0), plant=plant)
    o.set_smoother(smoother)
    ec.setup_oracle(o, p)
    maxp = o.residual()
    res = o.get_res()
    maxf = np.abs(res[p["cs"] == 1.0]).max()
    lines = ["maxres %d %.3g %.3g" % (N, maxf, maxp)]
    st = o.poisson(tolerance=tol, minlevel=4)
    assert not o.embed_bad()
    nn = ec.norms(p, o.get_a())
    lines.append("%d %.3g %.3g %.3g %.3g %.3g %.3g %d %d" % ((N,) + tuple(nn) + (st["i"], st["nrelax"])))
    return lines, o.get_a(), st


def test_neumann_ref():
    got = []
    for N in (32, 64, 128, 256, 512):
        got += _run(N, False)[0]
    assert got == cases.read_ref("neumann.ref")


def test_dirichlet_ref():
    ref = cases.read_ref("dirichlet.ref")
    got = []
    for N in (32, 64, 128, 256, 512):
        got += _run(N, True)[0]
    for g, r in zip(got, ref):
        if g.startswith("maxres"):
            assert g == r
        else:
            # N and the six norms; the cycle counts are the quadtree's
            assert g.split()[:7] == r.split()[:7], (g, r)


@pytest.mark.parametrize("dirichlet", [False, True])
@pytest.mark.parametrize("smoother", ["rbgs", "jacobi"])
def test_red_black_and_jacobi_reach_the_same_solution(dirichlet, smoother):
    """The order-independent smoothers (red-black: each pass from the state
    before it; Jacobi) solve the same discrete problem: at 1e-10 the
    solution is the GS one to 1e-9 and the error norms print as the log's."""
    ref = cases.read_ref("dirichlet.ref" if dirichlet else "neumann.ref")
    _, a_gs, _ = _run(64, dirichlet, "gs", tol=1e-10)
    lines, a, st = _run(64, dirichlet, smoother, tol=1e-10)
    assert st["resa"] <= 1e-10
    m = np.isfinite(a) & (np.abs(a_gs) < 1e29)
    assert np.abs(a[m] - a_gs[m]).max() < 1e-9
    assert lines[1].split()[1:7] == ref[3].split()[1:7]


@pytest.mark.parametrize("plant,dirichlet", [(8, False), (8, True), (9, False), (10, True)])
def test_planted_embed_mistakes_fail(plant, dirichlet):
    ref = cases.read_ref("dirichlet.ref" if dirichlet else "neumann.ref")
    lines, _, _ = _run(64, dirichlet, plant=plant)
    assert lines[0] != ref[2] or lines[1].split()[:7] != ref[3].split()[:7]


def test_library_geometry_equals_the_oracle():
    """bmg_embed_geometry (the library's host code, which also builds the
    cut-cell records) gives the oracle's centroids, lengths and normals."""
    import basilisk_amd as bmg
    from oracle.oracle import embed_geometry
    for dirichlet in (True, False):
        p = ec.problem(128, dirichlet)
        a = bmg.embed_geometry(p["cs"], p["fx"], p["fy"])
        b = embed_geometry(p["cs"], p["fx"], p["fy"])
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
