"""The CPU oracle against the reference's own known-answer files.

The `.ref` fixtures are the reference's test logs (tests/golden/README.md);
residuals there are printed with %.2g and errors with %g, so the oracle's
output is formatted the same way and compared as text.
"""
import numpy as np
import pytest

from oracle.oracle import OracleGrid
from tests import cases


def fmt_cycles(out):
    return ["%d %.2g" % (i, r) for i, r in out]


def test_poisson_cref_multigrid_2d():
    """test/poisson.cref: 2D 512^2 multigrid, GS, every line identical."""
    g = OracleGrid(2, 9, origin=(-0.5, -0.5, -0.5))
    out, err = cases.poisson_c(g)
    got = fmt_cycles(out) + ["# max error %g" % err]
    assert got == cases.read_ref("poisson.cref")


@pytest.mark.parametrize("dim,ref", [(2, "periodic.ref"), (3, "periodic.3D.ref")])
def test_periodic_ref(dim, ref):
    """test/periodic.c: poisson() on a periodic box, n = 8..64."""
    lines = []
    for n in (8, 16, 32, 64):
        g = OracleGrid(dim, int(np.log2(n)))
        e, st = cases.periodic_c(g)
        assert st["resa"] <= 1e-10
        lines.append("%d %g" % (n, e))
    assert lines == cases.read_ref(ref)


def test_poisson1d_ref():
    """test/poisson1D.c: poisson() in 1D, TOLERANCE 1e-4, depths 5..8."""
    lines = []
    for depth in range(5, 9):
        g = OracleGrid(1, depth, origin=(-0.5, -0.5, 0.0))
        e, st = cases.poisson1d_c(g)
        lines.append("%d %g" % (depth, e))
    assert lines == cases.read_ref("poisson1D.ref")


def test_variable_ref_final_error():
    """test/variable.c (the ref is a full quadtree, whose GS runs in Morton
    order): the converged error is identical at %g; the per-cycle residuals
    of the lexicographic multigrid GS stay within 10% of the quadtree's."""
    g = OracleGrid(2, 9, origin=(-0.5, -0.5, 0.0))
    out, err = cases.variable_c(g)
    ref = cases.read_ref("variable.ref")
    assert "# max error %g" % err == ref[-1]
    for (i, r), line in zip(out, ref[:-1]):
        assert abs(r / float(line.split()[1]) - 1) < 0.1


def test_poisson_3d_ref_final_error():
    """test/poisson.c in 3D (64^3; the ref is a full octree)."""
    g = OracleGrid(3, 6, origin=(-0.5, -0.5, -0.5))
    out, err = cases.poisson_c(g)
    ref = cases.read_ref("poisson.3D.ref")
    assert "# max error %g" % err == ref[-1]
    for (i, r), line in zip(out, ref[:-1]):
        assert abs(r / float(line.split()[1]) - 1) < 0.1


def _tree(dim, depth, origin):
    g = OracleGrid(dim, depth, origin=origin)
    g.set_tree()
    return g


@pytest.mark.parametrize("dim,depth,ref", [(2, 9, "poisson.ref"), (3, 6, "poisson.3D.ref")])
def test_poisson_ref_full_tree(dim, depth, ref):
    """test/poisson.c on a full quadtree / octree: Morton-order GS and the
    tree's corner ghosts reproduce every line of the reference log."""
    g = _tree(dim, depth, (-0.5, -0.5, -0.5))
    out, err = cases.poisson_c(g)
    assert fmt_cycles(out) + ["# max error %g" % err] == cases.read_ref(ref)


def test_variable_ref_full_tree():
    g = _tree(2, 9, (-0.5, -0.5, 0.0))
    out, err = cases.variable_c(g)
    assert fmt_cycles(out) + ["# max error %g" % err] == cases.read_ref("variable.ref")


def test_poisson_cadna_digits_full_tree():
    """test/poisson.CADNA.ref prints, per cycle, only the digits that CADNA's
    stochastic arithmetic certifies (9 to 12 significant digits for the
    first cycles).  The quadtree oracle reproduces all of them for cycles
    0-8; later cycles sit at the FP64 rounding floor, where CADNA itself
    keeps one or two digits."""
    g = _tree(2, 9, (-0.5, -0.5, -0.5))
    out, _ = cases.poisson_c(g)
    checked = 0
    for (i, r), line in zip(out, cases.read_ref("poisson.CADNA.ref")):
        if i > 8:
            break
        mant, exp = line.split()[1].split("E")
        digits = len(mant.replace("0.", "", 1))
        ref = float(line.split()[1])
        # within one unit of the last printed digit (CADNA truncates)
        assert abs(r - ref) < 10.0 ** (int(exp) - digits), (i, r, line)
        checked += digits
    assert checked == 70  # 12+10+9+8+7+7+6+7+4 certified digits
