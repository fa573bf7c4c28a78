"""The two-sweep kernel of the large FP32 levels (k_rb_stream3t: two
red-black sweeps in one pass over the coefficients) against two launches of
the one-sweep kernel, bit for bit: two mixed-precision grids whose solves
differ only in that kernel (BMG_SWEEP2_MIN) give the same statistics and
the same solution.  The one-sweep kernel equals the oracle's red-black
sweep (test_gpu_parity.py), so this carries the parity over.  The two-sweep
kernel is the default on FP32 levels of 256^3 cells and more
(BMG_SWEEP2_MIN=0 turns it off; 128 extends it to 128^3 levels)."""
import numpy as np
import pytest

import basilisk_amd as bmg
from tests import cases

pytestmark = pytest.mark.gpu

BCS = {
    "mixed": [cases.BC_DIRICHLET, cases.BC_NEUMANN, cases.BC_SYMMETRY, cases.BC_DIRICHLET,
              cases.BC_NEUMANN, cases.BC_SYMMETRY],
    "dirichlet": [cases.BC_DIRICHLET] * 6,
    "symmetry": [cases.BC_SYMMETRY] * 6,
}


def _grids(depth, smin, bcs, coef, monkeypatch):
    grids = []
    for knob in ("0", str(smin)):
        monkeypatch.setenv("BMG_SWEEP2_MIN", knob)
        g = bmg.Grid(3, depth, precision="mixed", smoother="rbgs")
        n = 1 << depth
        lam = np.random.default_rng(8).uniform(-2.0, -0.5, (n,) * 3) if coef == "field" else (
            -0.7 if coef == "lam_const" else 0.0)
        cases.random_problem(g, 77, lam=lam, alpha=coef != "alpha_const", bcs=BCS[bcs])
        g.restriction_alpha_lambda()
        grids.append(g)
    monkeypatch.delenv("BMG_SWEEP2_MIN", raising=False)
    return grids


@pytest.mark.parametrize("depth,smin", [(7, 128), (8, 256)])
@pytest.mark.parametrize("bcs", sorted(BCS))
@pytest.mark.parametrize("coef", ["field", "lam_const", "alpha_const"])
def test_two_sweeps_bitexact(depth, smin, bcs, coef, monkeypatch):
    print("case sweep2", depth, bcs, coef, flush=True)
    one, two = _grids(depth, smin, bcs, coef, monkeypatch)
    s1 = one.poisson(tolerance=1e-9)
    s2 = two.poisson(tolerance=1e-9)
    for key in ("i", "resb", "resa", "nrelax"):
        assert s1[key] == s2[key], (key, s1, s2)
    assert np.array_equal(one.get_a(), two.get_a())


@pytest.mark.parametrize("nrelax", [2, 3, 4, 5, 6])
def test_two_sweeps_every_count(nrelax, monkeypatch):
    """mg_cycle with odd and even sweep counts: pairs of sweeps in one launch
    and the odd one in the one-sweep kernel."""
    print("case sweep2 nrelax", nrelax, flush=True)
    one, two = _grids(7, 128, "mixed", "field", monkeypatch)
    assert one.residual() == two.residual()
    for it in range(2):
        one.mg_cycle(nrelax, 0)
        two.mg_cycle(nrelax, 0)
        assert one.residual() == two.residual(), it
    assert np.array_equal(one.get_a(), two.get_a())
