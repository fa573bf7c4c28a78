"""project() on adaptive trees in the oracle (OracleTree.project,
poisson.h:469-509 with the tree's halo_face and tree boundary), pinned by
what the mathematics fixes -- the reference has no fixture for it:

- a face velocity that is already dt times the discrete flux
  alpha (q[] - q[-1])/Delta of a pressure q with p's boundary conditions is
  removed entirely: p = q + const and uf -> 0, resolution jumps included;
- on a tree whose leaves all sit on one level it equals the uniform-grid
  oracle's project() in full-tree mode bit for bit (the uniform grid's
  project() is pinned separately, tests/test_project_oracle.py);
- dt = 0 and nrelax = 0 take the reference's defaults."""
import numpy as np
import pytest

from oracle.oracle import OracleGrid
from oracle.tree import OracleTree, T_G, GH
from tests import tree_cases


def face_fluxes(t):
    """alpha (a[] - a[-1])/Delta on the face cache (entry order, then
    dimension), read from the oracle's residual, which computes exactly
    these fluxes (poisson.h:355-358) before halo_face."""
    t.residual()
    out = []
    for lev, i, j, k, flags in t.faces():
        n = (1 << int(lev)) + 2 * GH
        for d in range(t.dim):
            if flags & (1 << d):
                g = t.lib.ot_level_field(t.t, int(lev), T_G + d)
                idx = ((i * n + j) * n + k) if t.dim == 3 else (i * n + j)
                out.append(g[idx])
    return np.array(out)


@pytest.mark.parametrize("dim,depth", [(2, 6), (3, 4)])
@pytest.mark.parametrize("dt", [1.0, 0.4])
def test_gradient_field_is_removed(dim, depth, dt):
    t = OracleTree(dim, 1.0, (0.0, 0.0, 0.0))
    t.set_smoother("rbgs")
    tree_cases.spheres_problem(t, depth, 11, nspheres=3, minlevel=2, lam=False)
    t.set_bc(0, 0, None)  # symmetry on the right
    t.set_bc(1, 0, None)  # and on the left: pure Neumann
    _, xyz = t.leaves()
    q = np.sin(2.0 * xyz[:, 0]) + np.cos(3.0 * xyz[:, 1]) + 0.5 * xyz[:, 2]
    t.set_a(q)
    t.restriction_alpha_lambda()
    uf = dt * face_fluxes(t)
    t.set_a(np.zeros_like(q))
    t.set_tolerance(1e-10)
    st = t.project(uf, dt=dt)
    p = t.get_a()
    assert st["resa"] <= 1e-10 / dt ** 2
    assert np.abs((p - p.mean()) - (q - q.mean())).max() < 1e-7
    assert np.abs(uf).max() < 1e-7 * max(1.0, dt)


def face_arrays_from_cache(t, values, n):
    """the face cache's values of a full tree as uniform face arrays"""
    dim = t.dim
    comps = [np.zeros(tuple(n + (c == e) for e in range(dim))) for c in range(dim)]
    v = 0
    for lev, i, j, k, flags in t.faces():
        for d in range(dim):
            if flags & (1 << d):
                comps[d][(i - GH, j - GH, k - GH)[:dim]] = values[v]
                v += 1
    return comps


@pytest.mark.parametrize("dim,depth", [(2, 5), (3, 3)])
def test_full_tree_equals_uniform_oracle(dim, depth):
    rng = np.random.default_rng(2)
    t = OracleTree(dim, 1.0, (0.0, 0.0, 0.0))
    t.set_smoother("rbgs")
    t.refine(lambda x, y, z, lev: lev < depth)
    o = OracleGrid(dim, depth)
    o.set_smoother("rbgs")
    o.set_tree(True)
    n = 1 << depth
    nv = sum(int(((t.faces()[:, 4] >> d) & 1).sum()) for d in range(dim))
    alpha_v = np.exp(0.3 * rng.standard_normal(nv))
    uf_v = rng.standard_normal(nv)
    t.set_alpha_faces(alpha_v)
    o.set_alpha(face_arrays_from_cache(t, alpha_v, n))
    t.set_a(np.zeros(t.nleaves()))
    o.set_a(np.zeros(o.shape()))
    for grid in (t, o):
        grid.set_tolerance(1e-8)
    uf_t = uf_v.copy()
    uf_o = face_arrays_from_cache(t, uf_v, n)
    st = t.project(uf_t, dt=0.5)
    so = o.project(uf_o, dt=0.5)
    assert st["i"] == so["i"] and st["resa"] == so["resa"]
    want = face_arrays_from_cache(t, uf_t, n)
    for c in range(dim):
        assert np.array_equal(want[c], uf_o[c]), c
    # leaves of a full tree are in Morton order; compare through coordinates
    _, xyz = t.leaves()
    idx = tuple(np.floor(xyz[:, d] * n).astype(int) for d in range(dim))
    assert np.array_equal(t.get_a(), o.get_a()[idx])


def test_defaults():
    t = OracleTree(2, 1.0, (0.0, 0.0, 0.0))
    t.set_smoother("rbgs")
    t.refine(lambda x, y, z, lev: (lev < 3) | ((lev < 5) & (x < 0.4)))
    nv = sum(int(((t.faces()[:, 4] >> d) & 1).sum()) for d in range(2))
    rng = np.random.default_rng(3)
    u0 = rng.standard_normal(nv)
    out = []
    for dt, nrelax in ((0.0, 0), (1.0, 4)):
        t.set_a(np.zeros(t.nleaves()))
        u = u0.copy()
        st = t.project(u, dt=dt, nrelax=nrelax)
        out.append((st, u, t.get_a()))
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][2], out[1][2])
