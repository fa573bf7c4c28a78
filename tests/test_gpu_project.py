"""project() on the GPU (bmg_project, poisson.h:469-509) against the
oracle's project(): the FP64 path gives the oracle's face velocities,
pressure and statistics bit for bit; the mixed path the same statistics
within one cycle and the same face velocities to the solver tolerance."""
import numpy as np
import pytest

import basilisk_amd as bmg
from oracle.oracle import OracleGrid
from tests import cases

pytestmark = pytest.mark.gpu

KINDS = {
    "symmetry": lambda dim: [cases.BC_SYMMETRY] * (2 * dim),
    "dirichlet_neumann": lambda dim: [cases.BC_DIRICHLET, cases.BC_NEUMANN] * dim,
    "periodic": lambda dim: [cases.BC_PERIODIC] * 2 + [cases.BC_SYMMETRY] * (2 * dim - 2),
}


def _setup(grid, dim, bcs, seed, alpha):
    rng = np.random.default_rng(seed)
    cases.random_problem(grid, seed, bcs=KINDS[bcs](dim), alpha=False)
    if alpha:
        al = [np.exp(0.5 * rng.standard_normal(grid.face_shape(c))) for c in range(dim)]
        if bcs == "periodic":
            al[0][-1] = al[0][0]
        grid.set_alpha(al)
    grid.set_a(np.zeros(grid.shape()))
    uf = []
    for c in range(dim):
        u = rng.standard_normal(grid.face_shape(c))
        if bcs != "dirichlet_neumann":
            idx = [slice(None)] * dim
            idx[c] = 0
            u[tuple(idx)] = 0.0
            idx[c] = -1
            u[tuple(idx)] = 0.0
        uf.append(np.ascontiguousarray(u))
    return uf


@pytest.mark.parametrize("dim,depth", [(2, 7), (3, 5)])
@pytest.mark.parametrize("bcs", sorted(KINDS))
@pytest.mark.parametrize("dt,alpha", [(1.0, True), (0.37, True), (0.5, False)])
def test_project_bitexact(dim, depth, bcs, dt, alpha):
    print("case project", dim, depth, bcs, dt, alpha, flush=True)
    g = bmg.Grid(dim, depth, smoother="rbgs")
    o = OracleGrid(dim, depth)
    o.set_smoother("rbgs")
    ug = _setup(g, dim, bcs, 17, alpha)
    uo = _setup(o, dim, bcs, 17, alpha)
    for grid in (g, o):
        grid.set_tolerance(1e-6)
    sg = g.project(ug, dt=dt)
    so = o.project(uo, dt=dt)
    for k in ("i", "resb", "resa", "nrelax", "minlevel"):
        assert sg[k] == so[k], (k, sg, so)
    assert np.array_equal(g.get_a(), o.get_a())
    for c in range(dim):
        assert np.array_equal(ug[c], uo[c]), c


def test_project_device_tensors_and_mixed():
    """torch CUDA face components are projected in place; the mixed-precision
    grid reaches the same tolerance within one cycle, leaves face velocities
    whose divergence over dt is its FP64 residual (so below the tolerance)
    and within 1e-7 (relative) of the FP64 ones."""
    import torch
    dim, depth = 3, 6
    out = {}
    for prec in ("fp64", "mixed"):
        g = bmg.Grid(dim, depth, smoother="rbgs", precision=prec)
        uf = _setup(g, dim, "symmetry", 23, True)
        g.set_tolerance(1e-9)
        ud = [torch.from_numpy(u).cuda() for u in uf]
        st = g.project(ud, dt=0.25)
        out[prec] = (st, [u.cpu().numpy() for u in ud])
    (s64, u64), (smx, umx) = out["fp64"], out["mixed"]
    assert smx["resa"] <= 1e-9 / 0.25 ** 2
    assert abs(smx["i"] - s64["i"]) <= 1
    scale = max(np.abs(u).max() for u in u64)
    assert max(np.abs(a - b).max() for a, b in zip(u64, umx)) <= 1e-7 * scale
    div = sum(np.diff(u, axis=c) for c, u in enumerate(umx)) * (1 << depth) / 0.25
    assert np.abs(div).max() <= 1.001 * smx["resa"] + 1e-12 * scale * (1 << depth)
    # the device path equals the host path bit for bit (FP64)
    g = bmg.Grid(dim, depth, smoother="rbgs")
    uf = _setup(g, dim, "symmetry", 23, True)
    g.set_tolerance(1e-9)
    sh = g.project(uf, dt=0.25)
    assert sh == s64
    assert all(np.array_equal(a, b) for a, b in zip(uf, u64))


@pytest.mark.parametrize("dim,depth", [(2, 7), (3, 5)])
def test_project_time_steps_with_nrelax_feedback(dim, depth):
    """The projection step of navier-stokes/centered.h over several time
    steps: each step adds a seeded perturbation to the face velocity and
    calls project() with the previous step's mgstats.nrelax (and the previous
    pressure as the initial guess), as `mgp = project (uf, p, alpha, dt,
    mgp.nrelax)` does; every step equals the oracle's bit for bit."""
    print("case project steps", dim, depth, flush=True)
    g = bmg.Grid(dim, depth, smoother="rbgs")
    o = OracleGrid(dim, depth)
    o.set_smoother("rbgs")
    ug = _setup(g, dim, "symmetry", 31, True)
    uo = _setup(o, dim, "symmetry", 31, True)
    for grid in (g, o):
        grid.set_tolerance(1e-7)
    rng = np.random.default_rng(77)
    ng = no = 0
    for step in range(4):
        for c in range(dim):
            pert = 0.1 * rng.standard_normal(ug[c].shape)
            idx = [slice(None)] * dim
            idx[c] = 0
            pert[tuple(idx)] = 0.0
            idx[c] = -1
            pert[tuple(idx)] = 0.0
            ug[c] += pert
            uo[c] += pert
        sg = g.project(ug, dt=0.05, nrelax=ng)
        so = o.project(uo, dt=0.05, nrelax=no)
        for k in ("i", "resb", "resa", "nrelax"):
            assert sg[k] == so[k], (step, k, sg, so)
        ng, no = sg["nrelax"], so["nrelax"]
        assert np.array_equal(g.get_a(), o.get_a())
        for c in range(dim):
            assert np.array_equal(ug[c], uo[c]), (step, c)


@pytest.mark.parametrize("dim,depth", [(2, 7), (3, 5)])
@pytest.mark.parametrize("bcs", sorted(KINDS))
@pytest.mark.parametrize("acc", [False, True])
def test_centered_gradient_bitexact(dim, depth, bcs, acc):
    """centered_gradient (navier-stokes/centered.h:387-406) of a random
    pressure with every boundary kind, a variable alpha and a face
    acceleration: the GPU's g equals the oracle's bit for bit, from host
    arrays and from device tensors."""
    import torch
    from oracle import oracle as orc
    print("case centered_gradient", dim, depth, bcs, acc, flush=True)
    g = bmg.Grid(dim, depth)
    o = OracleGrid(dim, depth)
    for grid in (g, o):
        _setup(grid, dim, bcs, 29, True)
    rng = np.random.default_rng(5)
    p = rng.standard_normal(g.shape())
    g.set_a(p)
    o.set_a(p)
    af = [np.ascontiguousarray(rng.standard_normal(g.face_shape(c))) for c in range(dim)] if acc else None
    alpha = [o.get_face(c) for c in range(dim)]
    want = orc.centered_gradient(o, alpha, af)
    out = [np.zeros(g.shape()) for _ in range(dim)]
    g.centered_gradient(out, af)
    for c in range(dim):
        assert np.array_equal(out[c], want[c]), c
    dev = [torch.zeros(g.shape(), dtype=torch.float64, device="cuda:0") for _ in range(dim)]
    g.centered_gradient(dev, None if af is None else [torch.from_numpy(x).cuda() for x in af])
    for c in range(dim):
        assert np.array_equal(dev[c].cpu().numpy(), want[c]), c


@pytest.mark.parametrize("dim,depth", [(2, 7), (3, 5)])
def test_centered_time_step_bitexact(dim, depth):
    """The solver's part of a navier-stokes/centered.h time step: the
    half-step project (uf, pf, alpha, dt/2.) of advection_term
    (centered.h:307-313) on the pressure pf, then the projection event
    (centered.h:412-420): project (uf, p, alpha, dt), centered_gradient (p,
    g), correction (dt) -- twice, each with its own pressure field and nrelax
    fed back as the reference does (mgpf, mgp).  Every face velocity,
    pressure, gradient and cell velocity equals the oracle's bit for bit."""
    from oracle import oracle as orc
    print("case centered time step", dim, depth, flush=True)
    dt = 0.3
    grids = {}
    for name in ("p", "pf"):
        g = bmg.Grid(dim, depth, smoother="rbgs")
        o = OracleGrid(dim, depth)
        o.set_smoother("rbgs")
        for grid in (g, o):
            _setup(grid, dim, "symmetry", 31, True)
            grid.set_tolerance(1e-6)
        grids[name] = (g, o)
    rng = np.random.default_rng(8)
    ug = [np.ascontiguousarray(x) for x in _setup(OracleGrid(dim, depth), dim, "symmetry", 37, False)]
    uo = [x.copy() for x in ug]
    u_g = [rng.standard_normal(grids["p"][0].shape()) for _ in range(dim)]
    u_o = [x.copy() for x in u_g]
    alpha = [grids["p"][1].get_face(c) for c in range(dim)]
    nr = {"p": {"g": 0, "o": 0}, "pf": {"g": 0, "o": 0}}
    for step in range(2):
        gpf, opf = grids["pf"]
        s1 = gpf.project(ug, dt=dt / 2.0, nrelax=nr["pf"]["g"])
        s2 = opf.project(uo, dt=dt / 2.0, nrelax=nr["pf"]["o"])
        assert s1["i"] == s2["i"] and s1["resa"] == s2["resa"]
        nr["pf"] = {"g": s1["nrelax"], "o": s2["nrelax"]}
        gp, op = grids["p"]
        s1 = gp.project(ug, dt=dt, nrelax=nr["p"]["g"])
        s2 = op.project(uo, dt=dt, nrelax=nr["p"]["o"])
        assert s1["i"] == s2["i"] and s1["resa"] == s2["resa"]
        nr["p"] = {"g": s1["nrelax"], "o": s2["nrelax"]}
        gg = [np.zeros(gp.shape()) for _ in range(dim)]
        gp.centered_gradient(gg)
        go = orc.centered_gradient(op, alpha)
        gp.correction(u_g, gg, dt)
        u_o = orc.correction(u_o, go, dt)
        for c in range(dim):
            assert np.array_equal(ug[c], uo[c]), (step, c)
            assert np.array_equal(gg[c], go[c]), (step, c)
            assert np.array_equal(u_g[c], u_o[c]), (step, c)
        assert np.array_equal(gp.get_a(), op.get_a())
        assert np.array_equal(gpf.get_a(), opf.get_a())
