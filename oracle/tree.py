"""ctypes front end of the adaptive-tree CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of
``bench.py`` may import this module.  It drives ``tree_oracle.c``, the
plain-C restatement of the reference's multigrid solver on quadtrees and
octrees (grid/tree.h, grid/tree-common.h, poisson.h); the header of that file
cites the reference per function.

Side numbering is the reference's: right=0, left=1, top=2, bottom=3,
front=4, back=5 (common.h:950-951).
"""
import ctypes
import os

import numpy as np

from oracle.oracle import MGStats, build, SMOOTHERS, BC_SYMMETRY, BC_DIRICHLET, BC_NEUMANN

_HERE = os.path.dirname(os.path.abspath(__file__))
GH = 2
T_A, T_B, T_RES, T_DA, T_LAM, T_C, T_G = range(7)


def t_s(dim, k=0):
    """Field index of user scalar k (0..NUS-1)."""
    i = _lib(dim).ot_user_scalar(int(k))  # the same in every build
    if i < 0:
        raise ValueError("user scalar %d outside 0..%d" % (k, NUS - 1))
    return i


NUS = 4  # user scalars
BK_A, BK_HOM, BK_SYM = range(3)
RIGHT, LEFT, TOP, BOTTOM, FRONT, BACK = range(6)

_libs = {}


def _planted_path(dim, plant):
    """Build tree_oracle.c with -DORC_PLANT=plant (one deliberate mistake,
    see the header of tree_oracle.c) into a temporary directory, for the
    tests that show their checks catch it."""
    import subprocess
    import tempfile
    out = os.path.join(tempfile.gettempdir(), "bmg_orc_plant_%d" % os.getpid())
    os.makedirs(out, exist_ok=True)
    path = os.path.join(out, "liborc_tree%d_plant%d.so" % (dim, plant))
    if not os.path.exists(path):
        src = os.path.join(_HERE, "tree_oracle.c")
        subprocess.run(["gcc", "-O2", "-std=gnu99", "-fPIC", "-shared", "-ffp-contract=off",
                        "-Wno-unknown-pragmas", "-DDIM=%d" % dim, "-DORC_PLANT=%d" % plant,
                        "-o", path + ".tmp", src, "-lm"], check=True)
        os.replace(path + ".tmp", path)
    return path


def _lib(dim, plant=0):
    if (dim, plant) not in _libs:
        if plant:
            path = _planted_path(dim, plant)
        else:
            path = os.path.join(_HERE, "liborc_tree%d.so" % dim)
            if not os.path.exists(path):
                build()
        lib = ctypes.CDLL(path)
        P, I, D, L = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_long
        vp = ctypes.c_void_p
        sigs = {
            "ot_new": (P, [D, D, D, D]), "ot_free": (None, [P]),
            "ot_depth": (I, [P]), "ot_set_niter": (None, [P, I, I]),
            "ot_set_tolerance": (None, [P, D]), "ot_set_smoother": (None, [P, I]),
            "ot_set_lambda_const": (None, [P, D]), "ot_set_alpha_const": (None, [P, D]),
            "ot_use_lambda_field": (None, [P]),
            "ot_update": (None, [P]), "ot_nleaves": (L, [P]),
            "ot_leaves": (None, [P, vp, vp]), "ot_refine_marked": (I, [P, vp]),
            "ot_cache_count": (I, [P, I, I]), "ot_cache_get": (None, [P, I, I, vp]),
            "ot_face_count": (I, [P]), "ot_face_get": (None, [P, vp]),
            "ot_level_meta": (None, [P, I, vp, vp, vp, vp]),
            "ot_set_leaf": (None, [P, I, vp]), "ot_get_leaf": (None, [P, I, vp]),
            "ot_set_faces": (None, [P, vp]), "ot_set_bc": (None, [P, I, I, vp]),
            "ot_level_field": (ctypes.POINTER(ctypes.c_double), [P, I, I]),
            "ot_restriction_alpha_lambda": (None, [P]),
            "ot_relax": (None, [P, I]), "ot_residual": (D, [P]),
            "ot_mg_cycle": (None, [P, I, I, I]), "ot_sum_b": (D, [P]),
            "ot_mg_solve": (I, [P, D, I, I, ctypes.POINTER(MGStats)]),
            "ot_poisson": (I, [P, D, I, I, ctypes.POINTER(MGStats)]),
            "ot_project": (I, [P, vp, D, I, ctypes.POINTER(MGStats)]),
            "ot_residual_transpose": (None, [P, vp, vp]),
            "ot_mg_cycle_T": (None, [P, vp, vp, I, I, I]),
            "ot_adjoint": (I, [P, vp, D, I, I, vp, vp, vp, ctypes.POINTER(MGStats)]),
            "ot_boundary": (None, [P, I, I, I]),
            "ot_adapt_wavelet": (I, [P, D, I, I, ctypes.POINTER(I)]),
            "ot_adapt_wavelet_list": (I, [P, I, vp, vp, I, I, ctypes.POINTER(I)]),
            "ot_user_scalar": (I, [I]),
            "ot_refine_cell": (I, [P, I, I, I, I]),
            "ot_nonleaf": (L, [P, vp, vp]), "ot_unrefine": (I, [P, vp]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _libs[(dim, plant)] = lib
    return _libs[(dim, plant)]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleTree:
    """An adaptive tree built like init_grid(1) followed by refine(cond).

    ``refine(cond)`` takes a vectorised condition ``cond(x, y, z, level) ->
    bool array`` over the current leaves and repeats passes until no leaf
    satisfies it (the refine() macro, grid/tree-common.h:308-326).
    """

    def __init__(self, dim, L0=1.0, origin=(0.0, 0.0, 0.0), plant=0):
        self.dim = dim
        self.plant = plant
        self.lib = _lib(dim, plant)
        o = list(origin) + [0.0] * (3 - len(origin))
        self.L0, self.origin = L0, o[:3]
        self.t = self.lib.ot_new(L0, *o[:3])

    def __del__(self):
        if getattr(self, "t", None):
            self.lib.ot_free(self.t)
            self.t = None

    # -- topology
    def depth(self):
        return self.lib.ot_depth(self.t)

    def nleaves(self):
        return self.lib.ot_nleaves(self.t)

    def leaves(self):
        n = self.nleaves()
        lev = np.zeros(n, np.int32)
        xyz = np.zeros((n, 3))
        self.lib.ot_leaves(self.t, _ptr(lev), _ptr(xyz))
        return lev, xyz

    def refine(self, cond):
        while True:
            lev, xyz = self.leaves()
            mark = np.ascontiguousarray(np.asarray(cond(xyz[:, 0], xyz[:, 1], xyz[:, 2], lev), dtype=np.uint8))
            if not mark.any():
                return
            self.lib.ot_refine_marked(self.t, _ptr(mark))

    def cache(self, which, level):
        """which: 'active', 'prolongation' or 'boundary' -> (n, 3) int array."""
        w = {"active": 0, "prolongation": 1, "boundary": 2}[which]
        n = self.lib.ot_cache_count(self.t, w, level)
        out = np.zeros((n, 3), np.int32)
        if n:
            self.lib.ot_cache_get(self.t, w, level, _ptr(out))
        return out

    def faces(self):
        n = self.lib.ot_face_count(self.t)
        out = np.zeros((n, 5), np.int32)
        if n:
            self.lib.ot_face_get(self.t, _ptr(out))
        return out

    def level_meta(self, level):
        m = (1 << level) + 2 * GH
        shape = (m,) * self.dim
        alloc = np.zeros(shape, np.uint8)
        flags = np.zeros(shape, np.uint8)
        nb = np.zeros(shape, np.uint16)
        pid = np.zeros(shape, np.int8)
        self.lib.ot_level_meta(self.t, level, _ptr(alloc), _ptr(flags), _ptr(nb), _ptr(pid))
        return alloc, flags, nb, pid

    def face_centres(self):
        """Coordinates of every face value of the face cache, in the order
        ot_set_faces consumes them (entry order, then dimension); returns
        (xyz (n,3), dim (n,))."""
        f = self.faces()
        xs, ds = [], []
        for lev, i, j, k, flags in f:
            D = self.L0 * (1.0 / (1 << int(lev)))
            idx = (i, j, k)
            for d in range(self.dim):
                if flags & (1 << d):
                    c = []
                    for e in range(3):
                        if e >= self.dim:
                            c.append(0.0)
                            continue
                        g = -1 if e == d else 0
                        c.append(((g / 2.0 + (idx[e] - GH)) + 0.5) * D + self.origin[e])
                    xs.append(c)
                    ds.append(d)
        return np.array(xs).reshape(-1, 3), np.array(ds, np.int32)

    # -- fields
    def set_leaf(self, field, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        assert v.size == self.nleaves()
        self.lib.ot_set_leaf(self.t, field, _ptr(v))

    def get_leaf(self, field):
        v = np.zeros(self.nleaves())
        self.lib.ot_get_leaf(self.t, field, _ptr(v))
        return v

    def set_a(self, v):
        self.set_leaf(T_A, v)

    def set_b(self, v):
        self.set_leaf(T_B, v)

    def get_a(self):
        return self.get_leaf(T_A)

    def get_res(self):
        return self.get_leaf(T_RES)

    def set_alpha_faces(self, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        self.lib.ot_set_faces(self.t, _ptr(v))

    def set_alpha_function(self, f):
        """alpha.x[] = f(x, y, z, d) on the face cache (a foreach_face loop)."""
        xyz, ds = self.face_centres()
        self.set_alpha_faces(f(xyz[:, 0], xyz[:, 1], xyz[:, 2], ds))

    def set_lambda(self, v):
        if np.isscalar(v):
            self.lib.ot_set_lambda_const(self.t, float(v))
        else:
            self.set_leaf(T_LAM, v)
            self.lib.ot_use_lambda_field(self.t)

    def set_alpha_const(self, v):
        self.lib.ot_set_alpha_const(self.t, float(v))

    def bc_face_centres(self, side):
        """Face centres of side `side` for every level 0..depth, concatenated
        in the order ot_set_bc reads them."""
        d = side // 2
        hi = side % 2 == 0
        out = []
        for lev in range(self.depth() + 1):
            n = 1 << lev
            D = self.L0 * (1.0 / n)
            tang = [e for e in range(self.dim) if e != d]
            grids = np.meshgrid(*[np.arange(n)] * (self.dim - 1), indexing="ij")
            c = np.zeros((n ** (self.dim - 1), 3))
            for t, e in enumerate(tang):
                c[:, e] = (grids[t].ravel() + 0.5) * D + self.origin[e]
            # face between the interior cell and its ghost: ig = +-1
            i_cell = n - 1 if hi else 0
            g = 1 if hi else -1
            c[:, d] = ((g / 2.0 + i_cell) + 0.5) * D + self.origin[d]
            out.append(c)
        return np.concatenate(out)

    def set_bc(self, side, kind, values=None):
        v = None if values is None else np.ascontiguousarray(values, dtype=np.float64)
        self.lib.ot_set_bc(self.t, side, kind, None if v is None else _ptr(v))
        self._keep = v

    def set_bc_function(self, side, kind, f):
        c = self.bc_face_centres(side)
        self.set_bc(side, kind, f(c[:, 0], c[:, 1], c[:, 2]))

    def set_smoother(self, name):
        self.lib.ot_set_smoother(self.t, SMOOTHERS[name])

    def set_niter(self, nmax, nmin=1):
        self.lib.ot_set_niter(self.t, nmax, nmin)

    def set_tolerance(self, tol):
        self.lib.ot_set_tolerance(self.t, tol)

    # -- solver
    def residual(self):
        return self.lib.ot_residual(self.t)

    def mg_cycle(self, nrelax=4, minlevel=0, maxlevel=None):
        if maxlevel is None:
            maxlevel = self.depth()
        self.lib.ot_mg_cycle(self.t, nrelax, minlevel, maxlevel)

    def restriction_alpha_lambda(self):
        self.lib.ot_restriction_alpha_lambda(self.t)

    def mg_solve(self, tolerance=0.0, nrelax=4, minlevel=0):
        s = MGStats()
        self.lib.ot_mg_solve(self.t, tolerance, nrelax, minlevel, ctypes.byref(s))
        return s.as_dict()

    def project(self, uf, dt=1.0, nrelax=0):
        """project(), poisson.h:469-509, with p = a: uf holds the face
        velocity on the face cache (one value per set flag bit, like
        set_alpha_faces) and is updated in place; returns the stats."""
        v = np.ascontiguousarray(uf, dtype=np.float64)
        if v.size != self.nfacevalues():
            raise ValueError("expected %d face values" % self.nfacevalues())
        s = MGStats()
        self.lib.ot_project(self.t, _ptr(v), float(dt), int(nrelax), ctypes.byref(s))
        uf[...] = v
        return s.as_dict()

    def residual_transpose(self, y):
        """r^T y for the leaf operator r(x) = -lambda x - div(alpha grad x)
        with homogeneous boundary conditions (leaf order in and out)."""
        yv = np.ascontiguousarray(y, dtype=np.float64)
        out = np.zeros(self.nleaves())
        self.lib.ot_residual_transpose(self.t, _ptr(yv), _ptr(out))
        return out

    def mg_cycle_transpose(self, y, nrelax=4, minlevel=1):
        """The transpose of one V-cycle from da = 0 (res -> da on the leaves,
        red-black relax) applied to y."""
        yv = np.ascontiguousarray(y, dtype=np.float64)
        out = np.zeros(self.nleaves())
        self.lib.ot_mg_cycle_T(self.t, _ptr(yv), _ptr(out), int(nrelax), int(minlevel), self.depth())
        return out

    def adjoint(self, g, tolerance=1e-10, nitermax=100, nrelax=4, alpha=True, lam=False):
        """The adjoint of the last poisson() solve: psi with L^T psi = g
        (L = -r), then dJ/db = psi, dJ/dalpha on the face values and
        dJ/dlambda at the leaves.  Returns (psi, galpha or None, glambda
        or None, stats)."""
        gv = np.ascontiguousarray(g, dtype=np.float64)
        psi = np.zeros(self.nleaves())
        ga = np.zeros(self.nfacevalues()) if alpha else None
        gl = np.zeros(self.nleaves()) if lam else None
        s = MGStats()
        self.lib.ot_adjoint(self.t, _ptr(gv), float(tolerance), int(nitermax), int(nrelax), _ptr(psi),
                            _ptr(ga) if alpha else None, _ptr(gl) if lam else None, ctypes.byref(s))
        return psi, ga, gl, s.as_dict()

    def nfacevalues(self):
        f = self.faces()
        return int(sum(((f[:, 4] >> d) & 1).sum() for d in range(self.dim)))

    def poisson(self, tolerance=0.0, nrelax=4, minlevel=0):
        s = MGStats()
        self.lib.ot_poisson(self.t, tolerance, nrelax, minlevel, ctypes.byref(s))
        return s.as_dict()

    # -- adaptation
    def set_s(self, v, k=0):
        """Leaf values of user scalar k (0..NUS-1), which adapt_wavelet may
        estimate and which follows the mesh."""
        self.set_leaf(t_s(self.dim, k), v)

    def get_s(self, k=0):
        return self.get_leaf(t_s(self.dim, k))

    def get_b(self):
        return self.get_leaf(T_B)

    def adapt_wavelet(self, max, maxlevel, minlevel=1, scalars=None):
        """adapt_wavelet (slist, max, maxlevel, minlevel) with slist the
        user scalars `scalars` (default [0]) and `max` one tolerance per
        scalar (or one number), every user scalar, a and b following the
        mesh (grid/tree-common.h:165-303); returns (nf, nc)."""
        us = [0] if scalars is None else list(scalars)
        mx = [float(max)] * len(us) if np.isscalar(max) else [float(m) for m in max]
        if len(mx) != len(us):
            raise ValueError("one tolerance per estimated scalar")
        ua = np.array(us, np.int32)
        ma = np.array(mx, np.float64)
        nf = ctypes.c_int(0)
        nc = self.lib.ot_adapt_wavelet_list(self.t, len(us), ua.ctypes.data, ma.ctypes.data, int(maxlevel),
                                            int(minlevel), ctypes.byref(nf))
        if nc < 0:
            raise ValueError("bad user scalar index")
        return nf.value, nc

    def refine_cell_at(self, level, c):
        """refine_cell of the leaf (level, i, j, k) with the 2:1 cascade."""
        return self.lib.ot_refine_cell(self.t, int(level), int(c[0]), int(c[1]), int(c[2]) if self.dim == 3 else 0)

    def nonleaf(self):
        """Levels and centres of the non-leaf cells foreach_cell() visits."""
        n = self.lib.ot_nonleaf(self.t, None, None)
        lev = np.zeros(n, np.int32)
        xyz = np.zeros((n, 3))
        if n:
            self.lib.ot_nonleaf(self.t, _ptr(lev), _ptr(xyz))
        return lev, xyz

    def unrefine(self, cond):
        """unrefine (cond), grid/tree-common.h:348-371: cond(x, y, z, level)
        on the non-leaf cells; returns the number of coarsened cells."""
        lev, xyz = self.nonleaf()
        mark = np.ascontiguousarray(cond(xyz[:, 0], xyz[:, 1], xyz[:, 2], lev), dtype=np.uint8)
        return self.lib.ot_unrefine(self.t, _ptr(mark))
