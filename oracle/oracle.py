"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of
``bench.py`` may import this module.  It drives ``mg_oracle.c``, the plain-C
restatement of the reference's multigrid Poisson-Helmholtz solver
(poisson.h:32-441, grid/multigrid.h, grid/multigrid-common.h); see the header
of that file for the per-function citations.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

F_A, F_B, F_RES, F_DA, F_LAMBDA, F_C, F_AX = range(7)
BC_SYMMETRY, BC_DIRICHLET, BC_NEUMANN, BC_PERIODIC = range(4)
BC_VALUE = 5  # the ghost is the expression itself (a[left] = expr)
SMOOTHERS = {"gs": 0, "jacobi": 1, "rbgs": 2}
GHOSTS = 2


class MGStats(ctypes.Structure):
    """mgstats, poisson.h:108-114."""
    _fields_ = [("i", ctypes.c_int), ("resb", ctypes.c_double),
                ("resa", ctypes.c_double), ("sum", ctypes.c_double),
                ("nrelax", ctypes.c_int), ("minlevel", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def build():
    """Compile the oracle libraries (make in this directory)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_libs = {}


def _planted_path(dim, plant):
    """Build mg_oracle.c with -DORC_PLANT=plant (one deliberate mistake, see
    the header of mg_oracle.c) into a temporary directory, for the tests
    that show their checks catch it."""
    import tempfile
    out = os.path.join(tempfile.gettempdir(), "bmg_orc_plant_%d" % os.getpid())
    os.makedirs(out, exist_ok=True)
    path = os.path.join(out, "liborc_mg%d_plant%d.so" % (dim, plant))
    if not os.path.exists(path):
        src = os.path.join(_HERE, "mg_oracle.c")
        subprocess.run(["gcc", "-O2", "-std=gnu99", "-fPIC", "-shared", "-ffp-contract=off",
                        "-Wno-unknown-pragmas", "-DDIM=%d" % dim, "-DORC_PLANT=%d" % plant,
                        "-o", path + ".tmp", src, "-lm"], check=True)
        os.replace(path + ".tmp", path)
    return path


def _lib(dim, omp=False, plant=0):
    key = (dim, omp, plant)
    if key not in _libs:
        if plant:
            path = _planted_path(dim, plant)
        else:
            name = "liborc_mg%d%s.so" % (dim, "_omp" if omp else "")
            path = os.path.join(_HERE, name)
            if not os.path.exists(path):
                build()
        lib = ctypes.CDLL(path)
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        dp = ctypes.POINTER(ctypes.c_double)
        sigs = {
            "orc_new": (P, [I]), "orc_free": (None, [P]),
            "orc_set_geometry": (None, [P, D, D, D, D]),
            "orc_set_niter": (None, [P, I, I]),
            "orc_set_tolerance": (None, [P, D]),
            "orc_set_smoother": (None, [P, I]),
            "orc_set_tree": (None, [P, I]),
            "orc_set_lambda_const": (None, [P, D]),
            "orc_use_lambda_field": (None, [P]),
            "orc_set_alpha_const": (None, [P, D]),
            "orc_use_alpha_field": (None, [P]),
            "orc_set_bc": (None, [P, I, I, dp]),
            "orc_set_cell": (None, [P, I, I, dp]),
            "orc_get_cell": (None, [P, I, I, dp]),
            "orc_set_face": (None, [P, I, I, dp]),
            "orc_get_face": (None, [P, I, I, dp]),
            "orc_level_data": (dp, [P, I]),
            "orc_boundary": (None, [P, I, I, I]),
            "orc_restriction_res": (None, [P]),
            "orc_restriction_alpha_lambda": (None, [P]),
            "orc_relax": (None, [P, I]),
            "orc_residual": (D, [P]),
            "orc_mg_cycle": (None, [P, I, I, I]),
            "orc_sum_b": (D, [P]),
            "orc_mg_solve": (I, [P, D, I, I, ctypes.POINTER(MGStats)]),
            "orc_poisson": (I, [P, D, I, I, ctypes.POINTER(MGStats)]),
            "orc_project": (I, [P, ctypes.POINTER(dp), D, I, ctypes.POINTER(MGStats)]),
            "orc_nrelax_update": (I, [D, D, D, I]),
            "orc_set_threads": (None, [I]),
            "orc_max_threads": (I, []),
            "orc_dimension": (I, []),
            "orc_nvar": (I, []),
        }
        if dim == 2:
            # embedded boundaries (embed.h), 2D only
            sigs.update({
                "orc_set_embed": (None, [P, dp, dp, dp, I, dp, I]),
                "orc_embed_bad": (I, [P]),
                "orc_fractions": (None, [I, dp, D, dp, dp, dp]),
                "orc_embed_geometry": (None, [I, dp, dp, dp, dp, dp, dp, dp, dp]),
                "orc_cut_centroids": (None, [I, dp, dp, dp, dp, dp]),
                "orc_line_alpha": (D, [D, D, D]),
                "orc_line_area": (D, [D, D, D]),
            })
        for name_, (res, args) in sigs.items():
            fn = getattr(lib, name_)
            fn.restype = res
            fn.argtypes = args
        _libs[key] = lib
    return _libs[key]


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class OracleGrid:
    """A uniform multigrid hierarchy of `depth` levels in `dim` dimensions,
    holding the fields a, b, res, da, lambda and the face vector alpha."""

    def __init__(self, dim, depth, omp=False, L0=1.0, origin=(0.0, 0.0, 0.0), plant=0):
        self.dim, self.depth = dim, depth
        self.lib = _lib(dim, omp, plant)
        self.g = self.lib.orc_new(depth)
        self.lib.orc_set_geometry(self.g, L0, *origin)
        self.L0, self.origin = L0, tuple(origin)
        self.nv = self.lib.orc_nvar()

    def __del__(self):
        try:
            self.lib.orc_free(self.g)
        except Exception:
            pass

    # -- geometry helpers (cartesian-common.h:9-34)
    def n(self, level=None):
        return 1 << (self.depth if level is None else level)

    def delta(self, level=None):
        return self.L0 * (1.0 / self.n(level))

    def centers(self, level=None):
        """Cell-centre coordinates per dimension (1D arrays)."""
        n, D = self.n(level), self.delta(level)
        return [(np.arange(n) + 0.5) * D + self.origin[d] for d in range(self.dim)]

    def shape(self, level=None):
        return (self.n(level),) * self.dim

    def face_shape(self, c, level=None):
        s = list(self.shape(level))
        s[c] += 1
        return tuple(s)

    # -- field access
    def set_cell(self, f, arr, level=None):
        level = self.depth if level is None else level
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.shape(level))
        self.lib.orc_set_cell(self.g, f, level, _dptr(a))

    def get_cell(self, f, level=None):
        level = self.depth if level is None else level
        a = np.empty(self.shape(level))
        self.lib.orc_get_cell(self.g, f, level, _dptr(a))
        return a

    def set_face(self, c, arr, level=None):
        level = self.depth if level is None else level
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.face_shape(c, level))
        self.lib.orc_set_face(self.g, c, level, _dptr(a))
        self.lib.orc_use_alpha_field(self.g)

    def get_face(self, c, level=None):
        level = self.depth if level is None else level
        a = np.empty(self.face_shape(c, level))
        self.lib.orc_get_face(self.g, c, level, _dptr(a))
        return a

    def set_a(self, arr):
        self.set_cell(F_A, arr)

    def set_b(self, arr):
        self.set_cell(F_B, arr)

    def get_a(self):
        return self.get_cell(F_A)

    def get_res(self):
        return self.get_cell(F_RES)

    def set_alpha(self, comps):
        for c, arr in enumerate(comps):
            self.set_face(c, arr)

    def set_lambda(self, lam):
        if np.isscalar(lam):
            self.lib.orc_set_lambda_const(self.g, float(lam))
        else:
            self.set_cell(F_LAMBDA, lam)
            self.lib.orc_use_lambda_field(self.g)

    def raw_level(self, level):
        """The AoS storage of one level, shape (m,)*dim + (nvar,), m = n+4."""
        m = self.n(level) + 2 * GHOSTS
        ptr = self.lib.orc_level_data(self.g, level)
        count = m ** self.dim * self.nv
        buf = np.ctypeslib.as_array(ptr, shape=(count,))
        return buf.reshape((m,) * self.dim + (self.nv,))

    # -- boundary conditions of `a`
    def set_bc(self, direction, kind, values=None):
        """direction: 0 right, 1 left, 2 top, 3 bottom, 4 front, 5 back.
        values: the BC expression over the full tangential extent (n+4 per
        tangential dimension) at the finest level, or None for 0."""
        if values is None:
            self.lib.orc_set_bc(self.g, direction, kind, None)
        else:
            v = np.ascontiguousarray(values, dtype=np.float64)
            m = self.n() + 2 * GHOSTS
            if v.size != m ** (self.dim - 1):
                raise ValueError("expected %d boundary values (ghost-inclusive extent), got %d"
                                 % (m ** (self.dim - 1), v.size))
            self.lib.orc_set_bc(self.g, direction, kind, _dptr(v))

    def set_bc_function(self, kind, fn, directions=None):
        """Evaluate fn(x, y, z) on each boundary face at the finest level over
        the full tangential extent, with the face coordinate shifted by half a
        cell like the qcc-generated BC functions (qcc.lex:1118-1130,
        cartesian-common.h:9-34)."""
        dirs = range(2 * self.dim) if directions is None else directions
        n, D, m = self.n(), self.delta(), self.n() + 2 * GHOSTS
        idx = np.arange(m)
        for b in dirs:
            d, side = b // 2, (1 if b % 2 == 0 else -1)
            plane = n + GHOSTS - 1 if side > 0 else GHOSTS
            coords = []
            for e in range(3):
                if e >= self.dim:
                    coords.append(np.zeros(1))
                elif e == d:
                    coords.append(np.array([(side / 2.0 + plane - GHOSTS + 0.5) * D + self.origin[e]]))
                else:
                    coords.append((idx - GHOSTS + 0.5) * D + self.origin[e])
            X, Y, Z = np.meshgrid(*coords, indexing="ij")
            vals = np.asarray(fn(X, Y, Z), dtype=np.float64) * np.ones_like(X)
            self.set_bc(b, kind, vals.reshape(-1))

    def set_periodic(self, d):
        self.lib.orc_set_bc(self.g, 2 * d, BC_PERIODIC, None)

    # -- solver
    def set_smoother(self, name):
        self.lib.orc_set_smoother(self.g, SMOOTHERS[name])

    def set_tree(self, tree=True):
        """Full quadtree/octree semantics (grid/tree.h): Morton-order
        Gauss-Seidel and the tree's edge/corner ghost rule."""
        self.lib.orc_set_tree(self.g, int(tree))

    def set_niter(self, nitermax=100, nitermin=1):
        self.lib.orc_set_niter(self.g, nitermax, nitermin)

    def set_tolerance(self, tol):
        self.lib.orc_set_tolerance(self.g, tol)

    def residual(self):
        return self.lib.orc_residual(self.g)

    def mg_cycle(self, nrelax, minlevel, maxlevel=None):
        self.lib.orc_mg_cycle(self.g, nrelax, minlevel,
                              self.depth if maxlevel is None else maxlevel)

    def relax(self, level):
        self.lib.orc_relax(self.g, level)

    def boundary(self, f, level, homogeneous):
        self.lib.orc_boundary(self.g, f, level, int(homogeneous))

    def restriction_res(self):
        self.lib.orc_restriction_res(self.g)

    def restriction_alpha_lambda(self):
        self.lib.orc_restriction_alpha_lambda(self.g)

    def mg_solve(self, tolerance=0.0, nrelax=0, minlevel=0):
        s = MGStats()
        self.lib.orc_mg_solve(self.g, tolerance, nrelax, minlevel, ctypes.byref(s))
        return s.as_dict()

    def project(self, uf, dt=1.0, nrelax=0):
        """project(), poisson.h:469-509, with p = a and the grid's alpha:
        updates the face components `uf` (float64 arrays, n+1 faces along
        their own direction) in place and returns the poisson() stats."""
        arrs = []
        for c, u in enumerate(uf):
            if u.dtype != np.float64 or not u.flags.c_contiguous or u.shape != self.face_shape(c):
                raise ValueError("face component %d: need a C-contiguous float64 array of shape %s"
                                 % (c, self.face_shape(c)))
            arrs.append(u)
        ptrs = (ctypes.POINTER(ctypes.c_double) * self.dim)(*[_dptr(u) for u in arrs])
        s = MGStats()
        self.lib.orc_project(self.g, ptrs, float(dt), int(nrelax), ctypes.byref(s))
        return s.as_dict()

    def poisson(self, tolerance=0.0, nrelax=0, minlevel=0):
        s = MGStats()
        self.lib.orc_poisson(self.g, tolerance, nrelax, minlevel, ctypes.byref(s))
        return s.as_dict()

    def set_embed(self, cs, fx, fy, kind, bc=None, third=False):
        """The embedded boundary (2D): finest-level cs, fs_x, fs_y; kind
        EMBED_DIRICHLET or EMBED_NEUMANN with the condition's value per
        cell (at the fragment centroid; None: homogeneous); third: embedded
        face gradients in the residual (embed.h:179-195)."""
        args = [np.ascontiguousarray(a, dtype=np.float64) for a in (cs, fx, fy)]
        bcv = None if bc is None else np.ascontiguousarray(bc, dtype=np.float64)
        self._emb = args + [bcv]
        self.lib.orc_set_embed(self.g, *[_dptr(a) for a in args], int(kind),
                               None if bcv is None else _dptr(bcv), 1 if third else 0)

    def embed_bad(self):
        return bool(self.lib.orc_embed_bad(self.g))


def centered_gradient(o, alpha=None, acc=None, plant=0):
    """centered_gradient (p, g) of navier-stokes/centered.h:387-406 with
    p = the oracle grid's `a` (ghosts filled with its boundary conditions,
    grid/multigrid.h:543-594), fm = unityf and SEPS = 0 (common.h:1281,
    1287): on every face gf.x[] = fm.x[]*a.x[] - alpha.x[]*(p[] - p[-1])/Delta,
    then g.x[] = (gf.x[] + gf.x[1])/(fm.x[] + fm.x[1] + SEPS) per cell.
    alpha, acc: dim face arrays (n+1 along their component) or None for
    unityf / zerof.  plant (tests only): 1 flips the sign of the pressure
    term, 2 averages a cell's low face twice."""
    dim, n, G = o.dim, o.n(), GHOSTS
    o.boundary(F_A, o.depth, False)
    p = o.raw_level(o.depth)[..., F_A]
    D = o.delta()
    g = []
    for d in range(dim):
        hi = [slice(G, G + n)] * dim
        lo = [slice(G, G + n)] * dim
        hi[d] = slice(G, G + n + 1)
        lo[d] = slice(G - 1, G + n)
        ph, pl = p[tuple(hi)], p[tuple(lo)]
        al = 1.0 if alpha is None else alpha[d]
        ac = 0.0 if acc is None else acc[d]
        gf = 1.0 * ac + al * (ph - pl) / D if plant == 1 else 1.0 * ac - al * (ph - pl) / D
        c0 = [slice(None)] * dim
        c1 = [slice(None)] * dim
        c0[d] = slice(0, n)
        c1[d] = slice(0, n) if plant == 2 else slice(1, n + 1)
        g.append((gf[tuple(c0)] + gf[tuple(c1)]) / (1.0 + 1.0 + 0.0))
    return g


def correction(u, g, dt):
    """correction (dt), navier-stokes/centered.h:322-327: u.x[] += dt*g.x[]
    per cell and component (returns new arrays)."""
    return [x + dt * y for x, y in zip(u, g)]


def nrelax_update(resb, resa, tolerance, nrelax, plant=0):
    """The oracle's nrelax adaptation after one cycle (poisson.h:198-204)."""
    return _lib(1, False, plant).orc_nrelax_update(resb, resa, tolerance, nrelax)


def set_threads(n, dim=3):
    _lib(dim, True).orc_set_threads(n)


# --- embedded boundaries (embed.h), 2D -----------------------------------

EMBED_DIRICHLET, EMBED_NEUMANN = 1, 2


def fractions(phi, val=0.0):
    """fractions() (fractions.h:129-290) in 2D from the vertex level set
    phi ((n+1, n+1), vertex (i, j) the lower-left corner of cell (i, j)):
    returns cs (n, n), fs_x (n+1, n) and fs_y (n, n+1)."""
    lib = _lib(2)
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    n = phi.shape[0] - 1
    cs = np.zeros((n, n))
    fx = np.zeros((n + 1, n))
    fy = np.zeros((n, n + 1))
    lib.orc_fractions(n, _dptr(phi), float(val), _dptr(cs), _dptr(fx), _dptr(fy))
    return cs, fx, fy


def embed_geometry(cs, fx, fy):
    """Per cut cell (0 < cs < 1): the fragment's centroid (px, py, relative
    to the cell centre in cell units, as embed_area_center shifts x and y),
    its length and its Euclidean unit normal (nx, ny); zero elsewhere."""
    lib = _lib(2)
    n = cs.shape[0]
    out = [np.zeros((n, n)) for _ in range(5)]
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in (cs, fx, fy)]
    lib.orc_embed_geometry(n, *[_dptr(a) for a in args], *[_dptr(o) for o in out])
    return tuple(out)


def cut_centroids(cs, fx, fy):
    """Per cut cell, the centroid of its fluid part (line_center), relative
    to the cell centre in cell units."""
    lib = _lib(2)
    n = cs.shape[0]
    px, py = np.zeros((n, n)), np.zeros((n, n))
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in (cs, fx, fy)]
    lib.orc_cut_centroids(n, *[_dptr(a) for a in args], _dptr(px), _dptr(py))
    return px, py
