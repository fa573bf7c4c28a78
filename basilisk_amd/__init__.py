"""basilisk_amd -- MI355X multigrid Poisson-Helmholtz solver for Basilisk.

The solver runs in ``libbmg.so`` (hand-written HIP for gfx950, built in-tree
by ``python -m basilisk_amd.build``); this module is its ctypes front end.
``Grid`` mirrors the reference's programming model: fields a, b, res, da,
lambda and the face vector alpha on a uniform multigrid hierarchy, and the
entry points ``poisson()`` / ``mg_solve()`` / ``mg_cycle()`` / ``residual()``
of poisson.h (poisson.h:32-441), with boundary conditions given like
``a[left] = dirichlet(...)`` (common.h:957-960).

There is no CPU fallback: importing the package without the built library,
or creating a Grid without a GPU, raises.
"""
import ctypes
import os

import numpy as np

__all__ = ["Grid", "Tree", "MGStats", "BMGError", "library", "LIB_PATH",
           "BC_SYMMETRY", "BC_DIRICHLET", "BC_NEUMANN", "BC_PERIODIC",
           "RIGHT", "LEFT", "TOP", "BOTTOM", "FRONT", "BACK"]

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libbmg.so")

BC_SYMMETRY, BC_DIRICHLET, BC_NEUMANN, BC_PERIODIC, BC_AFFINE = range(5)
RIGHT, LEFT, TOP, BOTTOM, FRONT, BACK = range(6)
SMOOTHERS = {"gs": 0, "jacobi": 1, "rbgs": 2}
PRECISIONS = {"fp64": 0, "mixed": 1}
F_A, F_B, F_RES, F_DA, F_LAMBDA, F_ALPHA = range(6)
F_S = 8  # trees: user scalar k is F_S + k, k < USER_SCALARS (adapt_wavelet's scalars)
USER_SCALARS = 4
PROFILE_KINDS = {"relax_fine": 0, "residual": 1, "restriction": 2,
                 "prolongation": 3, "correction": 4, "relax_coarse": 5, "relax_fine2": 6}

# every symbol declared in include/bmg.h
EXPORTS = [
    "bmg_version", "bmg_last_error", "bmg_device_count", "bmg_grid_new",
    "bmg_grid_free", "bmg_grid_info", "bmg_set_grid_kind", "bmg_set_precision", "bmg_set_smoother",
    "bmg_set_tolerance", "bmg_set_niter", "bmg_set_stream", "bmg_set_bc",
    "bmg_set_bc_affine", "bmg_set_field", "bmg_get_field",
    "bmg_set_alpha_const", "bmg_set_lambda_const",
    "bmg_restriction_alpha_lambda", "bmg_residual", "bmg_relax",
    "bmg_mg_cycle", "bmg_mg_solve", "bmg_poisson", "bmg_synchronize",
    "bmg_profile_enable", "bmg_profile_get", "bmg_poisson_view",
    "bmg_tree_new", "bmg_tree_free", "bmg_tree_info", "bmg_tree_leaves",
    "bmg_tree_refine", "bmg_tree_cache", "bmg_tree_faces", "bmg_tree_set_precision",
    "bmg_tree_set_smoother", "bmg_tree_set_tolerance", "bmg_tree_set_niter",
    "bmg_tree_set_bc", "bmg_tree_set_leaf_field", "bmg_tree_get_leaf_field",
    "bmg_tree_set_alpha_faces", "bmg_tree_set_alpha_const", "bmg_tree_set_lambda_const",
    "bmg_tree_restriction_alpha_lambda", "bmg_tree_residual", "bmg_tree_relax",
    "bmg_tree_mg_cycle", "bmg_tree_mg_solve", "bmg_tree_poisson", "bmg_tree_synchronize",
    "bmg_tree_adapt_wavelet", "bmg_tree_adapt_wavelet_list", "bmg_tree_dump", "bmg_tree_restore", "bmg_tree_nonleaf", "bmg_tree_unrefine",
    "bmg_coef_gradient", "bmg_project", "bmg_viscosity", "bmg_viscosity_explicit", "bmg_set_axi", "bmg_grid_dump", "bmg_grid_restore",
    "bmg_tree_profile_enable", "bmg_tree_profile_get", "bmg_view_configure", "bmg_view_release",
    "bmg_poisson_noalloc", "bmg_centered_gradient", "bmg_correction", "bmg_tree_project",
    "bmg_tree_residual_transpose", "bmg_tree_cycle_transpose", "bmg_tree_adjoint",
    "bmg_set_embed", "bmg_embed_geometry",
]


EMBED_NONE, EMBED_DIRICHLET, EMBED_NEUMANN = 0, 1, 2


def embed_geometry(cs, fs_x, fs_y):
    """Per cut cell of an n x n level: the fragment's centroid (px, py,
    relative to the cell centre in cell sizes), its length and its unit
    normal (nx, ny), as embed_geometry (embed.h:221-230) gives them; zero
    elsewhere."""
    lib = library()
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (cs, fs_x, fs_y)]
    n = arrs[0].shape[0]
    out = [np.zeros((n, n)) for _ in range(5)]
    _check(lib.bmg_embed_geometry(n, *[_host_ptr(a) for a in arrs], *[_host_ptr(o) for o in out]))
    return tuple(out)


class BMGError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("libbmg error %d: %s" % (code, msg))
        self.code = code


class MGStats(ctypes.Structure):
    """mgstats, poisson.h:108-114."""
    _fields_ = [("i", ctypes.c_int), ("resb", ctypes.c_double),
                ("resa", ctypes.c_double), ("sum", ctypes.c_double),
                ("nrelax", ctypes.c_int), ("minlevel", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MultigridView(ctypes.Structure):
    """bmg_multigrid_view: the reference's multigrid storage
    (grid/multigrid.h:16-73), one array of (2^l+4)^dim records per level."""
    _fields_ = [("dimension", ctypes.c_int), ("depth", ctypes.c_int),
                ("d", ctypes.POINTER(ctypes.c_void_p)), ("datasize", ctypes.c_size_t),
                ("nvarmax", ctypes.c_int), ("constant", ctypes.POINTER(ctypes.c_double)),
                ("L0", ctypes.c_double), ("X0", ctypes.c_double), ("Y0", ctypes.c_double),
                ("Z0", ctypes.c_double)]


class PoissonArgs(ctypes.Structure):
    """bmg_poisson_args: struct Poisson (poisson.h:258-265) with handles."""
    _fields_ = [("a", ctypes.c_int), ("b", ctypes.c_int), ("alpha", ctypes.c_int * 3),
                ("lambda_", ctypes.c_int), ("tolerance", ctypes.c_double),
                ("nrelax", ctypes.c_int), ("minlevel", ctypes.c_int), ("res", ctypes.c_int)]


class ViewBC(ctypes.Structure):
    """bmg_view_bc: per direction, kind, sigma and the face values c."""
    _fields_ = [("kind", ctypes.c_int * 6), ("sigma", ctypes.c_int * 6),
                ("c", ctypes.POINTER(ctypes.c_double) * 6)]


def view_records(dim, depth, nvar):
    """Zeroed record arrays laid out as the reference's multigrid storage:
    level l is ((2^l+4)^dim, nvar) doubles (GHOSTS = 2, datasize 8 nvar)."""
    return [np.zeros((((1 << l) + 4) ** dim, nvar)) for l in range(depth + 1)]


def poisson_view(levels, dim, depth, nvar, a, b, alpha, lam, tolerance, bc_sigma=None, TOLERANCE=1e-3,
                 NITERMAX=100, NITERMIN=1, res=-1):
    """bmg_poisson_view on record arrays from view_records(): field handles
    a, b, alpha (dim handles or 0), lam (handle or 0), res (-1: none);
    walls are affine with c = 0 and sigma bc_sigma[side] (default +1,
    symmetry).  Returns the mgstats dict."""
    lib = library()
    ptrs = (ctypes.c_void_p * (depth + 1))(*[lv.ctypes.data for lv in levels])
    const = np.zeros(4)
    v = MultigridView(dim, depth, ptrs, nvar * 8, 1 << 16, const.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                      1.0, 0.0, 0.0, 0.0)
    al = list(alpha) + [0] * (3 - len(alpha))
    args = PoissonArgs(a, b, (ctypes.c_int * 3)(*al), lam, tolerance, 0, 0, res)
    bc = ViewBC()
    for side in range(2 * dim):
        bc.kind[side] = BC_AFFINE
        bc.sigma[side] = 1 if bc_sigma is None else bc_sigma[side]
    st = MGStats()
    _check(lib.bmg_poisson_view(ctypes.byref(v), ctypes.byref(args), ctypes.byref(bc), TOLERANCE, NITERMAX,
                                NITERMIN, ctypes.byref(st)), allow_nan=True)
    return st.as_dict()


def view_configure(precision="fp64", smoother="rbgs"):
    """Precision and smoother of the following poisson_view calls."""
    _check(library().bmg_view_configure(PRECISIONS[precision], SMOOTHERS[smoother]))


_lib = None


def library():
    """Load libbmg.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libbmg.so is not built: run `python -m basilisk_amd.build` "
                          "(there is no CPU fallback)")
    # PyTorch, when installed, is loaded first: it carries its own copy of
    # the HIP runtime, and a process in which this library has already
    # initialised the GPU through the system runtime finds no device for
    # torch afterwards.  Loaded first, torch's runtime serves both.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    dp = ctypes.c_void_p
    sigs = {
        "bmg_version": (ctypes.c_char_p, []),
        "bmg_last_error": (ctypes.c_char_p, []),
        "bmg_device_count": (I, [ctypes.POINTER(I)]),
        "bmg_grid_new": (I, [ctypes.POINTER(P), I, I, D, D, D, D]),
        "bmg_grid_free": (I, [P]),
        "bmg_grid_info": (I, [P, ctypes.POINTER(I), ctypes.POINTER(I)]),
        "bmg_set_grid_kind": (I, [P, I]),
        "bmg_set_precision": (I, [P, I]),
        "bmg_set_smoother": (I, [P, I]),
        "bmg_set_tolerance": (I, [P, D]),
        "bmg_set_niter": (I, [P, I, I]),
        "bmg_set_stream": (I, [P, P]),
        "bmg_set_bc": (I, [P, I, I, dp]),
        "bmg_set_bc_affine": (I, [P, I, I, dp]),
        "bmg_set_embed": (I, [P, P, P, P, P, I, P, I]),
        "bmg_embed_geometry": (I, [I, P, P, P, P, P, P, P, P]),
        "bmg_set_field": (I, [P, I, I, dp, I]),
        "bmg_get_field": (I, [P, I, I, dp, I]),
        "bmg_set_alpha_const": (I, [P, D]),
        "bmg_set_lambda_const": (I, [P, D]),
        "bmg_restriction_alpha_lambda": (I, [P]),
        "bmg_residual": (I, [P, ctypes.POINTER(D)]),
        "bmg_relax": (I, [P, I]),
        "bmg_mg_cycle": (I, [P, I, I, I]),
        "bmg_mg_solve": (I, [P, D, I, I, ctypes.POINTER(MGStats)]),
        "bmg_poisson": (I, [P, D, I, I, ctypes.POINTER(MGStats)]),
        "bmg_synchronize": (I, [P]),
        "bmg_profile_enable": (I, [P, I]),
        "bmg_profile_get": (I, [P, I, ctypes.POINTER(D), ctypes.POINTER(ctypes.c_long)]),
        "bmg_poisson_view": (I, [P, P, P, D, I, I, ctypes.POINTER(MGStats)]),
        "bmg_view_configure": (I, [I, I]),
        "bmg_centered_gradient": (I, [P, P, P, P, P, P, P, I]),
        "bmg_tree_project": (I, [P, P, D, I, I, ctypes.POINTER(MGStats)]),
        "bmg_tree_residual_transpose": (I, [P, P, P, I]),
        "bmg_tree_cycle_transpose": (I, [P, P, P, I, I, I]),
        "bmg_tree_adjoint": (I, [P, P, D, I, I, P, P, P, I, ctypes.POINTER(MGStats)]),
        "bmg_correction": (I, [P, P, P, P, P, P, P, D, I]),
        "bmg_view_release": (I, []),
        "bmg_tree_new": (I, [ctypes.POINTER(P), I, D, D, D, D]),
        "bmg_tree_free": (I, [P]),
        "bmg_tree_info": (I, [P, ctypes.POINTER(I), ctypes.POINTER(ctypes.c_long),
                              ctypes.POINTER(ctypes.c_long)]),
        "bmg_tree_leaves": (I, [P, dp, dp]),
        "bmg_tree_refine": (I, [P, dp, ctypes.POINTER(ctypes.c_long)]),
        "bmg_tree_cache": (I, [P, I, I, ctypes.POINTER(I), dp]),
        "bmg_tree_faces": (I, [P, ctypes.POINTER(ctypes.c_long), dp]),
        "bmg_tree_set_precision": (I, [P, I]),
        "bmg_tree_set_smoother": (I, [P, I]),
        "bmg_tree_set_tolerance": (I, [P, D]),
        "bmg_tree_set_niter": (I, [P, I, I]),
        "bmg_tree_set_bc": (I, [P, I, I, dp]),
        "bmg_tree_set_leaf_field": (I, [P, I, dp, I]),
        "bmg_tree_get_leaf_field": (I, [P, I, dp, I]),
        "bmg_tree_set_alpha_faces": (I, [P, dp, I]),
        "bmg_tree_set_alpha_const": (I, [P, D]),
        "bmg_tree_set_lambda_const": (I, [P, D]),
        "bmg_tree_restriction_alpha_lambda": (I, [P]),
        "bmg_tree_residual": (I, [P, ctypes.POINTER(D)]),
        "bmg_tree_relax": (I, [P, I]),
        "bmg_tree_mg_cycle": (I, [P, I, I, I]),
        "bmg_tree_mg_solve": (I, [P, D, I, I, ctypes.POINTER(MGStats)]),
        "bmg_tree_poisson": (I, [P, D, I, I, ctypes.POINTER(MGStats)]),
        "bmg_tree_synchronize": (I, [P]),
        "bmg_tree_profile_enable": (I, [P, I]),
        "bmg_tree_profile_get": (I, [P, ctypes.POINTER(D), ctypes.POINTER(ctypes.c_long),
                                     ctypes.POINTER(ctypes.c_long)]),
        "bmg_tree_adapt_wavelet": (I, [P, D, I, I, ctypes.POINTER(I), ctypes.POINTER(I)]),
        "bmg_tree_adapt_wavelet_list": (I, [P, I, P, P, I, I, ctypes.POINTER(I), ctypes.POINTER(I)]),
        "bmg_tree_dump": (I, [P, ctypes.c_char_p, dp, dp, I, D, I]),
        "bmg_tree_nonleaf": (I, [P, ctypes.POINTER(ctypes.c_long), dp, dp]),
        "bmg_tree_unrefine": (I, [P, dp, ctypes.POINTER(ctypes.c_long)]),
        "bmg_tree_restore": (I, [P, ctypes.c_char_p, dp, dp, I, ctypes.POINTER(D), ctypes.POINTER(I)]),
        "bmg_coef_gradient": (I, [P, dp, dp, dp, dp, dp, I]),
        "bmg_project": (I, [P, dp, dp, dp, D, I, I, ctypes.POINTER(MGStats)]),
        "bmg_viscosity": (I, [P, dp, dp, dp, dp, dp, dp, D, dp, D, D, dp, I, I, ctypes.POINTER(MGStats)]),
        "bmg_set_axi": (I, [P, I]),
        "bmg_viscosity_explicit": (I, [P, dp, dp, dp, dp, dp, dp, D, dp, D, D, dp, I, ctypes.POINTER(MGStats)]),
        "bmg_grid_dump": (I, [P, ctypes.c_char_p, dp, dp, I, D, I]),
        "bmg_grid_restore": (I, [P, ctypes.c_char_p, dp, dp, I, ctypes.POINTER(D), ctypes.POINTER(I)]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(code, allow_nan=False):
    if code == 0 or (allow_nan and code == 5):
        return code
    raise BMGError(code, library().bmg_last_error().decode())


def _host_ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _device_ptr(t, numel, what):
    """Pointer of a contiguous float64 CUDA tensor of `numel` elements; any
    other object (a CPU tensor, another dtype, a strided view, the wrong
    size) raises ValueError before the library could write through it."""
    if not (hasattr(t, "data_ptr") and getattr(t, "is_cuda", False)):
        raise ValueError("%s must be a CUDA tensor" % what)
    if str(t.dtype) != "torch.float64" or not t.is_contiguous() or t.numel() != numel:
        raise ValueError("%s must be a contiguous float64 tensor of %d elements (got %s, %s, %d)"
                         % (what, numel, t.dtype, "contiguous" if t.is_contiguous() else "strided", t.numel()))
    return ctypes.c_void_p(t.data_ptr())


class Grid:
    """A uniform multigrid hierarchy (levels 0..depth) resident on the GPU.

    ``dim``: 1, 2 or 3; ``depth``: finest level (2^depth cells per
    dimension); ``L0``/``origin``: the box, like size() and origin()
    (common.h:857-866).  ``precision``: "fp64" or "mixed";
    ``smoother``: "rbgs" (default), "jacobi" (the reference's -DJACOBI
    relax) or "gs" (the reference's default in-place Gauss-Seidel, i-major
    order, run by hyperplanes; multigrid grids only)."""

    def __init__(self, dim, depth, L0=1.0, origin=(0.0, 0.0, 0.0),
                 precision="fp64", smoother="rbgs", kind="multigrid"):
        lib = library()
        self.dim, self.depth, self.L0 = dim, depth, float(L0)
        self.origin = tuple(float(o) for o in origin) + (0.0,) * (3 - len(origin))
        self._g = ctypes.c_void_p()
        _check(lib.bmg_grid_new(ctypes.byref(self._g), dim, depth, self.L0, *self.origin[:3]))
        self.lib = lib
        self.set_smoother(smoother)
        self.set_precision(precision)
        self.set_kind(kind)

    def set_kind(self, kind):
        """"multigrid" or "tree" (a full quadtree/octree): the ghost rule of
        edge and corner cells."""
        _check(self.lib.bmg_set_grid_kind(self._g, {"multigrid": 0, "tree": 1}[kind]))
        self.kind = kind

    def close(self):
        if getattr(self, "_g", None):
            self.lib.bmg_grid_free(self._g)
            self._g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- geometry (cartesian-common.h:9-34)
    def n(self, level=None):
        return 1 << (self.depth if level is None else level)

    def delta(self, level=None):
        return self.L0 * (1.0 / self.n(level))

    def centers(self, level=None):
        n, D = self.n(level), self.delta(level)
        return [(np.arange(n) + 0.5) * D + self.origin[d] for d in range(self.dim)]

    def shape(self, level=None):
        return (self.n(level),) * self.dim

    def face_shape(self, c, level=None):
        s = list(self.shape(level))
        s[c] += 1
        return tuple(s)

    # -- configuration
    def set_precision(self, name):
        _check(self.lib.bmg_set_precision(self._g, PRECISIONS[name]))
        self.precision = name

    def set_smoother(self, name):
        _check(self.lib.bmg_set_smoother(self._g, SMOOTHERS[name]))
        self.smoother = name

    def set_tolerance(self, tol):
        _check(self.lib.bmg_set_tolerance(self._g, float(tol)))

    def set_niter(self, nitermax=100, nitermin=1):
        _check(self.lib.bmg_set_niter(self._g, nitermax, nitermin))

    def set_stream(self, stream_ptr):
        _check(self.lib.bmg_set_stream(self._g, ctypes.c_void_p(stream_ptr)))

    # -- fields (numpy arrays on the host, or torch tensors on the GPU)
    def _put(self, field, arr, level, shape):
        level = self.depth if level is None else level
        if hasattr(arr, "data_ptr") and getattr(arr, "is_cuda", False):
            if tuple(arr.shape) != shape or not arr.is_contiguous() or str(arr.dtype) != "torch.float64":
                raise ValueError("device field must be a contiguous float64 tensor of shape %s" % (shape,))
            import torch
            torch.cuda.current_stream().synchronize()
            _check(self.lib.bmg_set_field(self._g, field, level, ctypes.c_void_p(arr.data_ptr()), 1))
            return
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.size != int(np.prod(shape)):
            raise ValueError("field shape %s does not match %s" % (a.shape, shape))
        _check(self.lib.bmg_set_field(self._g, field, level, _host_ptr(a), 0))

    def _get(self, field, level, shape, out=None):
        level = self.depth if level is None else level
        if out is not None:
            ptr = _device_ptr(out, int(np.prod(shape)), "out")
            _check(self.lib.bmg_get_field(self._g, field, level, ptr, 1))
            return out
        a = np.empty(shape)
        _check(self.lib.bmg_get_field(self._g, field, level, _host_ptr(a), 0))
        return a

    def set_a(self, arr):
        self._put(F_A, arr, None, self.shape())

    def set_b(self, arr):
        self._put(F_B, arr, None, self.shape())

    def get_a(self, out=None):
        return self._get(F_A, None, self.shape(), out)

    def get_b(self, out=None):
        return self._get(F_B, None, self.shape(), out)

    def get_res(self, level=None):
        return self._get(F_RES, level, self.shape(level))

    def set_res(self, arr, level=None):
        self._put(F_RES, arr, level, self.shape(level))

    def get_da(self, level=None):
        return self._get(F_DA, level, self.shape(level))

    def set_da(self, arr, level=None):
        self._put(F_DA, arr, level, self.shape(level))

    def set_alpha(self, comps, level=None):
        if np.isscalar(comps):
            _check(self.lib.bmg_set_alpha_const(self._g, float(comps)))
            return
        for c, arr in enumerate(comps):
            self._put(F_ALPHA + c, arr, level, self.face_shape(c, level))

    def get_alpha(self, c, level=None):
        return self._get(F_ALPHA + c, level, self.face_shape(c, level))

    def set_lambda(self, lam, level=None):
        if np.isscalar(lam):
            _check(self.lib.bmg_set_lambda_const(self._g, float(lam)))
        else:
            self._put(F_LAMBDA, lam, level, self.shape(level))

    def get_lambda(self, level=None):
        return self._get(F_LAMBDA, level, self.shape(level))

    # -- boundary conditions of `a`
    def set_bc(self, direction, kind, values=None):
        """values: the BC expression on the n^(dim-1) faces of that side of
        the finest level (C order over the tangential indices), or None."""
        if values is None:
            _check(self.lib.bmg_set_bc(self._g, direction, kind, None))
            return
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        if v.size == (self.n() + 4) ** (self.dim - 1) and self.dim > 1:
            # a full-extent (ghost-inclusive) array: keep the faces of cells
            m = self.n() + 4
            v = v.reshape((m,) * (self.dim - 1))[(slice(2, -2),) * (self.dim - 1)].reshape(-1)
        if v.size != self.n() ** (self.dim - 1):
            raise ValueError("expected %d boundary values" % self.n() ** (self.dim - 1))
        _check(self.lib.bmg_set_bc(self._g, direction, kind, _host_ptr(v)))

    def set_bc_affine(self, direction, sigma, c):
        """ghost = c + sigma s on that side (sigma +1, -1, or 0: the ghost
        is c itself); c: n^(dim-1) values of the finest level."""
        v = np.ascontiguousarray(c, dtype=np.float64).reshape(-1)
        _check(self.lib.bmg_set_bc_affine(self._g, direction, int(sigma), _host_ptr(v)))

    def set_embed(self, cs, fs_x, fs_y, kind, bc=None, third=False):
        """An embedded boundary (embed.h, 2D): the finest level's volume
        fractions cs (n, n) and face fractions fs_x (n+1, n), fs_y (n, n+1);
        kind EMBED_DIRICHLET or EMBED_NEUMANN with the condition's value in
        each cut cell at the fragment's centroid (None: 0); third: embedded
        face gradients in the residual.  kind EMBED_NONE removes it."""
        if kind == EMBED_NONE:
            _check(self.lib.bmg_set_embed(self._g, None, None, None, None, 0, None, 0))
            return
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (cs, fs_x, fs_y)]
        n = self.n()
        for a, shp in zip(arrs, ((n, n), (n + 1, n), (n, n + 1))):
            if a.shape != shp:
                raise ValueError("embed fractions: shape %s, expected %s" % (a.shape, shp))
        b = None if bc is None else np.ascontiguousarray(bc, dtype=np.float64).reshape(n, n)
        _check(self.lib.bmg_set_embed(self._g, *[_host_ptr(a) for a in arrs], None, int(kind),
                                      None if b is None else _host_ptr(b), 1 if third else 0))

    def set_bc_function(self, kind, fn, directions=None):
        """Evaluate fn(x, y, z) at the centre of each boundary face (the
        coordinates a qcc-generated BC function sees, qcc.lex:1118-1130)."""
        dirs = range(2 * self.dim) if directions is None else directions
        n, D = self.n(), self.delta()
        for b in dirs:
            d, side = b // 2, (1 if b % 2 == 0 else -1)
            plane = n - 1 if side > 0 else 0
            coords = []
            for e in range(3):
                if e >= self.dim:
                    coords.append(np.zeros(1))
                elif e == d:
                    coords.append(np.array([(side / 2.0 + plane + 0.5) * D + self.origin[e]]))
                else:
                    coords.append((np.arange(n) + 0.5) * D + self.origin[e])
            X, Y, Z = np.meshgrid(*coords, indexing="ij")
            vals = np.asarray(fn(X, Y, Z), dtype=np.float64) * np.ones_like(X)
            self.set_bc(b, kind, vals.reshape(-1))

    def set_periodic(self, d):
        _check(self.lib.bmg_set_bc(self._g, 2 * d, BC_PERIODIC, None))

    # -- solver (poisson.h)
    def restriction_alpha_lambda(self):
        _check(self.lib.bmg_restriction_alpha_lambda(self._g))

    def residual(self):
        m = ctypes.c_double()
        _check(self.lib.bmg_residual(self._g, ctypes.byref(m)))
        return m.value

    def relax(self, level):
        _check(self.lib.bmg_relax(self._g, level))

    def mg_cycle(self, nrelax, minlevel, maxlevel=None):
        _check(self.lib.bmg_mg_cycle(self._g, nrelax, minlevel,
                                     self.depth if maxlevel is None else maxlevel))

    def mg_solve(self, tolerance=0.0, nrelax=0, minlevel=0):
        s = MGStats()
        _check(self.lib.bmg_mg_solve(self._g, tolerance, nrelax, minlevel, ctypes.byref(s)))
        return s.as_dict()

    def poisson(self, tolerance=0.0, nrelax=0, minlevel=0):
        s = MGStats()
        _check(self.lib.bmg_poisson(self._g, tolerance, nrelax, minlevel, ctypes.byref(s)))
        return s.as_dict()

    def synchronize(self):
        _check(self.lib.bmg_synchronize(self._g))

    def coef_gradient(self, psi, out_alpha=None, out_lambda=None, want_lambda=True):
        """dJ/dalpha (per face component) and dJ/dlambda for the adjoint state
        psi, at the grid's current `a` (see bmg_coef_gradient).  With torch
        CUDA tensors (psi and preallocated outputs) everything stays on the
        device; with numpy input new host arrays are returned."""
        if hasattr(psi, "data_ptr") and getattr(psi, "is_cuda", False):
            ncell = int(np.prod(self.shape()))
            pp = _device_ptr(psi, ncell, "psi")
            ptrs = [None, None, None]
            for c in range(self.dim):
                if out_alpha is not None and out_alpha[c] is not None:
                    ptrs[c] = _device_ptr(out_alpha[c], int(np.prod(self.face_shape(c))), "out_alpha[%d]" % c)
            lam = _device_ptr(out_lambda, ncell, "out_lambda") if out_lambda is not None else None
            import torch
            torch.cuda.current_stream().synchronize()
            _check(self.lib.bmg_coef_gradient(self._g, pp, ptrs[0], ptrs[1], ptrs[2], lam, 1))
            return out_alpha, out_lambda
        p = np.ascontiguousarray(psi, dtype=np.float64)
        if p.size != int(np.prod(self.shape())):
            raise ValueError("psi must have one value per finest-level cell")
        ga = [np.empty(self.face_shape(c)) for c in range(self.dim)]
        gl = np.empty(self.shape()) if want_lambda else None
        ptrs = [_host_ptr(x) for x in ga] + [None] * (3 - self.dim)
        _check(self.lib.bmg_coef_gradient(self._g, _host_ptr(p), ptrs[0], ptrs[1], ptrs[2],
                                          _host_ptr(gl) if gl is not None else None, 0))
        return ga, gl

    def project(self, uf, dt=1.0, nrelax=0):
        """project(), poisson.h:469-509: makes the face velocity `uf` (one
        array per component, n+1 faces along its own direction) discretely
        divergence-free with p = the grid's `a` (initial guess in, pressure
        out) and the grid's alpha; returns the poisson() statistics.  torch
        CUDA tensors are updated on the device, numpy float64 arrays in
        place through host copies."""
        if len(uf) != self.dim:
            raise ValueError("need %d face components" % self.dim)
        for c, u in enumerate(uf):
            if tuple(u.shape) != self.face_shape(c):
                raise ValueError("face component %d: shape %s, expected %s" % (c, tuple(u.shape), self.face_shape(c)))
        s = MGStats()
        if all(getattr(u, "is_cuda", False) for u in uf):
            import torch
            for u in uf:
                if u.dtype != torch.float64 or not u.is_contiguous():
                    raise ValueError("face components must be contiguous float64 tensors")
            torch.cuda.current_stream().synchronize()
            ptrs = [ctypes.c_void_p(u.data_ptr()) for u in uf] + [None] * (3 - self.dim)
            _check(self.lib.bmg_project(self._g, ptrs[0], ptrs[1], ptrs[2], float(dt), int(nrelax), 1,
                                        ctypes.byref(s)))
            return s.as_dict()
        for u in uf:
            if not isinstance(u, np.ndarray) or u.dtype != np.float64 or not u.flags.c_contiguous:
                raise ValueError("face components must be C-contiguous float64 numpy arrays (or CUDA tensors)")
        ptrs = [_host_ptr(u) for u in uf] + [None] * (3 - self.dim)
        _check(self.lib.bmg_project(self._g, ptrs[0], ptrs[1], ptrs[2], float(dt), int(nrelax), 0, ctypes.byref(s)))
        return s.as_dict()

    def _arrays(self, arrs, shapes, what):
        """(pointers, on_device) of a list of per-component arrays: all
        contiguous float64 CUDA tensors, or all C-contiguous float64 numpy
        arrays, with the given shapes."""
        for c, x in enumerate(arrs):
            if tuple(x.shape) != tuple(shapes[c]):
                raise ValueError("%s component %d: shape %s, expected %s" % (what, c, tuple(x.shape), shapes[c]))
        if all(getattr(x, "is_cuda", False) for x in arrs):
            import torch
            for x in arrs:
                if x.dtype != torch.float64 or not x.is_contiguous():
                    raise ValueError("%s must be contiguous float64 tensors" % what)
            torch.cuda.current_stream().synchronize()
            return [ctypes.c_void_p(x.data_ptr()) for x in arrs] + [None] * (3 - len(arrs)), 1
        for x in arrs:
            if not isinstance(x, np.ndarray) or x.dtype != np.float64 or not x.flags.c_contiguous:
                raise ValueError("%s must be C-contiguous float64 numpy arrays (or CUDA tensors)" % what)
        return [_host_ptr(x) for x in arrs] + [None] * (3 - len(arrs)), 0

    def centered_gradient(self, g, acc=None):
        """centered_gradient (p, g), navier-stokes/centered.h:387-406, with
        p = the grid's `a` and its BCs, the grid's alpha, fm = 1: writes the
        cell field g (dim arrays of shape (n,)*dim) from the face
        acceleration `acc` (dim face arrays, or None for zero)."""
        gp, dev = self._arrays(g, [self.shape()] * self.dim, "g")
        if acc is None:
            ap = [None] * 3
        else:
            ap, adev = self._arrays(acc, [self.face_shape(c) for c in range(self.dim)], "acc")
            if adev != dev:
                raise ValueError("acc and g must both be on the device or both on the host")
        _check(self.lib.bmg_centered_gradient(self._g, ap[0], ap[1], ap[2], gp[0], gp[1], gp[2], dev))

    def correction(self, u, g, dt):
        """correction (dt), navier-stokes/centered.h:322-327: u += dt g."""
        up, dev = self._arrays(u, [self.shape()] * self.dim, "u")
        gp, gdev = self._arrays(g, [self.shape()] * self.dim, "g")
        if gdev != dev:
            raise ValueError("u and g must both be on the device or both on the host")
        _check(self.lib.bmg_correction(self._g, up[0], up[1], up[2], gp[0], gp[1], gp[2], float(dt), dev))

    def set_axi(self, axi=True):
        """viscosity() in its AXI form (2D: x the axis, y the radius; the y
        component's lambda gains dt/rho (sum of the cell's four mu)/2/y^2).
        mu and rho are the caller's metric-weighted fields, as axi.h gives
        them (mu times the face's y, rho times the cell's y)."""
        _check(self.lib.bmg_set_axi(self._g, 1 if axi else 0))

    def viscosity(self, u, mu=1.0, rho=1.0, dt=1.0, sigma=None, nrelax=0, explicit=False):
        """viscosity(), viscosity.h:163-178: the implicit viscous step
        u - dt/rho div(2 mu D(u)) = u_old for the cell components `u` (one
        (n,)*dim array per component, updated in place), with relax_viscosity
        run by the grid's smoother (red-black by default, the reference's
        JACOBI variant with smoother="jacobi", its default in-place
        Gauss-Seidel with smoother="gs"); mu: one face array per component (n+1 along
        it) or a constant; rho: a cell array or a constant; sigma: (dim, 2
        dim) wall factors (default: -1 for the normal component, +1 for the
        tangential ones).  torch CUDA tensors stay on the device, numpy
        float64 arrays are updated through host copies.  Returns the
        mg_solve statistics.  explicit=True: viscosity_explicit()
        (viscosity.h:180-190), u += residual_viscosity (u, u), returning
        resb only."""
        if len(u) != self.dim:
            raise ValueError("need %d velocity components" % self.dim)
        shape = self.shape()
        for c, x in enumerate(u):
            if tuple(x.shape) != shape:
                raise ValueError("component %d: shape %s, expected %s" % (c, tuple(x.shape), shape))
        dev = all(getattr(x, "is_cuda", False) for x in u)
        keep = []

        def ptr(x, what):
            if x is None:
                return None
            if dev:
                import torch
                if not getattr(x, "is_cuda", False) or x.dtype != torch.float64 or not x.is_contiguous():
                    raise ValueError("%s: contiguous float64 CUDA tensor expected" % what)
                return ctypes.c_void_p(x.data_ptr())
            if not isinstance(x, np.ndarray) or x.dtype != np.float64 or not x.flags.c_contiguous:
                raise ValueError("%s: C-contiguous float64 numpy array expected" % what)
            keep.append(x)
            return _host_ptr(x)

        up = [ptr(x, "u") for x in u] + [None] * (3 - self.dim)
        if np.isscalar(mu):
            mp, mu_c = [None] * 3, float(mu)
        else:
            if len(mu) != self.dim:
                raise ValueError("need %d mu components" % self.dim)
            for c, m in enumerate(mu):
                if tuple(m.shape) != self.face_shape(c):
                    raise ValueError("mu component %d: shape %s, expected %s" % (c, tuple(m.shape),
                                                                                self.face_shape(c)))
            mp, mu_c = [ptr(m, "mu") for m in mu] + [None] * (3 - self.dim), 0.0
        if np.isscalar(rho):
            rp, rho_c = None, float(rho)
        else:
            if tuple(rho.shape) != shape:
                raise ValueError("rho: shape %s, expected %s" % (tuple(rho.shape), shape))
            rp, rho_c = ptr(rho, "rho"), 0.0
        sp = None
        if sigma is not None:
            sg = np.ones((3, 6), np.int32)
            sa = np.asarray(sigma, np.int32)
            sg[:sa.shape[0], :sa.shape[1]] = sa
            keep.append(sg)
            sp = _host_ptr(sg)
        if dev:
            import torch
            torch.cuda.current_stream().synchronize()
        s = MGStats()
        if explicit:
            _check(self.lib.bmg_viscosity_explicit(self._g, up[0], up[1], up[2], mp[0], mp[1], mp[2], mu_c, rp, rho_c,
                                                   float(dt), sp, 1 if dev else 0, ctypes.byref(s)))
        else:
            _check(self.lib.bmg_viscosity(self._g, up[0], up[1], up[2], mp[0], mp[1], mp[2], mu_c, rp, rho_c,
                                          float(dt), sp, int(nrelax), 1 if dev else 0, ctypes.byref(s)))
        return s.as_dict()

    @staticmethod
    def _dump_fields(names):
        fmap = {"a": F_A, "b": F_B}
        items = list(names.items())
        for k, _ in items:
            if k not in fmap:
                raise ValueError("dump fields are a and b (got %r)" % k)
        ids = (ctypes.c_int * max(1, len(items)))(*[fmap[k] for k, _ in items])
        nm = (ctypes.c_char_p * max(1, len(items)))(*[str(v).encode() for _, v in items])
        return ids, nm, len(items)

    def dump(self, path, names=None, t=0.0, i=0):
        """dump() (output.h:1043-1118) of every level in the reference's file
        format; `names` maps "a" and "b" to the names written."""
        ids, nm, n = self._dump_fields(names if names is not None else {"a": "a", "b": "b"})
        _check(self.lib.bmg_grid_dump(self._g, str(path).encode(), ctypes.cast(ids, ctypes.c_void_p),
                                      ctypes.cast(nm, ctypes.c_void_p), n, float(t), int(i)))

    def restore(self, path, names=None):
        """restore() (output.h:1221-1390) of the finest level; returns the
        header's (t, i)."""
        ids, nm, n = self._dump_fields(names if names is not None else {"a": "a", "b": "b"})
        tt, ii = ctypes.c_double(0.0), ctypes.c_int(0)
        _check(self.lib.bmg_grid_restore(self._g, str(path).encode(), ctypes.cast(ids, ctypes.c_void_p),
                                         ctypes.cast(nm, ctypes.c_void_p), n, ctypes.byref(tt), ctypes.byref(ii)))
        return tt.value, ii.value

    # -- kernel timing
    def profile(self, enable=True):
        _check(self.lib.bmg_profile_enable(self._g, int(enable)))

    def profile_get(self, kind):
        ms, n = ctypes.c_double(), ctypes.c_long()
        _check(self.lib.bmg_profile_get(self._g, PROFILE_KINDS[kind], ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value


GHOSTS = 2


class Tree:
    """An adaptive quadtree (dim 2) or octree (dim 3) over the box
    [origin, origin + L0]^dim, refined on the host like the reference's
    init_grid(1) + refine(cond) (grid/tree.h:1488-1584,
    grid/tree-common.h:21-67, 308-326), solved on the GPU.

    Leaf fields are numpy arrays in leaf (cache) order, or float64 torch CUDA
    tensors.  The solver entry points mirror poisson.h on a tree grid.
    ``precision``: "fp64" or "mixed"; ``smoother``: "rbgs" or "jacobi"."""

    def __init__(self, dim, L0=1.0, origin=(0.0, 0.0, 0.0), precision="fp64", smoother="rbgs"):
        lib = library()
        self.dim, self.L0 = dim, float(L0)
        self.origin = tuple(float(o) for o in origin) + (0.0,) * (3 - len(origin))
        self._t = ctypes.c_void_p()
        _check(lib.bmg_tree_new(ctypes.byref(self._t), dim, self.L0, *self.origin[:3]))
        self.lib = lib
        self.set_precision(precision)
        self.set_smoother(smoother)

    def close(self):
        if getattr(self, "_t", None):
            self.lib.bmg_tree_free(self._t)
            self._t = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- topology
    def info(self):
        d, nl, nc = ctypes.c_int(), ctypes.c_long(), ctypes.c_long()
        _check(self.lib.bmg_tree_info(self._t, ctypes.byref(d), ctypes.byref(nl), ctypes.byref(nc)))
        return d.value, nl.value, nc.value

    def depth(self):
        return self.info()[0]

    def nleaves(self):
        return self.info()[1]

    def leaves(self):
        n = self.nleaves()
        lev = np.zeros(n, np.int32)
        xyz = np.zeros((n, 3))
        _check(self.lib.bmg_tree_leaves(self._t, _host_ptr(lev), _host_ptr(xyz)))
        return lev, xyz

    def refine_marked(self, mark):
        m = np.ascontiguousarray(np.asarray(mark, dtype=np.uint8))
        if m.size != self.nleaves():
            raise ValueError("one mark per leaf")
        nr = ctypes.c_long()
        _check(self.lib.bmg_tree_refine(self._t, _host_ptr(m), ctypes.byref(nr)))
        return nr.value

    def refine(self, cond):
        """refine(cond): passes until no leaf satisfies
        ``cond(x, y, z, level) -> bool array`` (vectorised over the leaves);
        the user scalars that are set, a and b follow the mesh."""
        while True:
            lev, xyz = self.leaves()
            mark = np.asarray(cond(xyz[:, 0], xyz[:, 1], xyz[:, 2], lev), dtype=np.uint8)
            if not mark.any():
                return
            self.refine_marked(mark)

    def cache(self, which, level):
        w = {"active": 0, "prolongation": 1, "boundary": 2}[which]
        n = ctypes.c_int()
        _check(self.lib.bmg_tree_cache(self._t, w, level, ctypes.byref(n), None))
        out = np.zeros((n.value, 3), np.int32)
        if n.value:
            _check(self.lib.bmg_tree_cache(self._t, w, level, ctypes.byref(n), _host_ptr(out)))
        if self.dim == 2:
            out[:, 2] = 0
        return out

    def faces(self):
        n = ctypes.c_long()
        _check(self.lib.bmg_tree_faces(self._t, ctypes.byref(n), None))
        out = np.zeros((n.value, 5), np.int32)
        if n.value:
            _check(self.lib.bmg_tree_faces(self._t, ctypes.byref(n), _host_ptr(out)))
        return out

    def face_centres(self):
        """Centres of the face values in the order set_alpha_faces reads them
        (x = (ig/2. + _I + 0.5)*Delta + X0 with ig = -1 along the normal)."""
        f = self.faces()
        rows, ds = [], []
        for d in range(self.dim):
            sel = (f[:, 4] & (1 << d)) != 0
            rows.append(np.nonzero(sel)[0])
            ds.append(np.full(sel.sum(), d))
        order = np.argsort(np.concatenate(rows) * 4 + np.concatenate(ds), kind="stable")
        idx = np.concatenate(rows)[order]
        dd = np.concatenate(ds)[order]
        lev = f[idx, 0]
        D = self.L0 * (1.0 / (1 << lev.astype(np.int64)))
        xyz = np.zeros((len(idx), 3))
        for e in range(self.dim):
            g = np.where(dd == e, -1, 0)
            xyz[:, e] = ((g / 2.0 + (f[idx, 1 + e] - GHOSTS)) + 0.5) * D + self.origin[e]
        return xyz, dd.astype(np.int32)

    # -- settings
    def set_precision(self, p):
        _check(self.lib.bmg_tree_set_precision(self._t, PRECISIONS[p]))
        self.precision = p

    def set_smoother(self, s):
        _check(self.lib.bmg_tree_set_smoother(self._t, SMOOTHERS[s]))

    def set_tolerance(self, tol):
        _check(self.lib.bmg_tree_set_tolerance(self._t, tol))

    def set_niter(self, nmax, nmin=1):
        _check(self.lib.bmg_tree_set_niter(self._t, nmax, nmin))

    def bc_face_centres(self, side):
        """Centres of the boundary faces of `side`, every level 0..depth,
        concatenated in the order set_bc reads them."""
        d, hi = side // 2, side % 2 == 0
        out = []
        for lev in range(self.depth() + 1):
            n = 1 << lev
            D = self.L0 * (1.0 / n)
            tang = [e for e in range(self.dim) if e != d]
            grids = np.meshgrid(*[np.arange(n)] * (self.dim - 1), indexing="ij")
            c = np.zeros((n ** (self.dim - 1), 3))
            for t, e in enumerate(tang):
                c[:, e] = (grids[t].ravel() + 0.5) * D + self.origin[e]
            g = 1 if hi else -1
            c[:, d] = ((g / 2.0 + (n - 1 if hi else 0)) + 0.5) * D + self.origin[d]
            out.append(c)
        return np.concatenate(out)

    def set_bc(self, side, kind, values=None):
        v = None if values is None else np.ascontiguousarray(values, dtype=np.float64)
        _check(self.lib.bmg_tree_set_bc(self._t, side, kind, None if v is None else _host_ptr(v)))

    def set_bc_function(self, side, kind, f):
        c = self.bc_face_centres(side)
        self.set_bc(side, kind, f(c[:, 0], c[:, 1], c[:, 2]))

    # -- fields
    def _put(self, field, v):
        if hasattr(v, "data_ptr") and getattr(v, "is_cuda", False):
            if not v.is_contiguous() or str(v.dtype) != "torch.float64" or v.numel() != self.nleaves():
                raise ValueError("device leaf field must be a contiguous float64 tensor, one value per leaf")
            import torch
            torch.cuda.current_stream().synchronize()
            _check(self.lib.bmg_tree_set_leaf_field(self._t, field, ctypes.c_void_p(v.data_ptr()), 1))
            return
        a = np.ascontiguousarray(v, dtype=np.float64)
        if a.size != self.nleaves():
            raise ValueError("one value per leaf (%d), got %d" % (self.nleaves(), a.size))
        _check(self.lib.bmg_tree_set_leaf_field(self._t, field, _host_ptr(a), 0))

    def _get(self, field, out=None):
        if out is not None:
            ptr = _device_ptr(out, self.nleaves(), "out")
            _check(self.lib.bmg_tree_get_leaf_field(self._t, field, ptr, 1))
            return out
        a = np.empty(self.nleaves())
        _check(self.lib.bmg_tree_get_leaf_field(self._t, field, _host_ptr(a), 0))
        return a

    def set_a(self, v):
        self._put(F_A, v)

    def set_b(self, v):
        self._put(F_B, v)

    def get_a(self, out=None):
        return self._get(F_A, out)

    def get_b(self):
        return self._get(F_B)

    def get_res(self):
        return self._get(F_RES)

    def get_da(self):
        return self._get(F_DA)

    # -- adaptation (grid/tree-common.h)
    def set_s(self, v, k=0):
        """Leaf values of user scalar k (0..USER_SCALARS-1): a field that
        adapt_wavelet may estimate and that follows the mesh."""
        if not 0 <= k < USER_SCALARS:
            raise ValueError("user scalar %d outside 0..%d" % (k, USER_SCALARS - 1))
        self._put(F_S + k, v)

    def get_s(self, out=None, k=0):
        if not 0 <= k < USER_SCALARS:
            raise ValueError("user scalar %d outside 0..%d" % (k, USER_SCALARS - 1))
        return self._get(F_S + k, out)

    def nonleaf(self):
        """Levels and centres of the non-leaf cells foreach_cell() visits."""
        n = ctypes.c_long()
        _check(self.lib.bmg_tree_nonleaf(self._t, ctypes.byref(n), None, None))
        lev = np.zeros(n.value, np.int32)
        xyz = np.zeros((n.value, 3))
        if n.value:
            _check(self.lib.bmg_tree_nonleaf(self._t, ctypes.byref(n), _host_ptr(lev), _host_ptr(xyz)))
        return lev, xyz

    def unrefine(self, cond):
        """unrefine (cond) (grid/tree-common.h:348-371): cond(x, y, z, level)
        on the non-leaf cells; returns the number of coarsened cells."""
        lev, xyz = self.nonleaf()
        m = np.ascontiguousarray(cond(xyz[:, 0], xyz[:, 1], xyz[:, 2], lev), dtype=np.uint8)
        if m.size == 0:
            m = np.zeros(1, np.uint8)
        nc = ctypes.c_long()
        _check(self.lib.bmg_tree_unrefine(self._t, _host_ptr(m), ctypes.byref(nc)))
        return nc.value

    @staticmethod
    def _dump_fields(names):
        fmap = {"a": F_A, "b": F_B, "s": F_S}
        for k in range(1, USER_SCALARS):
            fmap["s%d" % k] = F_S + k
        items = list(names.items())
        for k, _ in items:
            if k not in fmap:
                raise ValueError("dump fields are a, b, s and s1..s%d (got %r)" % (USER_SCALARS - 1, k))
        ids = (ctypes.c_int * max(1, len(items)))(*[fmap[k] for k, _ in items])
        nm = (ctypes.c_char_p * max(1, len(items)))(*[str(v).encode() for _, v in items])
        return ids, nm, len(items)

    def dump(self, path, names=None, t=0.0, i=0):
        """dump() (output.h:1043-1118): the tree and the leaf fields in the
        reference's file format; `names` maps "a", "b", "s" (user scalar 0)
        and "s1".."s3" to the names written (default: a and b under their
        own names)."""
        ids, nm, n = self._dump_fields(names if names is not None else {"a": "a", "b": "b"})
        _check(self.lib.bmg_tree_dump(self._t, str(path).encode(), ctypes.cast(ids, ctypes.c_void_p),
                                      ctypes.cast(nm, ctypes.c_void_p), n, float(t), int(i)))

    def restore(self, path, names=None):
        """restore() (output.h:1221-1390): rebuilds the tree from the file and
        sets "a", "b", "s" from the columns named in `names`; returns the
        header's (t, i)."""
        ids, nm, n = self._dump_fields(names if names is not None else {"a": "a", "b": "b"})
        tt, ii = ctypes.c_double(0.0), ctypes.c_int(0)
        _check(self.lib.bmg_tree_restore(self._t, str(path).encode(), ctypes.cast(ids, ctypes.c_void_p),
                                         ctypes.cast(nm, ctypes.c_void_p), n, ctypes.byref(tt), ctypes.byref(ii)))
        return tt.value, ii.value

    def adapt_wavelet(self, max, maxlevel, minlevel=1, scalars=None):
        """adapt_wavelet (slist, max, maxlevel, minlevel)
        (grid/tree-common.h:165-303): slist is the user scalars `scalars`
        (default [0]) with one tolerance each in `max` (or one number for
        all); every user scalar that is set, a and b follow the mesh.
        Returns (nf, nc), the refined and coarsened cell counts.  alpha and
        a lambda field must be set again afterwards."""
        nf, nc = ctypes.c_int(0), ctypes.c_int(0)
        if scalars is None and np.isscalar(max):
            _check(self.lib.bmg_tree_adapt_wavelet(self._t, float(max), int(maxlevel), int(minlevel),
                                                   ctypes.byref(nf), ctypes.byref(nc)))
            return nf.value, nc.value
        us = [0] if scalars is None else list(scalars)
        mx = [float(max)] * len(us) if np.isscalar(max) else [float(m) for m in max]
        if len(mx) != len(us):
            raise ValueError("one tolerance per estimated scalar")
        ids = np.array([F_S + k for k in us], np.int32)
        ma = np.array(mx, np.float64)
        _check(self.lib.bmg_tree_adapt_wavelet_list(self._t, len(us), _host_ptr(ids), _host_ptr(ma), int(maxlevel),
                                                    int(minlevel), ctypes.byref(nf), ctypes.byref(nc)))
        return nf.value, nc.value

    def set_lambda(self, v):
        if np.isscalar(v):
            _check(self.lib.bmg_tree_set_lambda_const(self._t, float(v)))
        else:
            self._put(F_LAMBDA, v)

    def set_alpha_const(self, v):
        _check(self.lib.bmg_tree_set_alpha_const(self._t, float(v)))

    def set_alpha_faces(self, values):
        a = np.ascontiguousarray(values, dtype=np.float64)
        f = self.faces()
        n = int(sum(((f[:, 4] >> d) & 1).sum() for d in range(self.dim)))
        if a.size != n:
            raise ValueError("one alpha value per face value (%d), got %d" % (n, a.size))
        _check(self.lib.bmg_tree_set_alpha_faces(self._t, _host_ptr(a), 0))

    def nfacevalues(self):
        f = self.faces()
        return int(sum(((f[:, 4] >> d) & 1).sum() for d in range(self.dim)))

    def _leafvec(self, x, n, what):
        """(pointer, on_device, keep-alive) of a leaf-order or face-value
        vector: a contiguous float64 CUDA tensor or a host array"""
        if getattr(x, "is_cuda", False):
            return _device_ptr(x, n, what), 1, x
        v = np.ascontiguousarray(x, dtype=np.float64)
        if v.size != n:
            raise ValueError("%s: expected %d values, got %d" % (what, n, v.size))
        return _host_ptr(v), 0, v

    def residual_transpose(self, y):
        """r^T y for the leaf operator r(x) = -lambda x - div(alpha grad x)
        with homogeneous boundary conditions (leaf order; host arrays)."""
        n = self.nleaves()
        yp, _, keep = self._leafvec(y, n, "y")
        out = np.zeros(n)
        _check(self.lib.bmg_tree_residual_transpose(self._t, yp, _host_ptr(out), 0))
        return out

    def cycle_transpose(self, y, nrelax=4, minlevel=1):
        """the transpose of one V-cycle from da = 0 (red-black relax)
        applied to y (leaf order; host arrays)."""
        n = self.nleaves()
        yp, _, keep = self._leafvec(y, n, "y")
        out = np.zeros(n)
        _check(self.lib.bmg_tree_cycle_transpose(self._t, yp, _host_ptr(out), int(nrelax), int(minlevel), 0))
        return out

    def adjoint(self, g, tolerance=1e-10, nitermax=100, nrelax=4, alpha=True, lam=False):
        """The adjoint of the last poisson() solve: psi with L^T psi = g (L
        the leaf operator), dJ/db = psi, dJ/dalpha on the face values and
        dJ/dlambda at the leaves.  g: numpy (results numpy) or a CUDA tensor
        (results CUDA tensors).  Returns (psi, galpha or None, glambda or
        None, stats)."""
        n, nv = self.nleaves(), self.nfacevalues()
        gp, dev, keep = self._leafvec(g, n, "g")
        if dev:
            import torch
            mk = lambda m: torch.zeros(m, dtype=torch.float64, device=g.device)
            ptr = lambda x: ctypes.c_void_p(x.data_ptr())
            torch.cuda.current_stream().synchronize()
        else:
            mk = np.zeros
            ptr = _host_ptr
        psi = mk(n)
        ga = mk(nv) if alpha else None
        gl = mk(n) if lam else None
        s = MGStats()
        _check(self.lib.bmg_tree_adjoint(self._t, gp, float(tolerance), int(nitermax), int(nrelax), ptr(psi),
                                         ptr(ga) if alpha else None, ptr(gl) if lam else None, dev,
                                         ctypes.byref(s)))
        return psi, ga, gl, s.as_dict()

    def project(self, uf, dt=1.0, nrelax=0):
        """project(), poisson.h:469-509, on the tree with p = a: `uf` holds
        the face velocity on the face cache (one value per set flag bit, as
        set_alpha_faces) -- a float64 numpy array updated in place, or a
        contiguous float64 CUDA tensor updated on the device; returns the
        poisson() statistics."""
        n = self.nfacevalues()
        s = MGStats()
        if getattr(uf, "is_cuda", False):
            ptr = _device_ptr(uf, n, "uf")
            import torch
            torch.cuda.current_stream().synchronize()
            _check(self.lib.bmg_tree_project(self._t, ptr, float(dt), int(nrelax), 1, ctypes.byref(s)))
            return s.as_dict()
        if not isinstance(uf, np.ndarray) or uf.dtype != np.float64 or not uf.flags.c_contiguous or uf.size != n:
            raise ValueError("uf must be a C-contiguous float64 array of %d face values" % n)
        _check(self.lib.bmg_tree_project(self._t, _host_ptr(uf), float(dt), int(nrelax), 0, ctypes.byref(s)))
        return s.as_dict()

    def set_alpha_function(self, f):
        """alpha.x[] = f(x, y, z, d) on every face of the face cache."""
        xyz, dd = self.face_centres()
        self.set_alpha_faces(f(xyz[:, 0], xyz[:, 1], xyz[:, 2], dd))

    # -- solver (poisson.h)
    def restriction_alpha_lambda(self):
        _check(self.lib.bmg_tree_restriction_alpha_lambda(self._t))

    def residual(self):
        m = ctypes.c_double()
        _check(self.lib.bmg_tree_residual(self._t, ctypes.byref(m)), allow_nan=True)
        return m.value

    def relax(self, level):
        _check(self.lib.bmg_tree_relax(self._t, level))

    def mg_cycle(self, nrelax=4, minlevel=0, maxlevel=None):
        _check(self.lib.bmg_tree_mg_cycle(self._t, nrelax, minlevel,
                                          self.depth() if maxlevel is None else maxlevel))

    def mg_solve(self, tolerance=0.0, nrelax=0, minlevel=0):
        s = MGStats()
        _check(self.lib.bmg_tree_mg_solve(self._t, tolerance, nrelax, minlevel, ctypes.byref(s)))
        return s.as_dict()

    def poisson(self, tolerance=0.0, nrelax=0, minlevel=0):
        s = MGStats()
        _check(self.lib.bmg_tree_poisson(self._t, tolerance, nrelax, minlevel, ctypes.byref(s)))
        return s.as_dict()

    def synchronize(self):
        _check(self.lib.bmg_tree_synchronize(self._t))

    def profile(self, enable):
        """Time every relax launch of more than 2048 cells with HIP events
        (the cycle runs without its graph meanwhile); enabling resets."""
        _check(self.lib.bmg_tree_profile_enable(self._t, 1 if enable else 0))

    def profile_get(self):
        """(total ms, launches, cells relaxed) of the relax launches timed
        since profiling was enabled."""
        ms, n, cells = ctypes.c_double(), ctypes.c_long(), ctypes.c_long()
        _check(self.lib.bmg_tree_profile_get(self._t, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(cells)))
        return ms.value, n.value, cells.value
