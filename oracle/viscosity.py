"""CPU oracle of viscosity() on a uniform multigrid grid.  TEST
INFRASTRUCTURE ONLY: only tests/, the smoke test and the benchmark's CPU leg
may import this module; the product library never calls it.

A numpy restatement, in the reference's evaluation order (elementwise
operations only, so every value is rounded exactly as the C expressions
round it), of

  viscosity()            viscosity.h:163-178 (r = u, restriction ({mu, rho}),
                         mg_solve with minlevel 0 and TOLERANCE)
  relax_viscosity        viscosity.h:27-96, lambda = 1 (or, with AXI, the
                         y component's lambda = 1 + dt/rho (mu.x[] + mu.x[1]
                         + mu.y[] + mu.y[0,1])/2/y^2, viscosity.h:12-14, y the
                         cell centre (j + 1/2) Delta + Y0): the default
                         in-place Gauss-Seidel (a serial loop over the cells
                         in i-major order, each component overwritten as it
                         is computed; the ghosts keep the sweep's start
                         values), the JACOBI variant (w from the old
                         iterate, then u = (u + 2 w)/3), or red-black
                         Gauss-Seidel (an extension): the cells with i+j+k
                         even, then the odd ones, take w computed from the
                         state before their pass (the cross terms read
                         same-colour diagonal neighbours, which keep their
                         old values)
  residual_viscosity     viscosity.h:98-160, the "naive" (!TREE) form, with
                         the same lambda
  viscosity_explicit     viscosity.h:180-190: resb = residual_viscosity
                         with b = u, then u += res (no restriction)
  mg_solve, mg_cycle     poisson.h:32-90, 139-235 (nrelax default 4, the
                         nrelax adaptation, NITERMIN/NITERMAX)
  restriction            restriction_average and face_average,
                         grid/multigrid-common.h:26-63 (children summed in
                         foreach_child order; the faces in the tangential
                         order the dimension rotation gives each component)
  bilinear               grid/multigrid-common.h:124-140
  box ghosts             grid/multigrid.h:543-594: one ghost layer filled
                         direction by direction (right, left, top, bottom,
                         front, back) over the full tangential extent, so an
                         edge or corner ghost is the product of the sides'
                         factors times the diagonal interior cell; every
                         component's ghost is sigma * the mirrored cell
                         (antisymmetry -1 / symmetry +1: the reference's
                         vector BCs at walls at rest)

  periodic               grid/multigrid.h:600-701: a periodic dimension's
                         ghost layers are copies of the opposite interior
                         layer (filled in the same direction order, so an
                         edge ghost across a periodic and a wall direction is
                         the wall's factor times the wrapped cell); on level
                         0 every ghost is the root value (655-664)

Fields are (n,)*dim arrays in C order (i slowest); face field d has n+1
entries along d (on a periodic dimension face n is face 0: the value given
for it is replaced by face 0's).
"""
import numpy as np


def default_sigma(dim):
    """The reference's default vector BCs: the normal component is
    antisymmetric, the tangential ones symmetric
    (grid/cartesian-common.h:536-563)."""
    s = np.ones((dim, 2 * dim), np.int64)
    for c in range(dim):
        s[c, 2 * c] = s[c, 2 * c + 1] = -1
    return s


def _pad(f, sig, dim, periodic=None):
    """One ghost layer, filled direction by direction; a periodic dimension
    copies the opposite interior layer, and on level 0 of a grid with a
    periodic dimension every ghost is the root value."""
    n = f.shape[0]
    periodic = periodic or [False] * dim
    root = n == 1 and any(periodic)
    p = np.zeros((n + 2,) * dim)
    p[(slice(1, n + 1),) * dim] = f
    for d in range(dim):
        hi = [slice(None)] * dim
        lo = [slice(None)] * dim
        hi_src = [slice(None)] * dim
        lo_src = [slice(None)] * dim
        if periodic[d]:
            hi[d], hi_src[d] = n + 1, 1
            lo[d], lo_src[d] = 0, n
            p[tuple(hi)] = p[tuple(hi_src)]
            p[tuple(lo)] = p[tuple(lo_src)]
            continue
        hi[d], hi_src[d] = n + 1, n
        lo[d], lo_src[d] = 0, 1
        p[tuple(hi)] = (1.0 if root else float(sig[2 * d])) * p[tuple(hi_src)]
        p[tuple(lo)] = (1.0 if root else float(sig[2 * d + 1])) * p[tuple(lo_src)]
    return p


def _sh(p, off, n, dim):
    """View of the padded array p shifted by `off` (one entry per dim)."""
    return p[tuple(slice(1 + off[d], 1 + off[d] + n) for d in range(dim))]


def _face(mu, d, hi, n, dim):
    sl = [slice(None)] * dim
    sl[d] = slice(1, n + 1) if hi else slice(0, n)
    return mu[tuple(sl)]


def _unit(dim, *pairs):
    o = [0] * dim
    for d, v in pairs:
        o[d] += v
    return o


def _others(c, dim):
    return [(c + 1) % dim, (c + 2) % dim][:dim - 1]


class Viscosity:
    """The viscous problem on a 2^depth grid of the box [0, L0]^dim."""

    def __init__(self, dim, depth, L0=1.0, Y0=0.0, axi=False, plant=0):
        self.dim, self.depth, self.L0 = dim, depth, L0
        # AXI (2D): x is the axis, y the radius; Y0 the box's lower y
        self.axi, self.Y0 = axi, Y0
        # planted mistake (tests only): 1 = the AXI term with its sign flipped
        self.plant = plant
        self.sigma = default_sigma(dim)
        self.TOLERANCE = 1e-3
        self.NITERMAX, self.NITERMIN = 100, 1
        self.smoother = "jacobi"
        self.periodic = [False] * dim

    def delta(self, l):
        return self.L0 / (1 << l)

    # -- coefficients on every level (restriction ({mu, rho}))
    def set_coefficients(self, mu, rho):
        """mu: list of dim face arrays or a constant; rho: cell array or a
        constant.  Restricted to every level."""
        dim, D = self.dim, self.depth
        self.mu_c = None if isinstance(mu, (list, tuple)) else float(mu)
        self.rho_c = None if isinstance(rho, np.ndarray) else float(rho)
        self.mu = [None] * (D + 1)
        self.rho = [None] * (D + 1)
        if self.mu_c is None:
            self.mu[D] = [np.array(m, np.float64) for m in mu]
            for d in range(dim):
                if self.periodic[d]:
                    # face n of a periodic dimension is face 0
                    hi, lo = [slice(None)] * dim, [slice(None)] * dim
                    hi[d], lo[d] = -1, 0
                    self.mu[D][d][tuple(hi)] = self.mu[D][d][tuple(lo)]
            for l in range(D - 1, -1, -1):
                self.mu[l] = [self._restrict_face(self.mu[l + 1][d], d) for d in range(dim)]
        if self.rho_c is None:
            self.rho[D] = np.asarray(rho, np.float64)
            for l in range(D - 1, -1, -1):
                self.rho[l] = self._restrict(self.rho[l + 1])

    def _restrict(self, f):
        """restriction_average: children in foreach_child order."""
        dim = self.dim
        s = np.zeros(tuple(m // 2 for m in f.shape))
        for ch in np.ndindex(*(2,) * dim):
            s = s + f[tuple(slice(o, None, 2) for o in ch)]
        return s / float(1 << dim)

    def _restrict_face(self, f, d):
        """face_average for component d (n+1 along d)."""
        dim = self.dim
        if dim == 2:
            t = 1 - d
            order = [(0,), (1,)]
            tdims = [t]
        else:
            tdims = {0: [1, 2], 1: [2, 0], 2: [0, 1]}[d]
            order = [(0, 0), (1, 0), (0, 1), (1, 1)]
        s = None
        for o in order:
            sl = [slice(None)] * dim
            sl[d] = slice(0, None, 2)
            for td, v in zip(tdims, o):
                sl[td] = slice(v, None, 2)
            term = f[tuple(sl)]
            s = term if s is None else s + term
        return s / float(len(order))

    def _mu(self, l, d, hi, n, at=None):
        if self.mu_c is not None:
            return self.mu_c
        f = _face(self.mu[l][d], d, hi, n, self.dim)
        return f if at is None else f[at]

    def _dtr(self, l, dt, at=None):
        if self.rho_c is not None:
            return dt / self.rho_c
        return dt / (self.rho[l] if at is None else self.rho[l][at])

    def _lambda(self, c, l, n, dt, at=None):
        """lambda of component c: 1, or with AXI for the y component
        1 + dt/rho (mu.x[] + mu.x[1] + mu.y[] + mu.y[0,1])/2/y^2 (at: one
        cell, else every cell)."""
        if not (self.axi and c == 1):
            return 1.
        y = (np.arange(n) + 0.5) * self.delta(l) + self.Y0
        y = y[None, :] if at is None else y[at[1]]
        s = (self._mu(l, 0, False, n, at) + self._mu(l, 0, True, n, at) + self._mu(l, 1, False, n, at) +
             self._mu(l, 1, True, n, at))
        t = self._dtr(l, dt, at) * s / 2. / (y * y)
        return 1. - t if self.plant == 1 else 1. + t

    # -- the stencils
    def _terms(self, P, c, l, n, residual, at=None):
        """The bracket of relax_viscosity (numerator terms, residual=False)
        or of residual_viscosity (residual=True) for component c; P are the
        padded components (at: one cell, else every cell)."""
        dim = self.dim
        if at is None:
            U = lambda comp, *pairs: _sh(P[comp], _unit(dim, *pairs), n, dim)
        else:
            U = lambda comp, *pairs: P[comp][tuple(1 + i + o for i, o in zip(at, _unit(dim, *pairs)))]
        u0 = U(c)
        mch, mcl = self._mu(l, c, True, n, at), self._mu(l, c, False, n, at)
        if residual:
            S = 2. * mch * (U(c, (c, 1)) - u0) - 2. * mcl * (u0 - U(c, (c, -1)))
        else:
            S = 2. * mch * U(c, (c, 1)) + 2. * mcl * U(c, (c, -1))
        for e in _others(c, dim):
            meh, mel = self._mu(l, e, True, n, at), self._mu(l, e, False, n, at)
            up = (U(e, (c, 1)) + U(e, (c, 1), (e, 1))) / 4.
            um = (U(e, (c, -1)) + U(e, (c, -1), (e, 1))) / 4.
            lp = (U(e, (c, 1), (e, -1)) + U(e, (c, 1))) / 4.
            lm = (U(e, (c, -1), (e, -1)) + U(e, (c, -1))) / 4.
            if residual:
                S = S + meh * (U(c, (e, 1)) - u0 + up - um)
                S = S - mel * (u0 - U(c, (e, -1)) + lp - lm)
            else:
                S = S + meh * (U(c, (e, 1)) + up - um)
                S = S - mel * (-U(c, (e, -1)) + lp - lm)
        return S

    def _diag(self, c, l, n, at=None):
        dim = self.dim
        s = 2. * self._mu(l, c, True, n, at) + 2. * self._mu(l, c, False, n, at)
        for e in _others(c, dim):
            s = s + self._mu(l, e, True, n, at) + self._mu(l, e, False, n, at)
        return s

    def _target(self, da, res, l, dt):
        """The value w of relax_viscosity for every cell and component."""
        dim = self.dim
        n = 1 << l
        D = self.delta(l)
        D2 = D * D
        P = [_pad(da[c], self.sigma[c], dim, self.periodic) for c in range(dim)]
        dtr = self._dtr(l, dt)
        return [(dtr * self._terms(P, c, l, n, False) + res[c] * D2) /
                (D2 * self._lambda(c, l, n, dt) + dtr * self._diag(c, l, n)) for c in range(dim)]

    def relax(self, da, res, l, dt):
        """relax_viscosity: returns the new components."""
        if self.smoother == "gs":
            # the serial loop of foreach_level (l), i-major, w = u: each
            # component is overwritten as it is computed; the ghosts keep
            # the values of the sweep's start
            dim = self.dim
            n = 1 << l
            D = self.delta(l)
            D2 = D * D
            P = [_pad(da[c], self.sigma[c], dim, self.periodic) for c in range(dim)]
            for at in np.ndindex(*(n,) * dim):
                dtr = self._dtr(l, dt, at)
                for c in range(dim):
                    w = (dtr * self._terms(P, c, l, n, False, at) + res[c][at] * D2) / \
                        (D2 * self._lambda(c, l, n, dt, at) + dtr * self._diag(c, l, n, at))
                    P[c][tuple(1 + i for i in at)] = w
            return [P[c][(slice(1, n + 1),) * dim].copy() for c in range(dim)]
        if self.smoother == "jacobi":
            w = self._target(da, res, l, dt)
            return [(da[c] + 2. * w[c]) / 3. for c in range(self.dim)]
        parity = np.indices((1 << l,) * self.dim).sum(0) & 1
        for colour in (0, 1):
            w = self._target(da, res, l, dt)
            da = [np.where(parity == colour, w[c], da[c]) for c in range(self.dim)]
        return da

    def residual(self, u, r, dt):
        """residual_viscosity on the finest level: (res list, max |res|)."""
        dim, l = self.dim, self.depth
        n = 1 << l
        D = self.delta(l)
        P = [_pad(u[c], self.sigma[c], dim, self.periodic) for c in range(dim)]
        dtr = self._dtr(l, dt)
        res = []
        m = 0.
        for c in range(dim):
            rc = (r[c] - self._lambda(c, l, n, dt) * u[c]) + dtr * self._terms(P, c, l, n, True) / (D * D)
            res.append(rc)
            m = max(m, float(np.abs(rc).max()))
        return res, m

    def _bilinear(self, coarse, sig):
        dim = self.dim
        nc = coarse.shape[0]
        P = _pad(coarse, sig, dim, self.periodic)
        n = 2 * nc
        idx = np.indices((n,) * dim)
        p = [(idx[d] >> 1) + 1 for d in range(dim)]
        o = [2 * (idx[d] & 1) - 1 for d in range(dim)]
        g = lambda *sh: P[tuple(p[d] + (o[d] if sh[d] else 0) for d in range(dim))]
        if dim == 2:
            return (9. * g(0, 0) + 3. * (g(1, 0) + g(0, 1)) + g(1, 1)) / 16.
        return (27. * g(0, 0, 0) + 9. * (g(1, 0, 0) + g(0, 1, 0) + g(0, 0, 1))
                + 3. * (g(1, 1, 0) + g(1, 0, 1) + g(0, 1, 1)) + g(1, 1, 1)) / 64.

    def mg_cycle(self, u, res, dt, nrelax, minlevel=0):
        dim, D = self.dim, self.depth
        # restriction (res) to every level
        rl = [None] * (D + 1)
        rl[D] = res
        for l in range(D - 1, -1, -1):
            rl[l] = [self._restrict(rl[l + 1][c]) for c in range(dim)]
        da = None
        for l in range(min(minlevel, D), D + 1):
            n = 1 << l
            if l == min(minlevel, D):
                da = [np.zeros((n,) * dim) for _ in range(dim)]
            else:
                da = [self._bilinear(da[c], self.sigma[c]) for c in range(dim)]
            for _ in range(nrelax):
                da = self.relax(da, rl[l], l, dt)
        return [u[c] + da[c] for c in range(dim)]

    def explicit(self, u, dt):
        """viscosity_explicit(): returns (u + res, stats with only resb)."""
        u = [np.array(x, np.float64) for x in u]
        res, resb = self.residual(u, u, dt)
        s = {"i": 0, "resb": resb, "resa": 0.0, "sum": 0.0, "nrelax": 0, "minlevel": 0}
        return [u[c] + res[c] for c in range(self.dim)], s

    def solve(self, u, dt, nrelax=0, tolerance=0.0):
        """viscosity(): returns (u, stats, residuals after each cycle)."""
        u = [np.array(x, np.float64) for x in u]
        r = [x.copy() for x in u]
        s = {"i": 0, "sum": float(sum(x.sum() for x in r)), "nrelax": nrelax if nrelax > 0 else 4,
             "minlevel": 0}
        res, resb = self.residual(u, r, dt)
        s["resb"] = s["resa"] = resb
        tol = tolerance if tolerance else self.TOLERANCE
        hist = []
        while s["i"] < self.NITERMAX and (s["i"] < self.NITERMIN or s["resa"] > tol):
            u = self.mg_cycle(u, res, dt, s["nrelax"])
            res, s["resa"] = self.residual(u, r, dt)
            hist.append(s["resa"])
            if s["resa"] > tol:
                if resb / s["resa"] < 1.2 and s["nrelax"] < 100:
                    s["nrelax"] += 1
                elif resb / s["resa"] > 10 and s["nrelax"] > 2:
                    s["nrelax"] -= 1
            resb = s["resa"]
            s["i"] += 1
        return u, s, hist
