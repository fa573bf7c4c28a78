"""The reference's embedded-boundary Poisson test case (test/neumann.c, and
test/dirichlet.c, the same file with DIRICHLET=1), generated on the CPU
with the oracle's fractions() and fragment geometry (test infrastructure):
the Johansen-Colella problems on [-0.5, 0.5]^2, exact solution
r^4 cos(3 theta).

- Dirichlet (Problem 1): the fluid is r <= 0.30 + 0.15 cos(6 theta), with
  a[embed] = exact at the fragment centroid;
- Neumann (Problem 3): the fluid is r >= 0.25 + 0.05 cos(6 theta), with the
  exact normal gradient at the fragment centroid, and the box ghosts set to
  the exact solution at the ghost centres (`a[left] = exact (x - Delta/2.,
  y)`, homogeneous form 0).

alpha = fs, lambda = 0, b = 7 r^2 cos(3 theta) cs with (r, theta) at the
centroid of the cell's fluid part, a = exact where cs > 0 and nodata (1e30)
elsewhere; third-order face gradients."""
import numpy as np

from oracle.oracle import (BC_VALUE, EMBED_DIRICHLET, EMBED_NEUMANN, cut_centroids, embed_geometry,
                           fractions)

NODATA = 1e30


def exact(x, y):
    theta = np.arctan2(y, x)
    r2 = x * x + y * y
    return r2 * r2 * np.cos(3. * theta)


def problem(N, dirichlet):
    D = 1.0 / N
    xv = -0.5 + np.arange(N + 1) * D
    X, Y = np.meshgrid(xv, xv, indexing="ij")
    theta = np.arctan2(Y, X)
    r = np.sqrt(X * X + Y * Y)
    phi = (0.30 + 0.15 * np.cos(6. * theta) - r) if dirichlet else (r - (0.25 + 0.05 * np.cos(6. * theta)))
    cs, fx, fy = fractions(phi)
    xc = -0.5 + (np.arange(N) + 0.5) * D
    XC, YC = np.meshgrid(xc, xc, indexing="ij")
    px, py, _, nx, ny = embed_geometry(cs, fx, fy)
    cut = (cs > 0) & (cs < 1)
    xe, ye = XC + px * D, YC + py * D
    if dirichlet:
        bc = np.where(cut, exact(xe, ye), 0.0)
    else:
        th = np.arctan2(ye, xe)
        rr = np.sqrt(xe * xe + ye * ye)
        dsdtheta = -3. * rr ** 3 * np.sin(3. * th)
        dsdr = 4. * rr ** 3 * np.cos(3. * th)
        bc = np.where(cut, nx * (dsdr * np.cos(th) - dsdtheta * np.sin(th)) +
                      ny * (dsdr * np.sin(th) + dsdtheta * np.cos(th)), 0.0)
    cx, cy = cut_centroids(cs, fx, fy)
    xb, yb = XC + np.where(cut, cx * D, 0.0), YC + np.where(cut, cy * D, 0.0)
    b = 7. * (xb * xb + yb * yb) * np.cos(3. * np.arctan2(yb, xb)) * cs
    h = D / 2
    box = {1: exact(-0.5 - h + 0 * xc, xc), 0: exact(0.5 + h + 0 * xc, xc),
           2: exact(xc, 0.5 + h + 0 * xc), 3: exact(xc, -0.5 - h + 0 * xc)}
    return {"N": N, "depth": int(round(np.log2(N))), "cs": cs, "fx": fx, "fy": fy, "bc": bc, "b": b,
            "a": np.where(cs > 0, exact(XC, YC), NODATA), "exact": exact(XC, YC), "box": box,
            "kind": EMBED_DIRICHLET if dirichlet else EMBED_NEUMANN, "vol": D * D * cs}


def setup_oracle(o, p):
    """Load problem p into an OracleGrid of depth log2 N."""
    N = p["N"]
    o.set_alpha([p["fx"], p["fy"]])
    o.set_embed(p["cs"], p["fx"], p["fy"], p["kind"], p["bc"], third=True)
    for side, vals in p["box"].items():
        v = np.zeros(N + 4)
        v[2:-2] = vals
        o.set_bc(side, BC_VALUE, v)
    o.set_a(p["a"])
    o.set_b(p["b"])


def norms(p, a):
    """normf of the error over all, partial and full cells (avg weighted by
    the cell volume cs Delta^2, and max), test/neumann.c:190-205."""
    cs = p["cs"]
    e = a - p["exact"]
    out = []
    for mask in (cs > 0, (cs > 0) & (cs < 1), cs >= 1.):
        ee, w = np.abs(e[mask]), p["vol"][mask]
        out += [(ee * w).sum() / w.sum(), ee.max()]
    return out
