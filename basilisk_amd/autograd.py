"""Differentiable Poisson-Helmholtz solves for PyTorch.

``PoissonSolver`` wraps two device-resident hierarchies of libbmg: the
forward one (the BCs of the unknown) and the adjoint one (the same BC kinds,
homogeneous).  Calling it solves

    div(alpha grad phi) + lambda phi = b      (poisson.h:402-441, poisson())

and returns phi as a float64 CUDA tensor.  The backward pass uses that the
discrete operator is self-adjoint: given g = dJ/dphi it solves the same
operator with homogeneous BCs and right-hand side g for the adjoint state
psi, then

    dJ/db      = psi
    dJ/dalpha  = grad(psi) . grad(phi) on each face (wall faces use the ghost
                 of phi), computed by one fused kernel on the device
    dJ/dlambda = -psi phi

(bmg_coef_gradient).  The gradients are FP64 whatever the solver precision.
With ``recompute=True`` the forward solution is not kept between forward and
backward; the backward solves the forward problem again first.

With no Dirichlet side and lambda = 0 the operator is singular (phi has a
free constant): the adjoint right-hand side then has its mean removed before
the adjoint solve, which is exact when J does not depend on that constant
(then sum(dJ/dphi) = 0 already) and otherwise differentiates J along
mean-free perturbations of phi only.  An adjoint solve that does not reach
the tolerance raises RuntimeError instead of returning a wrong gradient.
"""
import torch

import basilisk_amd as bmg


class PoissonSolver:
    """dim, depth: the uniform grid (2^depth cells per dimension) over
    [origin, origin + L0]^dim.  ``bc``: {side: (kind, values or None)} for the
    sides that are not symmetry (values: numpy array of the boundary
    expression on the finest-level faces, as bmg.Grid.set_bc takes).
    ``precision``: "fp64" or "mixed"; ``tolerance``: max |residual|."""

    def __init__(self, dim, depth, bc=None, L0=1.0, origin=(0.0, 0.0, 0.0), precision="mixed",
                 smoother="rbgs", tolerance=1e-9, nitermax=100, recompute=False):
        self.dim, self.depth = dim, depth
        self.tolerance = tolerance
        self.recompute = recompute
        self.fwd = bmg.Grid(dim, depth, L0, origin, precision=precision, smoother=smoother)
        self.adj = bmg.Grid(dim, depth, L0, origin, precision=precision, smoother=smoother)
        for g in (self.fwd, self.adj):
            g.set_niter(nitermax, 1)
        self.has_dirichlet = False
        for side, (kind, values) in (bc or {}).items():
            if kind == bmg.BC_PERIODIC:
                raise ValueError("periodic boundaries are not differentiable here")
            self.fwd.set_bc(side, kind, values)
            self.adj.set_bc(side, kind, None)
            self.has_dirichlet |= kind == bmg.BC_DIRICHLET
        self.shape = self.fwd.shape()
        self.last_stats = None  # mgstats of the last forward solve
        self.last_adjoint_stats = None  # ... and of the last adjoint solve

    def _coefficients(self, g, alpha, lam):
        if alpha is None:
            g.set_alpha(1.0)
        else:
            g.set_alpha([a.detach().contiguous() for a in alpha])
        if lam is None:
            g.set_lambda(0.0)
        elif isinstance(lam, (int, float)):
            g.set_lambda(float(lam))
        else:
            g.set_lambda(lam.detach().contiguous())

    def _solve(self, g, rhs, alpha, lam):
        self._coefficients(g, alpha, lam)
        g.set_b(rhs.detach().contiguous())
        g.set_a(torch.zeros(self.shape, dtype=torch.float64, device=rhs.device))
        st = g.poisson(tolerance=self.tolerance)
        out = torch.empty(self.shape, dtype=torch.float64, device=rhs.device)
        g.get_a(out)
        return out, st

    def __call__(self, b, alpha=None, lam=None):
        alpha = list(alpha) if alpha is not None else []
        return _PoissonFunction.apply(self, b, lam, *alpha)


class _PoissonFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, solver, b, lam, *alpha):
        alpha_l = list(alpha) if alpha else None
        phi, st = solver._solve(solver.fwd, b, alpha_l, lam)
        solver.last_stats = st
        ctx.solver = solver
        ctx.has_alpha = bool(alpha)
        ctx.lam_const = not torch.is_tensor(lam)
        ctx.lam_value = lam if ctx.lam_const else None
        tensors = [b] + ([lam] if torch.is_tensor(lam) else []) + list(alpha)
        if not solver.recompute:
            tensors.append(phi)
        ctx.save_for_backward(*tensors)
        return phi

    @staticmethod
    def backward(ctx, gphi):
        solver = ctx.solver
        saved = list(ctx.saved_tensors)
        b = saved.pop(0)
        lam = ctx.lam_value if ctx.lam_const else saved.pop(0)
        phi = None if solver.recompute else saved.pop()
        alpha = saved if ctx.has_alpha else None
        if phi is None:
            phi, _ = solver._solve(solver.fwd, b, alpha, lam)
        rhs = gphi.contiguous()
        singular = not solver.has_dirichlet and (
            (ctx.lam_const and not lam) or (not ctx.lam_const and not bool(torch.any(lam != 0))))
        if singular:
            rhs = rhs - rhs.mean()
        psi, solver.last_adjoint_stats = solver._solve(solver.adj, rhs, alpha, lam)
        if not solver.last_adjoint_stats["resa"] <= solver.tolerance:
            raise RuntimeError("adjoint solve did not reach the tolerance %g: %s"
                               % (solver.tolerance, solver.last_adjoint_stats))
        g = solver.fwd
        g.set_a(phi.detach().contiguous())
        need_a = ctx.has_alpha and any(ctx.needs_input_grad[3:])
        need_l = (not ctx.lam_const) and ctx.needs_input_grad[2]
        ga = [torch.empty(g.face_shape(c), dtype=torch.float64, device=phi.device) for c in range(solver.dim)] \
            if need_a else None
        gl = torch.empty(solver.shape, dtype=torch.float64, device=phi.device) if need_l else None
        if need_a or need_l:
            g.coef_gradient(psi, ga if need_a else [None] * 3, gl)
            torch.cuda.synchronize()
        grads_alpha = tuple(ga) if need_a else tuple(None for _ in range(len(alpha) if alpha else 0))
        return (None, psi if ctx.needs_input_grad[1] else None, gl) + grads_alpha


class TreePoissonSolver:
    """Differentiable poisson() on an adaptive tree (a basilisk_amd.Tree
    whose mesh, boundary conditions and settings the caller prepared).
    Calling it with b (leaf order), alpha (face values, the order of
    Tree.set_alpha_faces, or None for unity) and lambda (a leaf tensor, a
    constant or None) solves the tree problem and returns the leaf values
    of the solution.  The tree operator is not symmetric, so the backward
    pass solves the transposed system L^T psi = dJ/da (Tree.adjoint: the
    transposed V-cycle as defect correction) and returns dJ/db = psi,
    dJ/dalpha per face value and dJ/dlambda = -psi a.  ``recompute=True``
    solves the forward problem again in backward (the adjoint linearises
    about the tree's current solution)."""

    def __init__(self, tree, tolerance=1e-9, adjoint_tolerance=None, nitermax=100, recompute=False):
        self.tree = tree
        self.tolerance = tolerance
        self.adjoint_tolerance = tolerance if adjoint_tolerance is None else adjoint_tolerance
        self.nitermax = nitermax
        self.recompute = recompute
        tree.set_niter(nitermax, 1)
        self.last_stats = None
        self.last_adjoint_stats = None

    def _coefficients(self, alpha, lam):
        t = self.tree
        if alpha is None:
            t.set_alpha_const(1.0)
        else:
            t.set_alpha_faces(alpha.detach().cpu().numpy())
        if lam is None or not torch.is_tensor(lam):
            t.set_lambda(0.0 if lam is None else float(lam))
        else:
            t.set_lambda(lam.detach().contiguous())

    def _solve(self, b, alpha, lam):
        t = self.tree
        self._coefficients(alpha, lam)
        t.set_b(b.detach().contiguous())
        t.set_a(torch.zeros(t.nleaves(), dtype=torch.float64, device=b.device))
        st = t.poisson(tolerance=self.tolerance)
        a = torch.empty(t.nleaves(), dtype=torch.float64, device=b.device)
        t.get_a(a)
        return a, st

    def __call__(self, b, alpha=None, lam=None):
        return _TreePoissonFunction.apply(self, b, alpha, lam)


class _TreePoissonFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, solver, b, alpha, lam):
        a, st = solver._solve(b, alpha, lam)
        solver.last_stats = st
        ctx.solver = solver
        ctx.alpha, ctx.lam = alpha, lam
        if solver.recompute:
            ctx.save_for_backward(b)
        else:
            ctx.save_for_backward(b, a)
        return a

    @staticmethod
    def backward(ctx, ga_out):
        solver = ctx.solver
        saved = ctx.saved_tensors
        if solver.recompute:
            _, solver.last_stats = solver._solve(saved[0], ctx.alpha, ctx.lam)
        else:
            # the tree may have solved something else since: linearise about
            # this call's coefficients and solution
            solver._coefficients(ctx.alpha, ctx.lam)
            solver.tree.set_a(saved[1].detach().contiguous())
        lam_field = torch.is_tensor(ctx.lam)
        need_alpha = ctx.alpha is not None and ctx.needs_input_grad[2]
        need_lam = lam_field and ctx.needs_input_grad[3]
        psi, galpha, glam, st = solver.tree.adjoint(ga_out.contiguous(), tolerance=solver.adjoint_tolerance,
                                                    nitermax=solver.nitermax, alpha=need_alpha, lam=need_lam)
        solver.last_adjoint_stats = st
        if not st["resa"] <= solver.adjoint_tolerance:
            raise RuntimeError("tree adjoint did not reach the tolerance %g: %s" % (solver.adjoint_tolerance, st))
        return (None, psi if ctx.needs_input_grad[1] else None, galpha if need_alpha else None,
                glam if need_lam else None)
