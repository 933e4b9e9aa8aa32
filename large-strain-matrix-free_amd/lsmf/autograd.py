"""Differentiation of the residual for test and study harnesses.

``ResidualFunction`` wraps the FP64 residual kernel (``lsmf_residual``) as a
``torch.autograd.Function``; its backward is the hand-written vector-Jacobian
product kernel (``lsmf_residual_vjp``), which recomputes the quadrature-point
kinematics from the saved displacement instead of storing them.

``misfit_gradient`` differentiates an objective of the converged Newton
solution through one adjoint solve: with R(u, p) + f = 0 at equilibrium and
dR/du = -K (K symmetric), dJ/dp = lambda^T dR/dp with K lambda = dJ/du on the
free DoFs -- one GMG-preconditioned CG solve and one residual VJP.
"""
from __future__ import annotations

import numpy as np
import torch

from . import NeoHookOperator


class ResidualFunction(torch.autograd.Function):
    """R(u; mu, lambda) = -sum_cells int P(F) : Grad N dV (internal forces)"""

    @staticmethod
    def forward(ctx, u, mu, lam, op):
        op.set_material(mu.detach().cpu().numpy(), lam.detach().cpu().numpy())
        op.set_linearization(u.detach().contiguous())
        r = torch.empty_like(u)
        op.residual(r)
        ctx.op = op
        ctx.save_for_backward(u, mu, lam)
        return r

    @staticmethod
    def backward(ctx, w):
        u, mu, lam = ctx.saved_tensors
        op = ctx.op
        # recompute the state (the kernel recomputes F, J, ln J per qp)
        op.set_material(mu.detach().cpu().numpy(), lam.detach().cpu().numpy())
        op.set_linearization(u.detach().contiguous())
        du = torch.empty_like(u)
        dmu = torch.empty(op.n_cells, dtype=torch.float64, device=u.device)
        dlam = torch.empty_like(dmu)
        op.residual_vjp(w.contiguous(), du, dmu, dlam)
        return du, dmu.to(mu.device), dlam.to(lam.device), None


def residual(u, mu, lam, op):
    return ResidualFunction.apply(u, mu, lam, op)


def misfit_gradient(solver, mesh, u_star, mu, lam, target, nodes, cons, cg_tol=1e-12):
    """J = 1/2 sum_{DoFs of `nodes`, free} (u - target)^2 at the converged
    solution u_star of `solver`; returns (J, dJ/dmu, dJ/dlambda) per cell."""
    dim = mesh.dim
    dofs = (nodes[:, None] * dim + np.arange(dim)[None, :]).reshape(-1)
    dofs = dofs[cons[dofs] == 0]
    diff = u_star[dofs] - target[dofs]
    J = 0.5 * float(diff @ diff)
    g = np.zeros(mesh.n_dofs)
    g[dofs] = diff
    dev = torch.device("cuda")
    solver.set_state(u_star)
    gd = torch.tensor(g, device=dev)
    lam_adj = torch.zeros_like(gd)
    solver.cg(lam_adj, gd, cg_tol * float(np.linalg.norm(g)))
    op = NeoHookOperator(mesh, precision="fp64", caching="tensor4", constrained=cons)
    op.set_material(mu, lam)
    op.set_linearization(torch.tensor(u_star, device=dev))
    du = torch.empty_like(gd)
    dmu = torch.empty(mesh.n_cells, dtype=torch.float64, device=dev)
    dla = torch.empty_like(dmu)
    op.residual_vjp(lam_adj, du, dmu, dla)
    return J, dmu.cpu().numpy(), dla.cpu().numpy()
