"""torch.autograd wrapper of the internal force with the hand-written adjoint.

``internal_force(plan, u, mu, lam)`` returns f_int (3N, float64) computed by
``psx_assemble`` (force only); its backward calls ``psx_vjp``:
  du   <- K_T w,
  dmu  <- d(w^T f)/d mu_e,   dlam <- d(w^T f)/d lambda_e.
This is what parameter identification and shape sensing consume in place of
the 2-point finite-difference Jacobian that pSierra's
MaterialParameterIdentifier obtains from scipy.optimize.least_squares
(src/psierra/inverse/parameter_identification.py:370-379).
"""
from __future__ import annotations

import torch

from .engine import AssemblyPlan


class InternalForce(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, mu, lam, plan: AssemblyPlan, precision: str):
        f, _ = plan.assemble(u.detach().reshape(-1), mu.detach(), lam.detach(), want_k=False, precision=precision)
        ctx.save_for_backward(u, mu, lam)
        ctx.plan = plan
        ctx.precision = precision
        return f

    @staticmethod
    def backward(ctx, w):
        u, mu, lam = ctx.saved_tensors
        want_params = ctx.needs_input_grad[1] or ctx.needs_input_grad[2]
        gu, gmu, glam = ctx.plan.vjp(
            u.detach().reshape(-1), mu.detach(), lam.detach(), w.detach().contiguous(),
            precision=ctx.precision, want_params=want_params,
        )
        return (
            gu.reshape(u.shape) if ctx.needs_input_grad[0] else None,
            gmu if ctx.needs_input_grad[1] else None,
            glam if ctx.needs_input_grad[2] else None,
            None,
            None,
        )


def internal_force(plan: AssemblyPlan, u, mu, lam, precision: str = "fp64"):
    """Differentiable f_int(u; mu, lambda) on the GPU."""
    return InternalForce.apply(u, mu, lam, plan, precision)
