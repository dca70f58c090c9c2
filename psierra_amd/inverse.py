"""Parameter sensitivities of a Tet4 Neo-Hookean equilibrium, on the GPU.

pSierra's MaterialParameterIdentifier differentiates its forward model by
finite differences: scipy.optimize.least_squares(..., jac='2-point') re-solves
the whole nonlinear problem once per parameter and iteration
(src/psierra/inverse/parameter_identification.py:300-379).  At an equilibrium
f_int(u; mu, lambda)[free] = f_ext[free] (a load that does not depend on the
parameters) the implicit function theorem gives the same derivatives from one
factorisation of the free-DOF tangent K = K_T(u)[free, free]:

* adjoint: for an objective J(u), solve K lam = dJ/du[free]; then
  dJ/dp = -lam^T df_int/dp, one VJP with w = lam (zero on prescribed DOFs),
  which returns dJ/dmu_e and dJ/dlambda_e for every element at once;
* forward: for a parameter direction (dmu, dlam), du[free] = -K^{-1}
  df_int/dp . (dmu, dlam); f_int is linear in (mu, lambda) jointly
  (S = mu (...) + lambda (...)), so df_int/dp . (dmu, dlam) is f_int
  evaluated at (dmu, dlam) for the same u.

K_T is symmetric, so the adjoint system uses K itself.  The factorisation is
scipy's SuperLU on the host, as the reference's solves (src/psierra/solver/
sparse_linear_solver.py:16-19); assembly, the free-DOF extraction and the
VJP run on the GPU.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse.linalg as spla

from .engine import AssemblyPlan, _torch, device_f64, host_array
from .solver import FreeDofSystem


class EquilibriumSensitivity:
    """Sensitivities at a converged state u of ``plan`` (Tet4) with
    per-element (mu, lam) and free DOFs ``free_dofs``."""

    def __init__(self, plan: AssemblyPlan, u, mu, lam, free_dofs, precision: str = "fp64") -> None:
        if plan.npe != 4:
            raise ValueError("EquilibriumSensitivity takes a Tet4 plan (the adjoint kernel is Tet4)")
        torch = _torch()
        self.plan = plan
        self.precision = precision
        self.u = plan._device_f64(u, plan.n_dofs, "u")
        self.mu = plan._device_f64(mu, plan.n_elements, "mu")
        self.lam = plan._device_f64(lam, plan.n_elements, "lam")
        self.free = np.ascontiguousarray(free_dofs, dtype=np.int64)
        self._free_dev = torch.as_tensor(self.free, device=plan.device)
        _, kval = plan.assemble(self.u, self.mu, self.lam, precision=precision)
        K_ff = FreeDofSystem(plan, self.free).csc(kval)
        self._lu = spla.splu(K_ff)

    def adjoint(self, dJ_du):
        """(dJ/dmu, dJ/dlambda) per element, as device tensors, for the
        objective gradient dJ/du (a full (3N,) vector; its prescribed entries
        are not used)."""
        torch = _torch()
        g = np.asarray(dJ_du, dtype=np.float64).reshape(-1)
        if g.size != self.plan.n_dofs:
            raise ValueError(f"dJ/du has {g.size} entries, expected {self.plan.n_dofs}")
        lam_free = self._lu.solve(g[self.free])
        w = torch.zeros(self.plan.n_dofs, dtype=torch.float64, device=self.plan.device)
        w[self._free_dev] = device_f64(lam_free, self.plan.device)
        _, gmu, glam = self.plan.vjp(self.u, self.mu, self.lam, w, precision=self.precision)
        return -gmu, -glam

    def forward(self, dmu, dlam) -> np.ndarray:
        """du (3N,) for the parameter direction (dmu, dlam) (zero on the
        prescribed DOFs)."""
        df, _ = self.plan.assemble(self.u, dmu, dlam, precision=self.precision, want_k=False)
        du = np.zeros(self.plan.n_dofs)
        du[self.free] = -self._lu.solve(host_array(df)[self.free])
        return du
