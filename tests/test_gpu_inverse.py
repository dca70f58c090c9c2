"""Parameter sensitivities of a loaded Tet4 cantilever at equilibrium
(psierra_amd.inverse.EquilibriumSensitivity), against re-solves.

pSierra's identifier gets these derivatives by 2-point finite differences of
the whole nonlinear solve (src/psierra/inverse/parameter_identification.py:
370-379).  Here, for J(u) = w . u (the mean tip deflection) and a global scale
c of the shear modulus (mu_e -> c mu_e), three routes must agree:
* the adjoint (one K solve and one VJP): dJ/dc = sum_e mu_e dJ/dmu_e;
* the forward sensitivity (one K solve): dJ/dc = w . du for the direction
  (mu, 0);
* central differences of re-solved equilibria (FP64 Newton to 1e-11).
The same for lambda; and the adjoint's per-element gradient against the
forward sensitivity of a random per-element direction.
"""
from __future__ import annotations

import numpy as np
import pytest

from psierra_amd import workloads as W
from tests.gpu_helpers import banner

pytestmark = pytest.mark.gpu


def _setup():
    from psierra_amd.boundary import DirichletConstraint, DirichletManager, select_nodes_by_plane

    wl = W.cantilever_tet4()
    clamp = select_nodes_by_plane(wl.mesh, "yz", "min")
    mgr = DirichletManager(wl.mesh.n_dofs)
    mgr.add_constraint(DirichletConstraint(clamp, (0, 1, 2), 0.0))
    tip = select_nodes_by_plane(wl.mesh, "yz", "max")
    f_ext = np.zeros(wl.mesh.n_dofs)
    f_ext[3 * tip + 2] = -0.05 / len(tip)
    w = np.zeros(wl.mesh.n_dofs)
    w[3 * tip + 2] = 1.0 / len(tip)
    return wl, mgr, f_ext, w


def _solve(wl, mgr, f_ext, mu, lam):
    from psierra_amd.analysis import NonlinearStaticAnalysis
    from psierra_amd.domain import SolidDomain
    from psierra_amd.element import GpuTet4TotalLagrangian
    from psierra_amd.material import NeoHookeanField

    mat = NeoHookeanField(mu, lam)
    dom = SolidDomain(wl.mesh, mat, GpuTet4TotalLagrangian(mat))
    res = NonlinearStaticAnalysis(dom).solve(mgr, f_ext=f_ext, n_steps=2, nr_iterations=25, nr_tol=1e-11)
    assert res.converged
    return res.displacement


def test_adjoint_forward_and_resolved_differences_agree():
    import torch

    from psierra_amd.engine import AssemblyPlan
    from psierra_amd.inverse import EquilibriumSensitivity

    wl, mgr, f_ext, w = _setup()
    mu, lam = wl.mu.copy(), wl.lam.copy()
    banner("inverse: equilibrium and sensitivities")
    u = _solve(wl, mgr, f_ext, mu, lam)
    plan = AssemblyPlan(wl.mesh.nodes, wl.mesh.elements)
    sens = EquilibriumSensitivity(plan, u, mu, lam, mgr.free_dofs)
    gmu, glam = sens.adjoint(w)
    gmu, glam = gmu.cpu().numpy(), glam.cpu().numpy()
    dev = plan.device
    zeros = torch.zeros(plan.n_elements, dtype=torch.float64, device=dev)
    for which, g in (("mu", gmu), ("lam", glam)):
        p = mu if which == "mu" else lam
        adj = float(g @ p)
        pd = torch.as_tensor(p, device=dev)
        du = sens.forward(pd, zeros) if which == "mu" else sens.forward(zeros, pd)
        fwd = float(w @ du)
        h = 1e-4
        args_p = (mu * (1 + h), lam) if which == "mu" else (mu, lam * (1 + h))
        args_m = (mu * (1 - h), lam) if which == "mu" else (mu, lam * (1 - h))
        fd = (w @ _solve(wl, mgr, f_ext, *args_p) - w @ _solve(wl, mgr, f_ext, *args_m)) / (2 * h)
        print(f"dJ/dc_{which}: adjoint {adj:.10e} forward {fwd:.10e} re-solved FD {fd:.10e}")
        assert abs(adj - fwd) <= 1e-9 * abs(fwd)
        assert abs(adj - fd) <= 1e-5 * abs(fd)
    # the per-element gradient in a random direction
    rng = np.random.default_rng(3)
    dmu = rng.standard_normal(plan.n_elements) * mu * 1e-2
    dlam = rng.standard_normal(plan.n_elements) * lam * 1e-2
    du = sens.forward(torch.as_tensor(dmu, device=dev), torch.as_tensor(dlam, device=dev))
    lhs = float(gmu @ dmu + glam @ dlam)
    assert abs(lhs - float(w @ du)) <= 1e-9 * abs(float(w @ du))
