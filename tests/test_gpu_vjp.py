"""GPU adjoint of the Tet4 internal force.

Checked three independent ways:
* K_T w against the oracle's assembled K (FP64) times w;
* d(w^T f)/d(mu_e, lambda_e) against the oracle's element forces with unit
  parameters (f is exactly linear in (mu, lambda)) and against central finite
  differences on 100 seeded elements (step 1e-6 relative, FP64);
* the chain rule: sum_e mu_e0 d/dmu_e equals scipy's 2-point derivative of
  s -> w^T f(s mu0, lambda0), the Jacobian scheme pSierra's parameter
  identification uses through least_squares
  (src/psierra/inverse/parameter_identification.py:370-379).
Tolerance for the FD and chain-rule checks: 1e-4 relative (FP32 and FP64).
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import assembly as A
from oracle import fem
from psierra_amd import workloads as W
from tests.gpu_helpers import banner, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def vjp_case():
    import torch

    from psierra_amd.engine import AssemblyPlan

    wl = W.vjp_box_tet4(n=12, seed=3)
    plan = AssemblyPlan(wl.mesh.nodes, wl.mesh.elements)
    dev = plan.device
    t = {
        "u": torch.as_tensor(wl.u.reshape(-1), device=dev),
        "mu": torch.as_tensor(wl.mu, device=dev),
        "lam": torch.as_tensor(wl.lam, device=dev),
        "w": torch.as_tensor(wl.extra["w"], device=dev),
    }
    return wl, plan, t


def _elem_wf(wl, e, mu, lam):
    c = wl.mesh.elements[e]
    _, f = fem.tet4_force_and_tangent(wl.mesh.nodes[c], wl.u[c], fem.nh_material(mu, lam))
    return float(f @ wl.extra["w"].reshape(-1, 3)[c].ravel())


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_vjp_matches_oracle(vjp_case, precision):
    wl, plan, t = vjp_case
    banner(f"vjp {precision}")
    gu, gmu, glam = plan.vjp(t["u"], t["mu"], t["lam"], t["w"], precision=precision)
    gu, gmu, glam = gu.cpu().numpy(), gmu.cpu().numpy(), glam.cpu().numpy()
    _, K = A.assemble(wl.mesh.nodes, wl.mesh.elements, wl.u, wl.mu, wl.lam)
    w = wl.extra["w"]
    tol = 1e-12 if precision == "fp64" else 1e-5
    assert rel_err(gu, K @ w) < tol
    # f linear in (mu, lambda): d(w^T f_e)/dmu = w_e . f_e(mu=1, lam=0)
    E = wl.mesh.n_elements
    f_mu = A.evaluate_elements(wl.mesh.nodes, wl.mesh.elements, wl.u, np.ones(E), np.zeros(E), want_k=False)[1]
    f_lam = A.evaluate_elements(wl.mesh.nodes, wl.mesh.elements, wl.u, np.zeros(E), np.ones(E), want_k=False)[1]
    we = w.reshape(-1, 3)[wl.mesh.elements].reshape(E, -1)
    gmu_ref = (f_mu * we).sum(axis=1)
    glam_ref = (f_lam * we).sum(axis=1)
    ptol = 1e-11 if precision == "fp64" else 1e-5
    assert rel_err(gmu, gmu_ref) < ptol
    assert rel_err(glam, glam_ref) < ptol


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_vjp_params_match_central_fd(vjp_case, precision):
    wl, plan, t = vjp_case
    banner(f"vjp fd params {precision}")
    _, gmu, glam = plan.vjp(t["u"], t["mu"], t["lam"], t["w"], precision=precision)
    gmu, glam = gmu.cpu().numpy(), glam.cpu().numpy()
    rng = np.random.default_rng(33)
    for e in rng.choice(wl.mesh.n_elements, 100, replace=False):
        mu, lam = wl.mu[e], wl.lam[e]
        h = 1e-6 * mu
        d_mu = (_elem_wf(wl, e, mu + h, lam) - _elem_wf(wl, e, mu - h, lam)) / (2 * h)
        h = 1e-6 * lam
        d_lam = (_elem_wf(wl, e, mu, lam + h) - _elem_wf(wl, e, mu, lam - h)) / (2 * h)
        assert abs(gmu[e] - d_mu) <= 1e-4 * max(abs(d_mu), 1e-3 * np.abs(gmu).max())
        assert abs(glam[e] - d_lam) <= 1e-4 * max(abs(d_lam), 1e-3 * np.abs(glam).max())


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_vjp_displacement_matches_central_fd(vjp_case, precision):
    """d(w^T f)/du_i by central differences of the GPU FP64 forward on 100 DOFs."""
    import torch

    wl, plan, t = vjp_case
    banner(f"vjp fd u {precision}")
    gu, _, _ = plan.vjp(t["u"], t["mu"], t["lam"], t["w"], precision=precision, want_params=False)
    gu = gu.cpu().numpy()
    rng = np.random.default_rng(34)
    scale = np.abs(wl.u).max()
    for i in rng.choice(plan.n_dofs, 100, replace=False):
        h = 1e-6 * scale
        vals = []
        for sgn in (1.0, -1.0):
            u = t["u"].clone()
            u[i] += sgn * h
            f, _ = plan.assemble(u, t["mu"], t["lam"], want_k=False)
            vals.append(float(torch.dot(f, t["w"])))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(gu[i] - fd) <= 1e-4 * max(abs(fd), 1e-2 * np.abs(gu).max())


def test_global_parameter_gradient_matches_least_squares_fd(vjp_case):
    """Chain rule vs scipy's 2-point derivative (least_squares' default jac)."""
    import torch
    from scipy.optimize._numdiff import approx_derivative

    wl, plan, t = vjp_case
    banner("vjp chain rule vs 2-point")
    _, gmu, glam = plan.vjp(t["u"], t["mu"], t["lam"], t["w"], precision="fp32")

    def wf(x):
        mu = torch.as_tensor(wl.mu * x[0], device=plan.device)
        lam = torch.as_tensor(wl.lam * x[1], device=plan.device)
        f, _ = plan.assemble(t["u"], mu, lam, want_k=False)
        return np.array([float(torch.dot(f, t["w"]))])

    J = np.atleast_2d(approx_derivative(wf, np.array([1.0, 1.0]), method="2-point"))
    ref_mu = float((gmu * t["mu"]).sum())
    ref_lam = float((glam * t["lam"]).sum())
    assert abs(J[0, 0] - ref_mu) <= 1e-4 * abs(J[0, 0])
    assert abs(J[0, 1] - ref_lam) <= 1e-4 * abs(J[0, 1])


def test_autograd_function(vjp_case):
    import torch

    from psierra_amd.autograd import internal_force

    wl, plan, t = vjp_case
    banner("vjp autograd")
    u = t["u"].clone().requires_grad_(True)
    mu = t["mu"].clone().requires_grad_(True)
    lam = t["lam"].clone().requires_grad_(True)
    f = internal_force(plan, u, mu, lam, "fp64")
    (f * t["w"]).sum().backward()
    gu, gmu, glam = plan.vjp(t["u"], t["mu"], t["lam"], t["w"])
    assert torch.equal(u.grad, gu) and torch.equal(mu.grad, gmu) and torch.equal(lam.grad, glam)
