"""GPU consistent mass and linear stiffness against the oracle.

psx_assemble_mass (Tet4 and Tet10) against oracle.assembly.mass_matrix
(pattern bit-exact, values within 1e-13 relative: only FP64 sums of
rho V c differ in order), and the element's assemble_stiffness_matrix hook
(the FP64 tangent at u = 0) against oracle.assembly.linear_stiffness
(B^T C B V, within 1e-12).
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import assembly as A
from psierra_amd import workloads as W
from tests.gpu_helpers import banner, rel_err

pytestmark = pytest.mark.gpu


def _domain(wl, rho=1.0):
    from psierra_amd.domain import SolidDomain
    from psierra_amd.element import GpuTet4TotalLagrangian, GpuTet10TotalLagrangian
    from psierra_amd.material import NeoHookeanField

    mat = NeoHookeanField(wl.mu, wl.lam, mass_density=rho)
    el = GpuTet4TotalLagrangian(mat) if wl.mesh.elements.shape[1] == 4 else GpuTet10TotalLagrangian(mat)
    return SolidDomain(wl.mesh, mat, el)


@pytest.mark.parametrize("case", ["cantilever", "box_tet4", "box_tet10"])
def test_mass_matches_oracle(case):
    wl = {"cantilever": lambda: W.cantilever_tet4(), "box_tet4": lambda: W.jittered_box_tet4(n=8, seed=4),
          "box_tet10": lambda: W.jittered_box_tet10(n=4, seed=2)}[case]()
    rho = 7850.0
    banner(f"mass {case}")
    M = _domain(wl, rho).assemble_mass_matrix()
    M_ref = A.mass_matrix(wl.mesh.nodes, wl.mesh.elements, rho)
    assert np.array_equal(M.indptr, M_ref.indptr) and np.array_equal(M.indices, M_ref.indices)
    assert rel_err(M.data, M_ref.data) < 1e-13


def test_mass_per_element_density():
    from psierra_amd.engine import AssemblyPlan

    wl = W.jittered_box_tet4(n=6, seed=9)
    rho = np.random.default_rng(1).uniform(1e3, 8e3, wl.mesh.n_elements)
    plan = AssemblyPlan(wl.mesh.nodes, wl.mesh.elements)
    banner("mass per-element density")
    kval = plan.assemble_mass(rho).cpu().numpy()
    M_ref = A.mass_matrix(wl.mesh.nodes, wl.mesh.elements, rho)
    assert rel_err(kval, M_ref.data) < 1e-13


def test_linear_stiffness_matches_oracle():
    wl = W.jittered_box_tet4(n=8, seed=6)
    banner("linear stiffness")
    K = _domain(wl).assemble_stiffness_matrix()
    K_ref = A.linear_stiffness(wl.mesh.nodes, wl.mesh.elements, wl.mu, wl.lam)
    assert np.array_equal(K.indptr, K_ref.indptr) and np.array_equal(K.indices, K_ref.indices)
    assert rel_err(K.data, K_ref.data) < 1e-12


def test_element_level_mass_and_linear_stiffness(regular_tet_coords):
    from oracle import fem
    from psierra_amd.element import GpuTet4TotalLagrangian
    from psierra_amd.material import NeoHookeanField

    el = GpuTet4TotalLagrangian(NeoHookeanField(np.array([1e3]), np.array([5e3]), mass_density=2.0))
    banner("element mass / linear stiffness")
    me = el.compute_mass_matrix(1.0 / 6.0)
    ref = np.zeros((12, 12))
    for i in range(4):
        for j in range(4):
            ref[3 * i:3 * i + 3, 3 * j:3 * j + 3] = (2.0 if i == j else 1.0) * (2.0 / 6.0 / 20.0) * np.eye(3)
    assert np.allclose(me, ref, rtol=1e-15, atol=0)
    k = el.compute_linear_stiffness_matrix(regular_tet_coords, elem_index=0)
    V, H = fem.tet4_volume_and_gradients(regular_tet_coords)
    BL = fem.linear_B(H)
    assert rel_err(k, BL.T @ fem.hookean_C(5e3, 1e3) @ BL * V) < 1e-12
