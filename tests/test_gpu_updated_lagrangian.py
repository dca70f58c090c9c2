"""GPU Updated Lagrangian elements against the oracle's UL restatement.

GpuTet4UpdatedLagrangian runs the Total Lagrangian kernels; this checks the
assembled f and K against oracle.fem.tet4_ul_force_and_tangent (the
restatement of src/psierra/element/tet4_updated_lagrangian.py:31-109)
summed with add.at / coo.tocsr, within the FP64 tolerances (1e-11 / 1e-12),
and that an inversion raises the UL element's own message.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import assembly as A
from oracle import fem
from psierra_amd import workloads as W
from tests.gpu_helpers import banner, rel_err

pytestmark = pytest.mark.gpu


def _oracle_ul(wl):
    E = wl.mesh.n_elements
    ke = np.empty((E, 12, 12))
    fe = np.empty((E, 12))
    for e in range(E):
        c = wl.mesh.elements[e]
        ke[e], fe[e] = fem.tet4_ul_force_and_tangent(wl.mesh.nodes[c], wl.u[c], fem.nh_material(wl.mu[e], wl.lam[e]))
    dofs = A.compute_element_dofs(wl.mesh.elements)
    rows, cols = A.create_coo_sparsity_indices(dofs)
    n = wl.mesh.n_dofs
    return A.assemble_vector(fe, dofs, n), A.assemble_sparse_matrix(ke, rows, cols, n)


def _domain(wl):
    from psierra_amd.domain import SolidDomain
    from psierra_amd.element import GpuTet4UpdatedLagrangian
    from psierra_amd.material import NeoHookeanField

    mat = NeoHookeanField(wl.mu, wl.lam)
    return SolidDomain(wl.mesh, mat, GpuTet4UpdatedLagrangian(mat))


def test_ul_assembly_matches_oracle():
    wl = W.jittered_box_tet4(n=6, seed=21)
    f_ref, K_ref = _oracle_ul(wl)
    dom = _domain(wl)
    dom.update_state(u=wl.u.ravel())
    banner("updated lagrangian tet4 assembly")
    f = dom.assemble_internal_forces()
    K = dom.assemble_tangent_stiffness_matrix()
    assert np.array_equal(K.indptr, K_ref.indptr) and np.array_equal(K.indices, K_ref.indices)
    assert rel_err(f, f_ref) < 1e-11
    assert rel_err(K.data, K_ref.data) < 1e-12


def test_ul_inversion_message():
    wl = W.cantilever_tet4()
    u = wl.u.copy()
    conn = wl.mesh.elements[3]
    X = wl.mesh.nodes[conn]
    u[conn[0]] = 2.5 * (X[1:].mean(axis=0) - X[0])
    dom = _domain(wl)
    dom.update_state(u=u.ravel())
    banner("updated lagrangian inversion")
    with pytest.raises(ValueError, match=r"Element inversion detected in Tet4UpdatedLagrangian: det\(F\) = "):
        dom.assemble_internal_forces()
