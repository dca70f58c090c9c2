"""The oracle's ECSW training matrix, pinned by properties of the assembly.

oracle.ecsw.training_matrix restates ECSWTrainingMatrix.build
(src/psierra/hyperreduction/ecsw.py:171-365).  Checks that do not retype
its formula:
* per snapshot, the element columns sum to Phi^T f_int(u_k) with f_int from
  the oracle's global assembly (pinned in test_oracle.py);
* one column equals Phi_e^T f_e from the per-element routine of the
  reference's element (tet4_force_and_tangent), so a transposed or shifted
  block fails;
* the stiffness rows, unscaled and unpacked from numpy's triu order, sum to
  Phi^T K0 Phi with K0 the assembled tangent at u = 0, and the scaling makes
  the two blocks' RMS row norms equal (times the weight).
"""
from __future__ import annotations

import numpy as np

from oracle import assembly as A
from oracle import ecsw as OE
from oracle import fem
from psierra_amd import workloads as W


def _case(r=8, m=3, seed=21):
    wl = W.cantilever_tet4()
    rng = np.random.default_rng(seed)
    n = wl.mesh.n_dofs
    basis = np.linalg.qr(rng.standard_normal((n, r)))[0]
    snaps = 0.002 * rng.standard_normal((n, m)) + 0.5 * wl.u.reshape(-1, 1)
    return wl, basis, snaps


def test_force_block_sums_to_projected_global_force():
    wl, basis, snaps = _case()
    r, m = basis.shape[1], snaps.shape[1]
    Amat, b = OE.training_matrix(wl.mesh.nodes, wl.mesh.elements, basis, snaps, wl.mu, wl.lam)
    assert Amat.shape == (m * r, wl.mesh.n_elements)
    assert np.allclose(b, Amat.sum(axis=1))
    for k in range(m):
        f, _ = A.assemble(wl.mesh.nodes, wl.mesh.elements, snaps[:, k].reshape(-1, 3), wl.mu, wl.lam, want_k=False)
        ref = basis.T @ f
        assert np.linalg.norm(b[k * r:(k + 1) * r] - ref) <= 1e-12 * np.linalg.norm(ref)


def test_one_column_matches_element_routine():
    wl, basis, snaps = _case()
    r = basis.shape[1]
    Amat, _ = OE.training_matrix(wl.mesh.nodes, wl.mesh.elements, basis, snaps, wl.mu, wl.lam)
    for e in (0, 777, wl.mesh.n_elements - 1):
        conn = wl.mesh.elements[e]
        dofs = (3 * conn[:, None] + np.arange(3)).ravel()
        for k in range(snaps.shape[1]):
            ue = snaps[:, k].reshape(-1, 3)[conn]
            _, fe = fem.tet4_force_and_tangent(wl.mesh.nodes[conn], ue, fem.nh_material(wl.mu[e], wl.lam[e]))
            c = basis[dofs].T @ fe
            assert np.allclose(Amat[k * r:(k + 1) * r, e], c, rtol=1e-12, atol=1e-12 * np.abs(c).max())


def test_stiffness_block_is_projected_linear_stiffness():
    wl, basis, snaps = _case()
    r, m = basis.shape[1], snaps.shape[1]
    weight = 0.7
    Amat, _ = OE.training_matrix(wl.mesh.nodes, wl.mesh.elements, basis, snaps, wl.mu, wl.lam,
                                 include_stiffness=True, stiffness_weight=weight)
    nt = r * (r + 1) // 2
    assert Amat.shape == (m * r + nt, wl.mesh.n_elements)
    Af, As = Amat[:m * r], Amat[m * r:]
    rms_f = np.linalg.norm(Af) / np.sqrt(m * r)
    rms_s = np.linalg.norm(As) / np.sqrt(nt)
    assert abs(rms_s - weight * rms_f) <= 1e-12 * rms_f
    _, K0 = A.assemble(wl.mesh.nodes, wl.mesh.elements, np.zeros_like(wl.u), wl.mu, wl.lam)
    P = basis.T @ (K0 @ basis)
    iu = np.triu_indices(r)
    got = np.zeros((r, r))
    got[iu] = As.sum(axis=1)
    scale = got[0, 0] / P[0, 0]
    assert np.allclose(got[iu] / scale, P[iu], rtol=1e-10, atol=1e-10 * np.abs(P).max())
