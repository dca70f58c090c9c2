"""FP64 restatement of pSierra's Tet4Fbar domain assembly (test infrastructure only).

src/psierra/element/tet4_fbar.py:
* ``setup`` -- ``_setup_cache`` (:53-109): reference volumes and gradients per
  element, nodal reference volumes V0_a = sum_{e in a} V_e / 4;
* ``smoothed_jacobians`` -- the J loop and nodal averaging shared by
  ``prepare_for_assembly`` / ``assemble_internal_forces`` /
  ``assemble_tangent_stiffness_matrix`` (:226-255, :285-311, :369-395):
  F = I + u^T H, J_node = sum_{e in a} V_e J_e / 4 / V0_a,
  J_bar_e = mean of J_node over the element's nodes;
* ``internal_forces`` -- :264-330: F_bar = (J_bar/J)^(1/3) F, S_bar = S(F_bar),
  B_bar = ratio^2 (B - m b_vol^T) + m b_vol^T with b_vol = (B_0+B_1+B_2)/3,
  f_e = B_bar^T S_bar V;
* ``tangent_stiffness`` -- :332-459: K_dev from
  (ratio^2 B_dev)^T (C(F_bar) - K m m^T) (ratio^2 B_dev) V + G^T (I x S_bar) G V,
  plus the nodal volumetric part K_vol = sum_a K V0_a g_a g_a^T with
  g_a = sum_{e in a} (V_e J_e / 4 / V0_a) div_e and div_e = B_0+B_1+B_2;
  the result is scipy's CSR sum K_dev + K_vol (which drops exact-zero sums).

The material is the compressible Neo-Hookean model with per-element (mu_e,
lambda_e) (the element calls ``compute_stress_and_tangent(F_bar,
elem_index=e)`` when the material supports element indexing, :303-306);
``K_bulk`` is the single bulk modulus of ``_get_bulk_modulus`` (:112-118).

Pinning (tests/test_oracle_fbar.py): at u = 0 the assembled tangent is the
exact Jacobian of the assembled force (central finite differences, shown in
the test's docstring), under a homogeneous deformation J_bar = J and f equals
the Total Lagrangian force of ``oracle.assembly.assemble``; rigid motions give
f = 0; K is symmetric; the vectorised K_vol equals the reference's per-node
block loop.  K_dev away from u = 0 and away from homogeneous deformations is
pinned only through its building blocks (B, G and the Neo-Hookean C and S,
each pinned in the other oracle tests) and through the GPU engine's
independent closed form: "parity partially unpinned" for that term.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as spsp

from . import fem
from .assembly import assemble_sparse_matrix, assemble_vector, compute_element_dofs, create_coo_sparsity_indices

INVERSION_MSG = "Element inversion detected at element {}: det(F) = {:.4e} <= 0."
_M = np.array([1.0, 1.0, 1.0, 0.0, 0.0, 0.0])


def setup(nodes, elements):
    """(V (E,), H (E,4,3), V0 (N,)) as ``_setup_cache`` computes them (:66-80)."""
    nodes = np.asarray(nodes, dtype=np.float64)
    elements = np.asarray(elements, dtype=np.int64)
    Xe = nodes[elements]
    A = np.ones((len(elements), 4, 4))
    A[:, :, 1:] = Xe
    detA = np.linalg.det(A)
    if np.any(detA <= 0.0):
        raise ValueError("Negative or zero element volume detected")
    V = detA / 6.0
    H = np.linalg.inv(A)[:, 1:, :].transpose(0, 2, 1)
    V0 = np.bincount(elements.ravel(), weights=np.repeat(0.25 * V, 4), minlength=len(nodes))
    return V, H, V0


def smoothed_jacobians(elements, u, V, H, V0):
    """(F (E,3,3), J (E,), J_node (N,), J_bar (E,)) -- tet4_fbar.py:285-311."""
    elements = np.asarray(elements, dtype=np.int64)
    ue = np.asarray(u, dtype=np.float64).reshape(-1, 3)[elements]
    F = np.identity(3) + np.einsum("eai,eaj->eij", ue, H)
    J = np.linalg.det(F)
    bad = np.nonzero(J <= 0.0)[0]
    if bad.size:
        raise ValueError(INVERSION_MSG.format(int(bad[0]), J[bad[0]]))
    v_node = np.bincount(elements.ravel(), weights=np.repeat(0.25 * V * J, 4), minlength=len(V0))
    with np.errstate(invalid="ignore", divide="ignore"):
        J_node = v_node / V0
    J_bar = np.mean(J_node[elements], axis=1)
    return F, J, J_node, J_bar


def _element_terms(F, J, J_bar, H, mu, lam):
    """ratio, S_bar, C_tan(F_bar), B (nonlinear, actual F), G, b_vol."""
    ratio = (J_bar / J) ** (1.0 / 3.0)
    F_bar = ratio[:, None, None] * F
    Cv, S = fem.neohookean_batch(F_bar, mu, fem.bulk_from_lame(lam, mu))
    B, G = fem._batched_B_G(H, F - np.identity(3))
    b_vol = (B[:, 0] + B[:, 1] + B[:, 2]) / 3.0
    return ratio, S, Cv, B, G, b_vol


def internal_forces(nodes, elements, u, mu, lam):
    """Tet4Fbar.assemble_internal_forces (tet4_fbar.py:264-330) -> f (3N,)."""
    elements = np.asarray(elements, dtype=np.int64)
    V, H, V0 = setup(nodes, elements)
    F, J, _, J_bar = smoothed_jacobians(elements, u, V, H, V0)
    mu = np.broadcast_to(np.asarray(mu, dtype=np.float64), J.shape)
    lam = np.broadcast_to(np.asarray(lam, dtype=np.float64), J.shape)
    ratio, S, _, B, _, b_vol = _element_terms(F, J, J_bar, H, mu, lam)
    B_dev = B - _M[None, :, None] * b_vol[:, None, :]
    B_bar = (ratio**2)[:, None, None] * B_dev + _M[None, :, None] * b_vol[:, None, :]
    Sv = S[:, fem._VI, fem._VJ]
    f_local = np.einsum("evk,ev->ek", B_bar, Sv) * V[:, None]
    return assemble_vector(f_local, compute_element_dofs(elements), 3 * len(V0))


def volumetric_gradients(elements, V, J, V0, div):
    """g_a as a sparse (N, 3N) matrix: row a = sum_{e in a} w_{e,a} div_e
    placed at the element's dofs, w_{e,a} = V_e J_e / 4 / V0_a (:433-444)."""
    elements = np.asarray(elements, dtype=np.int64)
    E = len(elements)
    N = len(V0)
    dofs = compute_element_dofs(elements)                      # (E,12)
    rows = np.repeat(elements, 12, axis=1)                     # (E, 4*12): node a of incidence
    cols = np.tile(dofs, (1, 4))
    w = 0.25 * V * J                                            # (E,)
    vals = (w[:, None] / V0[elements])[:, :, None] * div[:, None, :]   # (E,4,12)
    return spsp.coo_matrix((vals.reshape(E, 48).ravel(), (rows.ravel(), cols.ravel())), shape=(N, 3 * N)).tocsr()


def tangent_stiffness(nodes, elements, u, mu, lam, K_bulk):
    """Tet4Fbar.assemble_tangent_stiffness_matrix (tet4_fbar.py:332-459) -> CSR."""
    elements = np.asarray(elements, dtype=np.int64)
    V, H, V0 = setup(nodes, elements)
    F, J, _, J_bar = smoothed_jacobians(elements, u, V, H, V0)
    mu = np.broadcast_to(np.asarray(mu, dtype=np.float64), J.shape)
    lam = np.broadcast_to(np.asarray(lam, dtype=np.float64), J.shape)
    ratio, S, Cv, B, G, b_vol = _element_terms(F, J, J_bar, H, mu, lam)
    C_dev = Cv - K_bulk * np.outer(_M, _M)[None]
    B_bar_dev = (ratio**2)[:, None, None] * (B - _M[None, :, None] * b_vol[:, None, :])
    k = np.einsum("evk,evw,ewl->ekl", B_bar_dev, C_dev, B_bar_dev, optimize=True)
    Shat = np.zeros((len(J), 9, 9))
    for c in range(3):
        Shat[:, 3 * c : 3 * c + 3, 3 * c : 3 * c + 3] = S
    k += np.einsum("evk,evw,ewl->ekl", G, Shat, G, optimize=True)
    k *= V[:, None, None]
    dofs = compute_element_dofs(elements)
    rows, cols = create_coo_sparsity_indices(dofs)
    n = 3 * len(V0)
    K_dev = assemble_sparse_matrix(k, rows, cols, n)
    return K_dev + _volumetric(elements, V, J, V0, 3.0 * b_vol, K_bulk)


def _volumetric(elements, V, J, V0, div, K_bulk):
    Gm = volumetric_gradients(elements, V, J, V0, div)
    K_vol = (Gm.T @ spsp.diags(K_bulk * V0) @ Gm).tocsr()
    K_vol.sort_indices()
    return K_vol


def volumetric_stiffness(nodes, elements, u, K_bulk):
    """K_vol alone (CSR), vectorised as G^T diag(K V0) G."""
    elements = np.asarray(elements, dtype=np.int64)
    V, H, V0 = setup(nodes, elements)
    F, J, _, _ = smoothed_jacobians(elements, u, V, H, V0)
    B, _ = fem._batched_B_G(H, F - np.identity(3))
    return _volumetric(elements, V, J, V0, B[:, 0] + B[:, 1] + B[:, 2], K_bulk)


def volumetric_blocks_loop(nodes, elements, u, K_bulk):
    """K_vol by the reference's per-node loop (tet4_fbar.py:432-452), dense."""
    elements = np.asarray(elements, dtype=np.int64)
    V, H, V0 = setup(nodes, elements)
    F, J, _, _ = smoothed_jacobians(elements, u, V, H, V0)
    B, _ = fem._batched_B_G(H, F - np.identity(3))
    div = B[:, 0] + B[:, 1] + B[:, 2]
    dofs = compute_element_dofs(elements)
    n = 3 * len(V0)
    K = np.zeros((n, n))
    for a in range(len(V0)):
        elems_a = np.nonzero(np.any(elements == a, axis=1))[0]
        if elems_a.size == 0:
            continue
        nodes_a = np.unique(elements[elems_a].ravel())
        dofs_a = np.repeat(nodes_a * 3, 3) + np.tile(np.arange(3), len(nodes_a))
        pos = {d: i for i, d in enumerate(dofs_a)}
        g_a = np.zeros(len(dofs_a))
        for e in elems_a:
            w_e = 0.25 * V[e] * J[e] / V0[a]
            for i, d in enumerate(dofs[e]):
                g_a[pos[d]] += w_e * div[e, i]
        K[np.ix_(dofs_a, dofs_a)] += K_bulk * V0[a] * np.outer(g_a, g_a)
    return K


def distance2_pattern(elements, n_nodes):
    """Structural CSR pattern of K_dev + K_vol: dof pairs of nodes that share an
    element or a common neighbouring node (the nodal patches of :86-109)."""
    elements = np.asarray(elements, dtype=np.int64)
    dofs = compute_element_dofs(elements)
    rows, cols = create_coo_sparsity_indices(dofs)
    n = 3 * n_nodes
    A = spsp.coo_matrix((np.ones(rows.size), (rows // 3, cols // 3)), shape=(n_nodes, n_nodes)).tocsr()
    A.data[:] = 1.0
    A2 = (A @ A).tocsr()
    A2.data[:] = 1.0
    P = spsp.kron(A2, np.ones((3, 3)), format="csr")
    P.sort_indices()
    return P.indptr, P.indices
