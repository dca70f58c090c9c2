"""FP64 ECSW online evaluation restated (test infrastructure only).

HyperReducedECSWModel.evaluate_forces_and_stiffness
(src/psierra/hyperreduction/model.py:254-318): for each RID element
u_e = Phi_e u_r, (k_e, f_e) from the Tet4 TL Neo-Hookean routine,
f_r += w Phi_e^T f_e, K_r += w Phi_e^T (k_e Phi_e), then K_r = (K_r + K_r^T)/2.
The element routine is vectorised over the RID (oracle.fem.tet4_batch).
"""
from __future__ import annotations

import numpy as np

from . import fem


def evaluate_forces_and_stiffness(nodes, elements, basis, rid, weights, mu, lam, u_r):
    conn = elements[rid]
    dofs = (3 * conn[:, :, None] + np.arange(3)).reshape(len(rid), 12)
    Phi = basis[dofs]  # (m, 12, r)
    ue = np.einsum("mdr,r->md", Phi, u_r).reshape(-1, 4, 3)
    m = len(rid)
    mu = np.broadcast_to(mu, (m,)).astype(float)
    lam = np.broadcast_to(lam, (m,)).astype(float)
    k, f = fem.tet4_batch(nodes[conn], ue, mu, lam)
    f_r = np.einsum("m,mdr,md->r", weights, Phi, f)
    K_r = np.einsum("m,mdr,mde,mes->rs", weights, Phi, k, Phi, optimize=True)
    return f_r, 0.5 * (K_r + K_r.T)


def training_matrix(nodes, elements, basis, snapshots, mu, lam, include_stiffness=False,
                    stiffness_weight=None):
    """ECSWTrainingMatrix.build, single-domain path
    (src/psierra/hyperreduction/ecsw.py:171-365, the loops at :319-365).

    A[k r + i, e] = (Phi_e^T f_e(u_k))_i for snapshot k (columns of
    ``snapshots``, shape (n_dofs, m)) and element e; with include_stiffness
    the rows of the upper triangle (numpy triu_indices order) of
    Phi_e^T K0_e Phi_e, K0_e the element tangent at u = 0, scaled by
    (s_force / s_stiff) * weight where s = ||block||_F / sqrt(rows of block)
    (:350-357, skipped when either is <= 1e-12); b = A summed over elements.
    """
    snaps = np.asarray(snapshots, dtype=np.float64)
    if snaps.ndim == 1:
        snaps = snaps.reshape(-1, 1)
    n_dofs, m = snaps.shape
    r = basis.shape[1]
    E = elements.shape[0]
    dofs = (3 * elements[:, :, None] + np.arange(3)).reshape(E, 12)
    Phi = basis[dofs]  # (E, 12, r)
    mu = np.broadcast_to(mu, (E,)).astype(float)
    lam = np.broadcast_to(lam, (E,)).astype(float)
    Xe = nodes[elements]
    A = np.zeros((m * r, E))
    for k in range(m):
        ue = snaps[:, k].reshape(-1, 3)[elements]
        _, f = fem.tet4_batch(Xe, ue, mu, lam, want_k=False)
        A[k * r:(k + 1) * r] = np.einsum("edr,ed->re", Phi, f)
    if include_stiffness:
        k0, _ = fem.tet4_batch(Xe, np.zeros_like(Xe), mu, lam)
        Kr = np.einsum("xdr,xdc,xcs->xrs", Phi, k0, Phi, optimize=True)
        iu = np.triu_indices(r)
        A_stiff = Kr[:, iu[0], iu[1]].T.copy()
        s_force = float(np.linalg.norm(A)) / np.sqrt(max(A.shape[0], 1))
        s_stiff = float(np.linalg.norm(A_stiff)) / np.sqrt(max(len(iu[0]), 1))
        if s_force > 1e-12 and s_stiff > 1e-12:
            weight = float(stiffness_weight) if stiffness_weight is not None else 1.0
            A_stiff *= (s_force / s_stiff) * weight
        A = np.vstack([A, A_stiff])
    return A, A.sum(axis=1)
