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
