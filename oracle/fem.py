"""FP64 restatement of pSierra's Tet4/Tet10 Total Lagrangian Neo-Hookean path.

Test infrastructure (see ``oracle/__init__.py``).  Two forms are provided:

* per-element functions that mirror the reference call by call (4x4
  ``det``/``inv`` geometry, dense B, B_NL, G, einsum C4, Voigt copy loop);
  these are used for single-element checks and the faithful CPU baseline;
* ``*_batch`` functions doing the same arithmetic vectorised over a chunk of
  elements with numpy's batched ``det``/``inv``/``einsum``.
"""
from __future__ import annotations

import numpy as np

# src/psierra/utility/tensor.py:5-12
VOIGT_MAP_3D = {0: (0, 0), 1: (1, 1), 2: (2, 2), 3: (0, 1), 4: (1, 2), 5: (2, 0)}
_VI = np.array([VOIGT_MAP_3D[a][0] for a in range(6)])
_VJ = np.array([VOIGT_MAP_3D[a][1] for a in range(6)])

INVERSION_MSG = "Element inversion detected: det(F) = {:.4e} <= 0."


def bulk_from_lame(lam, mu):
    """K = lambda + 2 mu / 3 (src/psierra/material/parameters.py:135)."""
    return lam + (2.0 / 3.0) * mu


def lame_from_young_poisson(E, nu):
    """src/psierra/material/parameters.py:66-68."""
    mu = E / (2.0 * (1.0 + nu))
    lam = (E * nu) / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return lam, mu


def tensor_to_voigt_stress(S):
    """src/psierra/utility/tensor.py:15-38."""
    return np.array([S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[1, 2], S[2, 0]], dtype=np.float64)


# ---------------------------------------------------------------------------
# Material (src/psierra/material/neohookean.py, hookean.py)
# ---------------------------------------------------------------------------

def hookean_C(lam, mu):
    """Small-strain isotropic C (src/psierra/material/hookean.py:30-60)."""
    C = np.zeros((6, 6))
    C[0, 0] = C[1, 1] = C[2, 2] = lam + 2.0 * mu
    C[0, 1] = C[1, 0] = C[0, 2] = C[2, 0] = C[1, 2] = C[2, 1] = lam
    C[3, 3] = C[4, 4] = C[5, 5] = mu
    return C


def hookean_stress_and_tangent(F, lam, mu):
    """Hookean.compute_stress_and_tangent (src/psierra/material/hookean.py:62-78)."""
    E = 0.5 * (F.T @ F - np.identity(3))
    S = lam * np.trace(E) * np.identity(3) + 2.0 * mu * E
    return hookean_C(lam, mu), S


def neohookean_stress_and_tangent(F, mu, K):
    """NeoHookean.compute_stress_and_tangent (src/psierra/material/neohookean.py:55-122)."""
    J = float(np.linalg.det(F))
    if J <= 0.0:
        raise ValueError(INVERSION_MSG.format(J))
    C = F.T @ F
    C_inv = np.linalg.inv(C)
    I1 = float(np.trace(C))
    I3 = np.identity(3)
    factor = mu * (J ** (-2.0 / 3.0))
    S = factor * (I3 - (1.0 / 3.0) * I1 * C_inv) + K * np.log(J) * C_inv
    C_inv_outer = np.einsum("ij,kl->ijkl", C_inv, C_inv)
    I_Cinv = 0.5 * (np.einsum("ik,jl->ijkl", C_inv, C_inv) + np.einsum("il,jk->ijkl", C_inv, C_inv))
    C4_vol = K * C_inv_outer - 2.0 * K * np.log(J) * I_Cinv
    C_inv_I = np.einsum("ij,kl->ijkl", C_inv, I3) + np.einsum("ij,kl->ijkl", I3, C_inv)
    C4_iso = (
        -(2.0 / 3.0) * factor * C_inv_I
        + (2.0 / 9.0) * factor * I1 * C_inv_outer
        + (2.0 / 3.0) * factor * I1 * I_Cinv
    )
    C4 = C4_vol + C4_iso
    C_voigt = np.zeros((6, 6))
    for a, (i, j) in VOIGT_MAP_3D.items():
        for b, (k, l) in VOIGT_MAP_3D.items():
            C_voigt[a, b] = C4[i, j, k, l]
    return C_voigt, S


def neohookean_energy(F, mu, K):
    """Psi = mu/2 (J^-2/3 I1 - 3) + K/2 (ln J)^2 (src/psierra/material/neohookean.py:13-15)."""
    J = np.linalg.det(F)
    I1 = np.trace(F.T @ F)
    return 0.5 * mu * (J ** (-2.0 / 3.0) * I1 - 3.0) + 0.5 * K * np.log(J) ** 2


def _voigt_of(C4):
    Cv = np.zeros((6, 6))
    for a, (i, j) in VOIGT_MAP_3D.items():
        for b, (k, l) in VOIGT_MAP_3D.items():
            Cv[a, b] = C4[i, j, k, l]
    return Cv


def mooney_rivlin_stress_and_tangent(F, c10, c01, K):
    """MooneyRivlin.compute_stress_and_tangent
    (src/psierra/material/mooney_rivlin.py:69-152)."""
    J = float(np.linalg.det(F))
    if J <= 0.0:
        raise ValueError(INVERSION_MSG.format(J))
    C = F.T @ F
    Ci = np.linalg.inv(C)
    I1 = float(np.trace(C))
    I2 = 0.5 * (I1 ** 2 - float(np.trace(C @ C)))
    I3 = np.identity(3)
    f1 = 2.0 * c10 * J ** (-2.0 / 3.0)
    f2 = 2.0 * c01 * J ** (-4.0 / 3.0)
    S = f1 * (I3 - I1 / 3.0 * Ci) + f2 * (I1 * I3 - C - 2.0 / 3.0 * I2 * Ci) + K * np.log(J) * Ci
    outer = np.einsum("ij,kl->ijkl", Ci, Ci)
    ICi = 0.5 * (np.einsum("ik,jl->ijkl", Ci, Ci) + np.einsum("il,jk->ijkl", Ci, Ci))
    C4 = K * outer - 2.0 * K * np.log(J) * ICi
    CiI = np.einsum("ij,kl->ijkl", Ci, I3) + np.einsum("ij,kl->ijkl", I3, Ci)
    C4 += -2.0 / 3.0 * f1 * CiI + 2.0 / 9.0 * f1 * I1 * outer + 2.0 / 3.0 * f1 * I1 * ICi
    Iout = np.einsum("ij,kl->ijkl", I3, I3)
    Isym = 0.5 * (np.einsum("ik,jl->ijkl", I3, I3) + np.einsum("il,jk->ijkl", I3, I3))
    B2 = I1 * I3 - C
    CiB2 = np.einsum("ij,kl->ijkl", Ci, B2) + np.einsum("ij,kl->ijkl", B2, Ci)
    C4 += (c01 * J ** (-4.0 / 3.0)) * (4.0 * (Iout - Isym) - 8.0 / 3.0 * CiB2 + 16.0 / 9.0 * I2 * outer
                                       + 8.0 / 3.0 * I2 * ICi)
    return _voigt_of(C4), S


def mooney_rivlin_energy(F, c10, c01, K):
    """Psi = c10 (I1bar - 3) + c01 (I2bar - 3) + K/2 (ln J)^2 (mooney_rivlin.py:15-19)."""
    J = np.linalg.det(F)
    C = F.T @ F
    I1 = np.trace(C)
    I2 = 0.5 * (I1 ** 2 - np.trace(C @ C))
    return c10 * (J ** (-2.0 / 3.0) * I1 - 3.0) + c01 * (J ** (-4.0 / 3.0) * I2 - 3.0) + 0.5 * K * np.log(J) ** 2


def yeoh_stress_and_tangent(F, c10, c20, c30, K):
    """Yeoh.compute_stress_and_tangent (src/psierra/material/yeoh.py:69-148)."""
    J = float(np.linalg.det(F))
    if J <= 0.0:
        raise ValueError(INVERSION_MSG.format(J))
    C = F.T @ F
    Ci = np.linalg.inv(C)
    I1 = float(np.trace(C))
    Jp = J ** (-2.0 / 3.0)
    d = Jp * I1 - 3.0
    psi1 = c10 + 2.0 * c20 * d + 3.0 * c30 * d ** 2
    psi11 = 2.0 * c20 + 6.0 * c30 * d
    I3 = np.identity(3)
    A1 = 2.0 * Jp * (I3 - I1 / 3.0 * Ci)
    S = psi1 * A1 + K * np.log(J) * Ci
    outer = np.einsum("ij,kl->ijkl", Ci, Ci)
    ICi = 0.5 * (np.einsum("ik,jl->ijkl", Ci, Ci) + np.einsum("il,jk->ijkl", Ci, Ci))
    C4 = K * outer - 2.0 * K * np.log(J) * ICi
    CiI = np.einsum("ij,kl->ijkl", Ci, I3) + np.einsum("ij,kl->ijkl", I3, Ci)
    C4 += psi1 * Jp * (-4.0 / 3.0 * CiI + 4.0 / 9.0 * I1 * outer + 4.0 / 3.0 * I1 * ICi)
    C4 += psi11 * np.einsum("ij,kl->ijkl", A1, A1)
    return _voigt_of(C4), S


def yeoh_energy(F, c10, c20, c30, K):
    """Psi = sum_i c_i0 (I1bar - 3)^i + K/2 (ln J)^2 (yeoh.py:13-16)."""
    J = np.linalg.det(F)
    d = J ** (-2.0 / 3.0) * np.trace(F.T @ F) - 3.0
    return c10 * d + c20 * d ** 2 + c30 * d ** 3 + 0.5 * K * np.log(J) ** 2


def material_fn(model, params):
    """(C_voigt, S) = material(F) for model 0 = Neo-Hookean (mu, lambda),
    1 = Mooney-Rivlin (c10, c01, K), 2 = Yeoh (c10, c20, c30, K)."""
    if model == 0:
        return nh_material(*params)
    if model == 1:
        return lambda F: mooney_rivlin_stress_and_tangent(F, *params)
    return lambda F: yeoh_stress_and_tangent(F, *params)


def neohookean_batch(F, mu, K):
    """Vectorised NeoHookean over F (m,3,3), mu/K (m,) -> (C_voigt (m,6,6), S (m,3,3))."""
    J = np.linalg.det(F)
    bad = np.nonzero(~(J > 0.0) & ~np.isnan(J))[0]
    if bad.size:
        raise ValueError(INVERSION_MSG.format(J[bad[0]]))
    C = np.einsum("eki,ekj->eij", F, F)
    Ci = np.linalg.inv(C)
    I1 = np.trace(C, axis1=1, axis2=2)
    lnJ = np.log(J)
    fac = mu * J ** (-2.0 / 3.0)
    I3 = np.identity(3)
    S = fac[:, None, None] * (I3 - (I1 / 3.0)[:, None, None] * Ci) + (K * lnJ)[:, None, None] * Ci
    outer = np.einsum("eij,ekl->eijkl", Ci, Ci)
    ICi = 0.5 * (np.einsum("eik,ejl->eijkl", Ci, Ci) + np.einsum("eil,ejk->eijkl", Ci, Ci))
    CiI = np.einsum("eij,kl->eijkl", Ci, I3) + np.einsum("ij,ekl->eijkl", I3, Ci)
    c_outer = K + (2.0 / 9.0) * fac * I1
    c_sym = -2.0 * K * lnJ + (2.0 / 3.0) * fac * I1
    c_mix = -(2.0 / 3.0) * fac
    C4 = (
        c_outer[:, None, None, None, None] * outer
        + c_sym[:, None, None, None, None] * ICi
        + c_mix[:, None, None, None, None] * CiI
    )
    Cv = C4[:, _VI[:, None], _VJ[:, None], _VI[None, :], _VJ[None, :]]
    return Cv, S


# ---------------------------------------------------------------------------
# Tet4 (src/psierra/element/tet4_base.py, tet4_total_lagrangian.py)
# ---------------------------------------------------------------------------

def tet4_volume_and_gradients(X):
    """Tet4Base.compute_volume_and_gradients (src/psierra/element/tet4_base.py:34-67)."""
    A = np.ones((4, 4))
    A[:, 1:] = X
    detA = np.linalg.det(A)
    if detA <= 0.0:
        raise ValueError(
            f"Negative or zero element volume detected (det(A) = {detA}). "
            "Check element vertex ordering."
        )
    return detA / 6.0, np.linalg.inv(A)[1:, :].T


def linear_B(H):
    """compute_linear_B_matrix (tet4_base.py:73-95; tet10_base.py:249-276)."""
    n = H.shape[0]
    BL = np.zeros((6, 3 * n))
    for i in range(n):
        b, c, d = H[i]
        col = 3 * i
        BL[0, col] = b
        BL[1, col + 1] = c
        BL[2, col + 2] = d
        BL[3, col] = c
        BL[3, col + 1] = b
        BL[4, col + 1] = d
        BL[4, col + 2] = c
        BL[5, col] = d
        BL[5, col + 2] = b
    return BL


def geometric_G(H):
    """compute_geometric_matrix (tet4_base.py:97-123; tet10_base.py:278-304)."""
    n = H.shape[0]
    G = np.zeros((9, 3 * n))
    for i in range(n):
        for c in range(3):
            G[3 * c : 3 * c + 3, 3 * i + c] = H[i]
    return G


def nonlinear_B(H, u):
    """compute_nonlinear_B_matrix (tet4_total_lagrangian.py:24-75; tet10_total_lagrangian.py:22-61)."""
    BL = linear_B(H)
    n = H.shape[0]
    BNL = np.zeros((6, 3 * n))
    l = u.T @ H
    du, dv, dw = l[0], l[1], l[2]
    for i in range(n):
        b, c, d = H[i]
        col = 3 * i
        BNL[0, col : col + 3] = du[0] * b, dv[0] * b, dw[0] * b
        BNL[1, col : col + 3] = du[1] * c, dv[1] * c, dw[1] * c
        BNL[2, col : col + 3] = du[2] * d, dv[2] * d, dw[2] * d
        BNL[3, col : col + 3] = du[0] * c + du[1] * b, dv[0] * c + dv[1] * b, dw[0] * c + dw[1] * b
        BNL[4, col : col + 3] = du[1] * d + du[2] * c, dv[1] * d + dv[2] * c, dw[1] * d + dw[2] * c
        BNL[5, col : col + 3] = du[0] * d + du[2] * b, dv[0] * d + dv[2] * b, dw[0] * d + dw[2] * b
    return BL + BNL


def tet4_force_and_tangent(X, u, material):
    """Tet4TotalLagrangian.compute_nonlinear_force_and_tangent_stiffness
    (src/psierra/element/tet4_total_lagrangian.py:77-139).

    ``material(F) -> (C_voigt, S)``.  Returns (k_tan (12,12), f_int (12,)).
    """
    V, H = tet4_volume_and_gradients(X)
    F = (X + u).T @ H
    J = np.linalg.det(F)
    if J <= 0.0:
        raise ValueError(INVERSION_MSG.format(J))
    C, S = material(F)
    B = nonlinear_B(H, u)
    G = geometric_G(H)
    k_mat = (B.T @ C @ B) * V
    k_geo = (G.T @ np.kron(np.identity(3), S) @ G) * V
    f = (B.T @ tensor_to_voigt_stress(S)) * V
    return k_mat + k_geo, f


def tensor_to_voigt_strain(E):
    """src/psierra/utility/tensor.py:41-64 (engineering shear)."""
    return np.array([E[0, 0], E[1, 1], E[2, 2], 2 * E[0, 1], 2 * E[1, 2], 2 * E[2, 0]])


def voigt_strain_to_tensor(v):
    """src/psierra/utility/tensor.py:89-111."""
    return np.array([[v[0], 0.5 * v[3], 0.5 * v[5]], [0.5 * v[3], v[1], 0.5 * v[4]], [0.5 * v[5], 0.5 * v[4], v[2]]])


def strain_pushforward(F):
    """compute_strain_pushforward_matrix (src/psierra/utility/tensor.py:113-137):
    column k = Voigt(F^T eps_k F) for the k-th unit Voigt strain."""
    T = np.zeros((6, 6))
    for k in range(6):
        e = np.zeros(6)
        e[k] = 1.0
        T[:, k] = tensor_to_voigt_strain(F.T @ voigt_strain_to_tensor(e) @ F)
    return T


UL_INVERSION_MSG = "Element inversion detected in Tet4UpdatedLagrangian: det(F) = {:.4e} <= 0."


def tet4_ul_force_and_tangent(X, u, material):
    """Tet4UpdatedLagrangian.compute_nonlinear_force_and_tangent_stiffness
    (src/psierra/element/tet4_updated_lagrangian.py:31-109): spatial
    gradients h = H F^-1, v = J V, sigma = F S F^T / J, f = b_L^T sigma v,
    k = b_L^T (T^T C T / J) b_L v + g^T (I3 x sigma) g v."""
    V, H = tet4_volume_and_gradients(X)
    F = (X + u).T @ H
    J = np.linalg.det(F)
    if J <= 0.0:
        raise ValueError(UL_INVERSION_MSG.format(J))
    h = H @ np.linalg.inv(F)
    v = J * V
    C, S = material(F)
    sigma = (F @ S @ F.T) / J
    bL = linear_B(h)
    f = (bL.T @ tensor_to_voigt_stress(sigma)) * v
    T = strain_pushforward(F)
    k_mat = (bL.T @ ((T.T @ C @ T) / J) @ bL) * v
    g = geometric_G(h)
    k_geo = (g.T @ np.kron(np.identity(3), sigma) @ g) * v
    return k_mat + k_geo, f


def nh_material(mu, lam):
    K = bulk_from_lame(lam, mu)
    return lambda F: neohookean_stress_and_tangent(F, mu, K)


def _batched_B_G(H, gu):
    """Batched B = B_L + B_NL(u) (m,6,3n) and G (m,9,3n) from H (m,n,3) and grad u (m,3,3)."""
    m, n, _ = H.shape
    F = gu + np.identity(3)
    # B row v, column 3a+k:  F[k, I] H[a, J] symmetrised over Voigt pair (I, J)
    B = np.empty((m, 6, n, 3))
    for v in range(6):
        I, Jv = VOIGT_MAP_3D[v]
        if I == Jv:
            B[:, v] = H[:, :, I, None] * F[:, None, :, I]
        else:
            B[:, v] = H[:, :, Jv, None] * F[:, None, :, I] + H[:, :, I, None] * F[:, None, :, Jv]
    G = np.zeros((m, 9, n, 3))
    for c in range(3):
        G[:, 3 * c : 3 * c + 3, :, c] = H.transpose(0, 2, 1)
    return B.reshape(m, 6, 3 * n), G.reshape(m, 9, 3 * n)


def tet4_batch(Xe, ue, mu, lam, want_k=True):
    """Vectorised Tet4 TL NH element routine over (m,4,3) arrays.

    Same arithmetic as ``tet4_force_and_tangent`` (4x4 det/inv geometry,
    F = (X+u)^T H, B = B_L + B_NL, k = B^T C B V + G^T (I3 x S) G V,
    f = B^T S_v V), batched.  Returns (k (m,12,12) or None, f (m,12)).
    """
    m = Xe.shape[0]
    A = np.ones((m, 4, 4))
    A[:, :, 1:] = Xe
    detA = np.linalg.det(A)
    bad = np.nonzero(detA <= 0.0)[0]
    if bad.size:
        raise ValueError(
            f"Negative or zero element volume detected (det(A) = {detA[bad[0]]}). "
            "Check element vertex ordering."
        )
    V = detA / 6.0
    H = np.linalg.inv(A)[:, 1:, :].transpose(0, 2, 1)  # (m,4,3)
    F = np.einsum("eai,eaj->eij", Xe + ue, H)
    gu = np.einsum("eai,eaj->eij", ue, H)
    K = bulk_from_lame(lam, mu)
    Cv, S = neohookean_batch(F, mu, K)
    B, G = _batched_B_G(H, gu)
    Sv = S[:, _VI, _VJ]
    f = np.einsum("evk,ev->ek", B, Sv) * V[:, None]
    if not want_k:
        return None, f
    k = np.einsum("evk,evw,ewl->ekl", B, Cv, B, optimize=True)
    Shat = np.zeros((m, 9, 9))
    for c in range(3):
        Shat[:, 3 * c : 3 * c + 3, 3 * c : 3 * c + 3] = S
    k += np.einsum("evk,evw,ewl->ekl", G, Shat, G, optimize=True)
    k *= V[:, None, None]
    return k, f


# ---------------------------------------------------------------------------
# Tet10 (src/psierra/element/tet10_base.py, tet10_total_lagrangian.py,
#        src/psierra/element/quadrature.py)
# ---------------------------------------------------------------------------

def tetrahedral_quadrature_4():
    """4-point rule (src/psierra/element/quadrature.py:36-50)."""
    a = (5.0 - np.sqrt(5.0)) / 20.0
    b = (5.0 + 3.0 * np.sqrt(5.0)) / 20.0
    qp = np.array([[a, a, a], [b, a, a], [a, b, a], [a, a, b]])
    return qp, np.full(4, 1.0 / 24.0)


def tet10_shape_gradients_ref(xi, eta, zeta):
    """Tet10Base.evaluate_shape_gradients_ref (src/psierra/element/tet10_base.py:117-179)."""
    L1, L2, L3 = xi, eta, zeta
    L0 = 1.0 - L1 - L2 - L3
    dN = np.zeros((10, 3))
    dN[0, :] = -(4.0 * L0 - 1.0)
    dN[1, 0] = 4.0 * L1 - 1.0
    dN[2, 1] = 4.0 * L2 - 1.0
    dN[3, 2] = 4.0 * L3 - 1.0
    dN[4, 0] = 4.0 * (L0 - L1)
    dN[4, 1] = -4.0 * L1
    dN[4, 2] = -4.0 * L1
    dN[5, 0] = 4.0 * L2
    dN[5, 1] = 4.0 * L1
    dN[6, 0] = -4.0 * L2
    dN[6, 1] = 4.0 * (L0 - L2)
    dN[6, 2] = -4.0 * L2
    dN[7, 0] = -4.0 * L3
    dN[7, 1] = -4.0 * L3
    dN[7, 2] = 4.0 * (L0 - L3)
    dN[8, 0] = 4.0 * L3
    dN[8, 2] = 4.0 * L1
    dN[9, 1] = 4.0 * L3
    dN[9, 2] = 4.0 * L2
    return dN


def tet10_force_and_tangent(X, u, material):
    """Tet10TotalLagrangian.compute_nonlinear_force_and_tangent_stiffness
    (src/psierra/element/tet10_total_lagrangian.py:63-136)."""
    qp, qw = tetrahedral_quadrature_4()
    x = X + u
    k = np.zeros((30, 30))
    f = np.zeros(30)
    for q in range(4):
        dN = tet10_shape_gradients_ref(*qp[q])
        J = X.T @ dN
        detJ = np.linalg.det(J)
        if detJ <= 0.0:
            raise ValueError(f"Negative or zero Jacobian determinant: det(J) = {detJ:.4e}.")
        H = dN @ np.linalg.inv(J)
        F = x.T @ H
        detF = np.linalg.det(F)
        if detF <= 0.0:
            raise ValueError(INVERSION_MSG.format(detF))
        C, S = material(F)
        B = nonlinear_B(H, u)
        G = geometric_G(H)
        wd = qw[q] * detJ
        k += (B.T @ C @ B + G.T @ np.kron(np.identity(3), S) @ G) * wd
        f += (B.T @ tensor_to_voigt_stress(S)) * wd
    return k, f


def tet10_batch(Xe, ue, mu, lam, want_k=True):
    """Vectorised Tet10 routine over (m,10,3) arrays (same arithmetic as above)."""
    m = Xe.shape[0]
    qp, qw = tetrahedral_quadrature_4()
    K = bulk_from_lame(lam, mu)
    k = np.zeros((m, 30, 30)) if want_k else None
    f = np.zeros((m, 30))
    xe = Xe + ue
    for q in range(4):
        dN = tet10_shape_gradients_ref(*qp[q])
        J = np.einsum("eai,aj->eij", Xe, dN)
        detJ = np.linalg.det(J)
        bad = np.nonzero(detJ <= 0.0)[0]
        if bad.size:
            raise ValueError(f"Negative or zero Jacobian determinant: det(J) = {detJ[bad[0]]:.4e}.")
        H = np.einsum("aj,ejk->eak", dN, np.linalg.inv(J))
        F = np.einsum("eai,eaj->eij", xe, H)
        gu = np.einsum("eai,eaj->eij", ue, H)
        Cv, S = neohookean_batch(F, mu, K)
        B, G = _batched_B_G(H, gu)
        wd = qw[q] * detJ
        f += np.einsum("evk,ev->ek", B, S[:, _VI, _VJ]) * wd[:, None]
        if want_k:
            kq = np.einsum("evk,evw,ewl->ekl", B, Cv, B, optimize=True)
            Shat = np.zeros((m, 9, 9))
            for c in range(3):
                Shat[:, 3 * c : 3 * c + 3, 3 * c : 3 * c + 3] = S
            kq += np.einsum("evk,evw,ewl->ekl", G, Shat, G, optimize=True)
            k += kq * wd[:, None, None]
    return k, f


def element_batch(Xe, ue, mu, lam, want_k=True):
    if Xe.shape[1] == 4:
        return tet4_batch(Xe, ue, mu, lam, want_k)
    return tet10_batch(Xe, ue, mu, lam, want_k)


# ---------------------------------------------------------------------------
# Strain energy (for finite-difference pinning of f and k)
# ---------------------------------------------------------------------------

def element_energy(X, u, mu, lam):
    """W_e = sum_q w_q detJ_q Psi(F_q) (Tet4: one point with weight V)."""
    K = bulk_from_lame(lam, mu)
    if X.shape[0] == 4:
        V, H = tet4_volume_and_gradients(X)
        return V * neohookean_energy((X + u).T @ H, mu, K)
    qp, qw = tetrahedral_quadrature_4()
    W = 0.0
    for q in range(4):
        dN = tet10_shape_gradients_ref(*qp[q])
        J = X.T @ dN
        H = dN @ np.linalg.inv(J)
        W += qw[q] * np.linalg.det(J) * neohookean_energy((X + u).T @ H, mu, K)
    return W
