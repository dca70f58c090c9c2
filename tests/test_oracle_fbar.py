"""Pins of the F-bar oracle (oracle/fbar.py), CPU only.

Why the tangent at u = 0 is the exact Jacobian of the assembled force.  At
u = 0 every ratio is 1 and B_bar = B.  With Hooke's C = 2 mu P_dev + K I x I
(K = lambda + 2 mu / 3 = K_bulk for a homogeneous material), the linearised
F-bar strain is dev(dE) + dJ_bar/3 I, so
  df_a = sum_e V_e B_e^T [2 mu dev(dE_e) + K dJ_bar,e I].
The first term is B_dev^T (C - K m m^T) B_dev V (K_dev at u = 0, S = 0).  With
dJ_bar,e = 1/4 sum_{n in e} dJ_node,n and dJ_node,n = g_n . du
(w_{e,n} = V_e / 4 / V0_n at J = 1), the second term is
  K sum_e V_e div_e 1/4 sum_{n in e} g_n . du = sum_n K V0_n g_n (g_n . du),
which is K_vol du.  A wrong weight, a lost factor 1/4, J in place of 1, or a
sign error in either term breaks the equality.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import assembly as A
from oracle import fbar as FB
from oracle import fem
from psierra_amd import workloads as W


def _box(n=2, seed=21):
    wl = W.jittered_box_tet4(n=n, seed=seed)
    return wl.mesh.nodes, wl.mesh.elements, wl


def test_tangent_at_rest_is_jacobian_of_force():
    X, conn, _ = _box(2, 21)
    mu, lam = 1.0e3, 5.0e3
    K_bulk = fem.bulk_from_lame(lam, mu)
    n = 3 * len(X)
    u0 = np.zeros(n)
    K = FB.tangent_stiffness(X, conn, u0, mu, lam, K_bulk).toarray()
    h = 1e-6
    Kfd = np.zeros_like(K)
    for j in range(n):
        e = np.zeros(n)
        e[j] = h
        Kfd[:, j] = (FB.internal_forces(X, conn, e, mu, lam) - FB.internal_forces(X, conn, -e, mu, lam)) / (2 * h)
    assert np.linalg.norm(K - Kfd) / np.linalg.norm(K) < 1e-7
    # the volumetric part matters here: without it the check fails
    K_vol = FB.volumetric_stiffness(X, conn, u0, K_bulk).toarray()
    assert np.linalg.norm(K - K_vol - Kfd) / np.linalg.norm(K) > 1e-2


def test_homogeneous_deformation_equals_total_lagrangian_force():
    X, conn, _ = _box(3, 22)
    rng = np.random.default_rng(5)
    F0 = np.identity(3) + 0.05 * rng.standard_normal((3, 3))
    u = ((X @ F0.T) - X).reshape(-1)
    mu = 1.0e3 * np.exp(0.3 * rng.standard_normal(len(conn)))
    lam = 5.0 * mu
    V, H, V0 = FB.setup(X, conn)
    _, J, J_node, J_bar = FB.smoothed_jacobians(conn, u, V, H, V0)
    assert np.allclose(J_bar, J, rtol=1e-13) and np.allclose(J_node, np.linalg.det(F0), rtol=1e-13)
    f = FB.internal_forces(X, conn, u, mu, lam)
    f_tl, _ = A.assemble(X, conn, u.reshape(-1, 3), mu, lam, want_k=False)
    assert np.linalg.norm(f - f_tl) <= 1e-11 * np.linalg.norm(f_tl)


def test_rigid_motions_give_zero_force():
    X, conn, _ = _box(2, 23)
    mu, lam = 1.0e3, 5.0e3
    t = np.array([0.3, -0.2, 0.1])
    f = FB.internal_forces(X, conn, np.tile(t, len(X)), mu, lam)
    assert np.abs(f).max() < 1e-10
    c, s = np.cos(np.pi / 6), np.sin(np.pi / 6)
    R = np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])
    u = (X @ R.T - X).reshape(-1)
    f = FB.internal_forces(X, conn, u, mu, lam)
    assert np.abs(f).max() < 1e-9


def test_tangent_symmetric_and_on_distance2_pattern():
    X, conn, wl = _box(3, 24)
    K_bulk = 5.0e3
    K = FB.tangent_stiffness(X, conn, wl.u.reshape(-1), wl.mu, wl.lam, K_bulk)
    D = K.toarray()
    assert np.abs(D - D.T).max() <= 1e-12 * np.abs(D).max()
    indptr, indices = FB.distance2_pattern(conn, len(X))
    P = np.zeros_like(D, dtype=bool)
    for r in range(D.shape[0]):
        P[r, indices[indptr[r]:indptr[r + 1]]] = True
    assert not np.any((D != 0) & ~P)
    # the CSR sum drops exact-zero entries (a few couplings of the unjittered
    # boundary nodes cancel exactly); everything else of the pattern is there
    assert indices.size - 24 <= K.nnz <= indices.size


def test_volumetric_part_matches_per_node_loop():
    X, conn, wl = _box(2, 25)
    K_bulk = 4.2e3
    u = wl.u.reshape(-1)
    K_vec = FB.volumetric_stiffness(X, conn, u, K_bulk).toarray()
    K_loop = FB.volumetric_blocks_loop(X, conn, u, K_bulk)
    assert np.abs(K_vec - K_loop).max() <= 1e-13 * np.abs(K_loop).max()


def test_inversion_raises_with_element_index():
    X, conn, _ = _box(2, 26)
    u = np.zeros_like(X)
    u[conn[3, 0]] = (X[conn[3, 1]] - X[conn[3, 0]]) * 2.0   # push node 0 through its face
    with pytest.raises(ValueError, match="Element inversion detected at element"):
        FB.internal_forces(X, conn, u.reshape(-1), 1e3, 5e3)
