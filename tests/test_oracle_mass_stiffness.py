"""The oracle's consistent mass and linear stiffness, pinned.

* The Tet4 mass tensor equals the closed form of tet4_base.py:138-151,
  (1 + delta_ab) / 20; both tensors integrate the product of partitions of
  unity to 1 (total mass rho V), are symmetric, and give the textbook Tet10
  row sums (corner nodes -1/20, mid-edge nodes 1/5 of the element mass).
* The Tet10 tensor has the classic denominators: 2520 c is integer-valued
  with corner diagonal 36, mid-edge diagonal 192 (Zienkiewicz & Taylor).
* Assembled mass: 1^T M_x 1 = rho * total volume; M is symmetric.
* Linear stiffness (B^T C B V) equals the oracle's Neo-Hookean tangent at
  u = 0 (S = 0, C4 = Hooke) and annihilates rigid translations and
  infinitesimal rotations.
"""
from __future__ import annotations

import numpy as np

from oracle import assembly as A
from psierra_amd import workloads as W


def test_mass_tensors():
    c4 = A.reference_mass_tensor(4)
    assert np.allclose(c4, (np.ones((4, 4)) + np.eye(4)) / 20.0, rtol=0, atol=1e-15)
    c10 = A.reference_mass_tensor(10)
    for c in (c4, c10):
        assert np.allclose(c, c.T, atol=1e-15)
        assert abs(c.sum() - 1.0) < 1e-13
    rows = c10.sum(axis=1)
    assert np.allclose(rows[:4], -1.0 / 20.0, atol=1e-14)
    assert np.allclose(rows[4:], 1.0 / 5.0, atol=1e-14)
    ci = 2520.0 * c10
    assert np.allclose(ci, np.round(ci), atol=1e-9)
    assert np.allclose(np.diag(ci)[:4], 36.0) and np.allclose(np.diag(ci)[4:], 192.0)


def test_assembled_mass_total_and_symmetry():
    for wl in (W.cantilever_tet4(), W.jittered_box_tet10(n=3, seed=2)):
        M = A.mass_matrix(wl.mesh.nodes, wl.mesh.elements, 7.5)
        V = A.element_volumes(wl.mesh.nodes, wl.mesh.elements).sum()
        ex = np.zeros(M.shape[0])
        ex[0::3] = 1.0
        assert abs(ex @ (M @ ex) - 7.5 * V) <= 1e-12 * 7.5 * V
        assert abs(M - M.T).max() <= 1e-15 * abs(M).max()


def test_linear_stiffness_is_tangent_at_rest():
    wl = W.cantilever_tet4()
    K = A.linear_stiffness(wl.mesh.nodes, wl.mesh.elements, wl.mu, wl.lam)
    _, K0 = A.assemble(wl.mesh.nodes, wl.mesh.elements, np.zeros_like(wl.u), wl.mu, wl.lam)
    assert np.array_equal(K.indptr, K0.indptr) and np.array_equal(K.indices, K0.indices)
    assert np.linalg.norm(K.data - K0.data) <= 1e-13 * np.linalg.norm(K0.data)
    X = wl.mesh.nodes
    for t in np.eye(3):
        assert np.abs(K @ np.tile(t, len(X))).max() <= 1e-9 * abs(K).max()
    w = np.array([0.3, -0.2, 0.5])
    rot = np.cross(w, X).ravel()
    assert np.abs(K @ rot).max() <= 1e-9 * abs(K).max()
