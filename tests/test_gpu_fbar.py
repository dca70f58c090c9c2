"""F-bar Tet4 assembly on the GPU against the F-bar oracle (oracle/fbar.py).

* f and K_dev + K_vol on jittered boxes with per-element mu: FP64 within
  1e-11 (f) / 1e-12 (K), FP32 element arithmetic within 1e-5, relative to
  the norm of the oracle's vector / matrix; the pattern equals the oracle's
  distance-2 pattern bit for bit;
* u = 0: f = 0 exactly and K equal to the oracle's (the exact Jacobian there);
* tiles 64/128/256 give the same bits; inversion raises the reference's
  message with the element index;
* the element hooks, the free-dof CSC over the wider pattern and the Newton
  solve: a nearly incompressible cantilever (nu = 0.499) bends more with
  F-bar than with the standard Tet4, as in the reference's own locking test
  (tests/test_element_tet4_fbar.py:304-364), and each converged solution
  satisfies the oracle's equilibrium.
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as spsp

from oracle import fbar as FB
from psierra_amd import workloads as W
from tests.gpu_helpers import banner, rel_err

pytestmark = pytest.mark.gpu

BULK = 5.7e3


def _gpu(wl, precision, tile=0, want_k=True, bulk=BULK):
    import torch

    from psierra_amd.engine import AssemblyPlan, FbarAssembly

    plan = AssemblyPlan(wl.mesh.nodes, wl.mesh.elements, tile=tile)
    fb = FbarAssembly(plan)
    dev = plan.device
    f, k = fb.assemble(torch.as_tensor(wl.u.reshape(-1), device=dev), torch.as_tensor(wl.mu, device=dev),
                       torch.as_tensor(wl.lam, device=dev), bulk, precision=precision, want_k=want_k)
    torch.cuda.synchronize()
    return fb, f.cpu().numpy(), (k.cpu().numpy() if k is not None else None)


def _k_err(fb, k, K_ref):
    indptr, indices = fb.pattern()
    n = fb.n_dofs
    K = spsp.csr_matrix((k, indices, indptr), shape=(n, n))
    return spsp.linalg.norm(K - K_ref) / spsp.linalg.norm(K_ref)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fbar_matches_oracle(precision):
    wl = W.jittered_box_tet4(n=5, seed=31)
    banner(f"fbar: box 5^3 {precision}")
    fb, f, k = _gpu(wl, precision)
    X, conn, u = wl.mesh.nodes, wl.mesh.elements, wl.u.reshape(-1)
    f_ref = FB.internal_forces(X, conn, u, wl.mu, wl.lam)
    K_ref = FB.tangent_stiffness(X, conn, u, wl.mu, wl.lam, BULK)
    indptr, indices = fb.pattern()
    ip_ref, ix_ref = FB.distance2_pattern(conn, len(X))
    assert np.array_equal(indptr, ip_ref) and np.array_equal(indices, ix_ref)
    tf, tk = (1e-11, 1e-12) if precision == "fp64" else (1e-5, 1e-5)
    assert rel_err(f, f_ref) < tf
    assert _k_err(fb, k, K_ref) < tk


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fbar_larger_box_and_force_only(precision):
    wl = W.jittered_box_tet4(n=10, seed=32)
    banner(f"fbar: box 10^3 {precision}")
    fb, f, k = _gpu(wl, precision)
    X, conn, u = wl.mesh.nodes, wl.mesh.elements, wl.u.reshape(-1)
    f_ref = FB.internal_forces(X, conn, u, wl.mu, wl.lam)
    K_ref = FB.tangent_stiffness(X, conn, u, wl.mu, wl.lam, BULK)
    tf, tk = (1e-11, 1e-12) if precision == "fp64" else (1e-5, 1e-5)
    assert rel_err(f, f_ref) < tf
    assert _k_err(fb, k, K_ref) < tk
    # the f-only path (line-search trials; its own summation order)
    _, f2, _ = _gpu(wl, precision, want_k=False)
    assert rel_err(f2, f_ref) < tf


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fbar_at_rest(precision):
    wl = W.jittered_box_tet4(n=4, seed=33)
    wl.u = np.zeros_like(wl.u)
    mu = np.full_like(wl.mu, 1.0e3)
    wl.mu, wl.lam = mu, 5.0 * mu
    bulk = 5.0e3 + 2.0e3 / 3.0
    banner(f"fbar: at rest {precision}")
    fb, f, k = _gpu(wl, precision, bulk=bulk)
    assert np.all(f == 0.0)
    K_ref = FB.tangent_stiffness(wl.mesh.nodes, wl.mesh.elements, np.zeros(wl.mesh.n_dofs), wl.mu, wl.lam, bulk)
    assert _k_err(fb, k, K_ref) < (1e-12 if precision == "fp64" else 1e-5)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fbar_tiles_bit_identical(precision):
    wl = W.jittered_box_tet4(n=7, seed=34)
    banner(f"fbar: tiles 64/128/256 {precision}")
    ref = None
    for tile in (64, 128, 256):
        _, f, k = _gpu(wl, precision, tile=tile)
        if ref is None:
            ref = (f, k)
        else:
            assert np.array_equal(f, ref[0]) and np.array_equal(k, ref[1])


def test_fbar_inversion_raises():
    wl = W.jittered_box_tet4(n=3, seed=35)
    conn = wl.mesh.elements
    X = wl.mesh.nodes
    u = np.zeros_like(X)
    u[conn[5, 0]] = 2.0 * (X[conn[5, 1]] - X[conn[5, 0]])
    wl.u = u
    banner("fbar: inversion")
    with pytest.raises(ValueError, match=r"Element inversion detected at element \d+: det\(F\) = "):
        _gpu(wl, "fp64")


def _beam(nx=4, ny=2, nz=2):
    """Structured 5-tet-per-cube beam with alternating splits, unit cubes
    (the mesh of tests/test_element_tet4_fbar.py:66-118)."""
    from psierra_amd.mesh import TetMesh

    ids = {}
    nodes = []

    def nid(i, j, k):
        if (i, j, k) not in ids:
            ids[(i, j, k)] = len(nodes)
            nodes.append((float(i), float(j), float(k)))
        return ids[(i, j, k)]

    tets = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                n0, n1, n2, n3 = nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k)
                n4, n5, n6, n7 = nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1)
                if (i + j + k) % 2 == 0:
                    tets += [[n0, n1, n3, n4], [n1, n2, n3, n6], [n1, n4, n5, n6], [n3, n4, n6, n7], [n1, n3, n4, n6]]
                else:
                    tets += [[n0, n1, n2, n5], [n0, n2, n3, n7], [n0, n4, n5, n7], [n2, n5, n6, n7], [n2, n0, n5, n7]]
    return TetMesh(np.array(nodes), np.array(tets, dtype=np.int64))


def test_fbar_relieves_locking_in_newton_solve():
    from psierra_amd.analysis import NonlinearStaticAnalysis
    from psierra_amd.boundary import DirichletConstraint, DirichletManager, select_nodes_by_plane
    from psierra_amd.domain import SolidDomain
    from psierra_amd.element import GpuTet4Fbar, GpuTet4TotalLagrangian
    from psierra_amd.material import IsotropicElasticParameters, NeoHookean

    mesh = _beam()
    params = IsotropicElasticParameters()
    params.set_from_youngs_poisson(youngs_modulus=1e5, poissons_ratio=0.499, mass_density=1000.0)
    mat = NeoHookean(params)
    fixed = select_nodes_by_plane(mesh, plane="yz", bound="min")
    mgr = DirichletManager(mesh.n_dofs)
    mgr.add_constraint(DirichletConstraint(fixed, (0, 1, 2), 0.0))
    tip = select_nodes_by_plane(mesh, plane="yz", bound="max")
    f_ext = np.zeros(mesh.n_dofs)
    f_ext[3 * tip + 1] = 5.0 / len(tip)
    defl = {}
    for name, elem in (("tl", GpuTet4TotalLagrangian(mat)), ("fbar", GpuTet4Fbar(mat))):
        banner(f"fbar: locking beam {name}")
        dom = SolidDomain(mesh, mat, elem)
        res = NonlinearStaticAnalysis(dom).solve(dirichlet=mgr, f_ext=f_ext, n_steps=4, nr_iterations=15,
                                                 nr_tol=1e-4)
        assert res.converged and res.steps_completed == 4
        defl[name] = float(np.mean(dom.displacement[tip, 1]))
        if name == "fbar":
            # equilibrium of the converged state in the oracle's F-bar force
            K = params.bulk_modulus
            mu = np.full(mesh.n_elements, params.shear_modulus)
            f_int = FB.internal_forces(mesh.nodes, mesh.elements, dom.u, mu, K - 2.0 / 3.0 * mu)
            free = mgr.free_dofs
            assert np.linalg.norm((f_int - f_ext)[free]) < 1e-4 * np.linalg.norm(f_ext[free])
    assert defl["fbar"] > 1.15 * defl["tl"]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fbar_isolated_nodes_and_single_element(precision):
    from psierra_amd.mesh import TetMesh
    from psierra_amd.workloads import Workload

    nodes = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [5, 5, 5], [6, 5, 5.0]])
    conn = np.array([[0, 1, 2, 3]], dtype=np.int64)
    rng = np.random.default_rng(41)
    wl = Workload("fbar_one", TetMesh(nodes, conn), 0.02 * rng.standard_normal(nodes.shape),
                  np.array([1e3]), np.array([5e3]))
    banner(f"fbar: single element + isolated nodes {precision}")
    fb, f, k = _gpu(wl, precision)
    f_ref = FB.internal_forces(nodes, conn, wl.u.reshape(-1), wl.mu, wl.lam)
    K_ref = FB.tangent_stiffness(nodes, conn, wl.u.reshape(-1), wl.mu, wl.lam, BULK)
    tf, tk = (1e-11, 1e-12) if precision == "fp64" else (1e-5, 1e-5)
    assert rel_err(f, f_ref) < tf and _k_err(fb, k, K_ref) < tk
    assert np.all(f[12:] == 0.0)
    indptr, _ = fb.pattern()
    assert np.all(np.diff(indptr)[12:] == 0)


@pytest.mark.parametrize("n_elem", [12, 20])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fbar_kvol_by_slot_with_many_units(precision, n_elem):
    """Fans of 12 and 20 elements: node 0 has 37 / 61 element neighbours
    (units), 181 / 301 K_vol terms and 37 / 61 slots, inside the limits of
    K_vol by slot (64 units, 512 terms, 128 slots), so the unit records of
    lanes beyond 32 feed the terms' g indices; the 33-element fan below
    (100 units) takes the unit-step kernel."""
    from tests.test_gpu_edge_cases import _fan

    wl = _fan(n_elem, seed=42)
    banner(f"fbar: fan of {n_elem} elements {precision}")
    fb, f, k = _gpu(wl, precision)
    X, conn, u = wl.mesh.nodes, wl.mesh.elements, wl.u.reshape(-1)
    f_ref = FB.internal_forces(X, conn, u, wl.mu, wl.lam)
    K_ref = FB.tangent_stiffness(X, conn, u, wl.mu, wl.lam, BULK)
    tf, tk = (1e-11, 1e-12) if precision == "fp64" else (1e-5, 1e-5)
    assert rel_err(f, f_ref) < tf and _k_err(fb, k, K_ref) < tk


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fbar_node_with_33_incidences(precision):
    # a fan of 33 elements around node 0: the 64-bit feed masks of the row
    # kernel, and K_vol by unit steps (100 units: beyond the slot layout)
    from tests.test_gpu_edge_cases import _fan

    wl = _fan(33, seed=42)
    banner(f"fbar: fan of 33 elements {precision}")
    fb, f, k = _gpu(wl, precision)
    X, conn, u = wl.mesh.nodes, wl.mesh.elements, wl.u.reshape(-1)
    f_ref = FB.internal_forces(X, conn, u, wl.mu, wl.lam)
    K_ref = FB.tangent_stiffness(X, conn, u, wl.mu, wl.lam, BULK)
    tf, tk = (1e-11, 1e-12) if precision == "fp64" else (1e-5, 1e-5)
    assert rel_err(f, f_ref) < tf and _k_err(fb, k, K_ref) < tk
