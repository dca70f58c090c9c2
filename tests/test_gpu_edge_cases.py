"""Edge cases of the Tet4 assembly on the GPU, against the FP64 oracle.

* nodes that no element references (empty CSR rows, f = 0 exactly) and a
  one-element mesh;
* a node shared by 32, 33 and 64 elements (the kernel's feed masks are
  32-bit up to 32 incidences per node, 64-bit above), and by 65, 128 and 200
  (a wide plan: the mask-free row kernel, f-only, adjoint and free CSC;
  the kernels that need the masks or rank bytes refuse it);
* zero displacement: f = 0 bit-exactly and K equal to the small-strain
  stiffness of the oracle (S = 0, C = Hookean);
* every supported tile size gives the same bits on a mesh whose node runs
  straddle tile boundaries in many ways.
Tolerances as in test_gpu_tet4.py: FP64 1e-12 (K) / 1e-11 (f), FP32 1e-5.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import assembly as A
from psierra_amd.mesh import TetMesh
from psierra_amd.workloads import Workload
from tests.gpu_helpers import banner, gpu_assemble, rel_err

pytestmark = pytest.mark.gpu


def _check(wl, precision):
    f_ref, K_ref = A.assemble(wl.mesh.nodes, wl.mesh.elements, wl.u, wl.mu, wl.lam)
    plan, f, k = gpu_assemble(wl, precision)
    indptr, indices = plan.pattern()
    assert np.array_equal(indptr, K_ref.indptr) and np.array_equal(indices, K_ref.indices)
    tk, tf = (1e-12, 1e-11) if precision == "fp64" else (1e-5, 1e-5)
    assert rel_err(k, K_ref.data) < tk
    assert rel_err(f, f_ref) < tf
    return plan, f, k, f_ref


def _fan(n_elem: int, seed: int) -> Workload:
    """n_elem tetrahedra sharing node 0 and nothing else, all positively oriented."""
    rng = np.random.default_rng(seed)
    nodes = [np.zeros(3)]
    conn = []
    for e in range(n_elem):
        th = 2 * np.pi * e / n_elem
        d = np.array([np.cos(th), np.sin(th), 0.1 * rng.standard_normal()])
        base = 1.0 + 0.05 * rng.standard_normal()
        p1 = base * d
        p2 = p1 + 0.3 * np.array([-np.sin(th), np.cos(th), 0.0])
        p3 = p1 + np.array([0.0, 0.0, 0.4])
        pts = [p1, p2, p3]
        ids = list(range(len(nodes), len(nodes) + 3))
        nodes.extend(pts)
        X = np.array([nodes[0], *pts])
        if np.linalg.det((X[1:] - X[0]).T) < 0:
            ids[1], ids[2] = ids[2], ids[1]
        conn.append([0, *ids])
    nodes = np.array(nodes)
    conn = np.array(conn, dtype=np.int64)
    u = 0.01 * rng.standard_normal(nodes.shape)
    mu = 1e3 * np.exp(0.3 * rng.standard_normal(n_elem))
    return Workload(f"fan_{n_elem}", TetMesh(nodes, conn), u, mu, 5.0 * mu)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_isolated_nodes_and_single_element(precision):
    nodes = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [5, 5, 5], [6, 5, 5], [7, 7, 7.0]])
    # one positively oriented element on nodes 0..3; nodes 4..6 are unreferenced
    conn = np.array([[0, 1, 2, 3]], dtype=np.int64)
    rng = np.random.default_rng(7)
    u = 0.02 * rng.standard_normal(nodes.shape)
    wl = Workload("one_element", TetMesh(nodes, conn), u, np.array([1e3]), np.array([5e3]))
    banner(f"edge: isolated nodes + single element {precision}")
    plan, f, _, _ = _check(wl, precision)
    assert np.all(f[12:] == 0.0)
    indptr, _ = plan.pattern()
    assert np.all(np.diff(indptr)[12:] == 0)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("n_elem", [32, 33, 64])
def test_node_with_many_incidences(precision, n_elem):
    """32 is the largest run the 32-bit feed masks take; 33 and 64 use the
    64-bit masks.  The fan's tiles also cover both ways a tile reads its
    nodes: the tile holding the centre of the 64-element fan has 193 own and
    halo nodes, beyond the 128 a tile stages in LDS, so it gathers from the
    device copy, while the other tiles (and the smaller fans) stage theirs.
    f + K, the f-only kernel and the adjoint K_T w against the oracle."""
    import torch

    wl = _fan(n_elem, seed=8)
    banner(f"edge: fan of {n_elem} elements {precision}")
    plan, f, _, _ = _check(wl, precision)
    assert plan.info["tile"] == 256
    dev = plan.device
    u = torch.as_tensor(wl.u.reshape(-1), device=dev)
    mu = torch.as_tensor(wl.mu, device=dev)
    lam = torch.as_tensor(wl.lam, device=dev)
    f_only, _ = plan.assemble(u, mu, lam, precision=precision, want_k=False)
    assert rel_err(f_only.cpu().numpy(), f) < (1e-13 if precision == "fp64" else 1e-6)
    _, K_ref = A.assemble(wl.mesh.nodes, wl.mesh.elements, wl.u, wl.mu, wl.lam)
    w = np.random.default_rng(n_elem).standard_normal(plan.n_dofs)
    gu, _, _ = plan.vjp(u, mu, lam, torch.as_tensor(w, device=dev), precision=precision, want_params=False)
    assert rel_err(gu.cpu().numpy(), K_ref @ w) < (1e-12 if precision == "fp64" else 1e-5)


@pytest.mark.parametrize("n_elem", [21, 64])
def test_free_csc_of_high_degree_node(n_elem):
    """The free-DOF CSC when a column node has more than 21 neighbours (its
    3 x neighbours (row, neighbour) pairs take several passes of a wave): bit for
    bit scipy's (K + empty)[free][:, free].tocsc(), all DOFs free and a random
    third prescribed."""
    import scipy.sparse as spsp
    import torch

    from psierra_amd.solver import FreeDofSystem

    wl = _fan(n_elem, seed=12)
    banner(f"edge: free CSC of a fan of {n_elem} elements")
    plan, _, k, _ = _check(wl, "fp64")
    kt = torch.as_tensor(k, device=plan.device)
    indptr, indices = plan.pattern()
    n = plan.n_dofs
    K = spsp.csr_matrix((k, indices, indptr), shape=(n, n))
    rng = np.random.default_rng(13)
    for free in (np.arange(n), np.setdiff1d(np.arange(n), rng.choice(n, n // 3, replace=False))):
        ref = A.free_dof_csc(K, free)
        got = FreeDofSystem(plan, free).csc(kt)
        assert np.array_equal(got.indptr, ref.indptr) and np.array_equal(got.indices, ref.indices)
        assert np.array_equal(got.data, ref.data)


@pytest.mark.parametrize("n_elem", [33, 64])
def test_mass_of_high_degree_node(n_elem):
    """Consistent mass on a fan (one node with 33 or 64 incidences and up to
    193 neighbours) against the oracle's coo.tocsr assembly."""
    import torch

    from psierra_amd.engine import AssemblyPlan

    wl = _fan(n_elem, seed=14)
    banner(f"edge: mass of a fan of {n_elem} elements")
    plan = AssemblyPlan(wl.mesh.nodes, wl.mesh.elements)
    rho = 1.0e3 * np.exp(0.2 * np.random.default_rng(15).standard_normal(plan.n_elements))
    got = plan.assemble_mass(torch.as_tensor(rho, device=plan.device)).cpu().numpy()
    ref = A.mass_matrix(wl.mesh.nodes, wl.mesh.elements, rho)
    indptr, indices = plan.pattern()
    assert np.array_equal(indptr, ref.indptr) and np.array_equal(indices, ref.indices)
    assert rel_err(got, ref.data) < 1e-13


def test_free_csc_degenerate_free_sets():
    """The free-DOF CSC with no free dof, a single free dof and a single free
    node, against scipy (shapes and every array)."""
    import scipy.sparse as spsp
    import torch

    from psierra_amd.solver import FreeDofSystem

    wl = _fan(5, seed=16)
    banner("edge: free CSC of degenerate free sets")
    plan, _, k, _ = _check(wl, "fp64")
    kt = torch.as_tensor(k, device=plan.device)
    indptr, indices = plan.pattern()
    n = plan.n_dofs
    K = spsp.csr_matrix((k, indices, indptr), shape=(n, n))
    for free in (np.array([], dtype=np.int64), np.array([4]), np.array([0, 1, 2])):
        ref = A.free_dof_csc(K, free)
        got = FreeDofSystem(plan, free).csc(kt)
        assert got.shape == ref.shape == (free.size, free.size)
        assert np.array_equal(got.indptr, ref.indptr) and np.array_equal(got.indices, ref.indices)
        assert np.array_equal(got.data, ref.data)
    # the residual of an empty free set is empty with norm 0
    sysm = FreeDofSystem(plan, np.array([], dtype=np.int64))
    sysm.sumsq.fill_(123.0)
    r, rn = sysm.residual(torch.zeros(n, dtype=torch.float64, device=plan.device))
    assert r.numel() == 0 and rn == 0.0


def _fan_and_box(n_elem: int, seed: int) -> Workload:
    """A fan of n_elem elements around one node (3 n_elem neighbours) next to
    a jittered 5^3 box (750 elements in many ordinary tiles), one mesh."""
    from psierra_amd import workloads as W

    fan = _fan(n_elem, seed)
    box = W.jittered_box_tet4(n=5, seed=seed)
    off = box.mesh.nodes.shape[0]
    nodes = np.vstack([box.mesh.nodes, fan.mesh.nodes + np.array([3.0, 0.0, 0.0])])
    conn = np.vstack([box.mesh.elements, fan.mesh.elements + off])
    u = np.vstack([box.u, fan.u])
    mu = np.concatenate([box.mu, fan.mu])
    return Workload(f"fan{n_elem}_box5", TetMesh(nodes, conn), u, mu, 5.0 * mu)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("n_elem", [65, 128, 200])
def test_wide_node_matches_oracle(precision, n_elem):
    """A node shared by 65, 128 or 200 elements (beyond the 64-bit feed masks;
    200 also beyond the 255-neighbour slot-rank bytes): a wide plan, whose
    Tet4 row kernel stages feeds at positions from the plan.  Pattern
    bit-exact, K and f against the oracle, the f-only kernel and the adjoint
    K_T w too, and the same bits for every tile that holds the run."""
    import torch

    from psierra_amd.engine import AssemblyPlan

    wl = _fan_and_box(n_elem, seed=9)
    banner(f"edge: wide plan, fan of {n_elem} elements {precision}")
    plan, f, k, f_ref = _check(wl, precision)
    assert plan.info["wide"] == 1 and plan.info["max_incidences"] == n_elem
    dev = plan.device
    u = torch.as_tensor(wl.u.reshape(-1), device=dev)
    mu = torch.as_tensor(wl.mu, device=dev)
    lam = torch.as_tensor(wl.lam, device=dev)
    f_only, _ = plan.assemble(u, mu, lam, precision=precision, want_k=False)
    assert rel_err(f_only.cpu().numpy(), f) < (1e-13 if precision == "fp64" else 1e-6)
    _, K_ref = A.assemble(wl.mesh.nodes, wl.mesh.elements, wl.u, wl.mu, wl.lam)
    w = np.random.default_rng(n_elem).standard_normal(plan.n_dofs)
    gu, _, _ = plan.vjp(u, mu, lam, torch.as_tensor(w, device=dev), precision=precision, want_params=False)
    assert rel_err(gu.cpu().numpy(), K_ref @ w) < (1e-12 if precision == "fp64" else 1e-5)
    for tile in (64, 128, 256):
        if tile < n_elem:
            with pytest.raises(Exception, match="exceeds the tile"):
                AssemblyPlan(wl.mesh.nodes, wl.mesh.elements, tile=tile)
            continue
        p2 = AssemblyPlan(wl.mesh.nodes, wl.mesh.elements, tile=tile)
        f2, k2 = p2.assemble(u, mu, lam, precision=precision)
        assert np.array_equal(f2.cpu().numpy(), f) and np.array_equal(k2.cpu().numpy(), k), tile


def test_wide_plan_free_csc_and_refusals():
    """The free-DOF CSC of a wide plan (a column node with 601 neighbours) is
    scipy's bit for bit; the mass and F-bar kernels refuse a wide plan with a
    clear message."""
    import torch

    from psierra_amd.engine import FbarAssembly
    from psierra_amd.solver import FreeDofSystem

    wl = _fan_and_box(200, seed=10)
    banner("edge: wide plan free CSC and refusals")
    plan, _, k, _ = _check(wl, "fp64")
    indptr, indices = plan.pattern()
    import scipy.sparse as spsp

    n = plan.n_dofs
    K = spsp.csr_matrix((k, indices, indptr), shape=(n, n))
    free = np.sort(np.random.default_rng(3).choice(n, 2 * n // 3, replace=False))
    ref = (K + spsp.csr_matrix((n, n)))[free][:, free].tocsc()
    sysm = FreeDofSystem(plan, free)
    got = sysm.csc(torch.as_tensor(k, device=plan.device))
    assert np.array_equal(got.indptr, ref.indptr) and np.array_equal(got.indices, ref.indices)
    assert np.array_equal(got.data, ref.data)
    with pytest.raises(Exception, match="at most 64 incident elements"):
        plan.assemble_mass(1.0)
    with pytest.raises(Exception, match="at most 64 incident elements"):
        FbarAssembly(plan)


def _tet10_wide(n_elem: int, seed: int) -> Workload:
    wl = _fan_and_box(n_elem, seed)
    from psierra_amd.mesh import tet4_to_tet10

    X10, c10 = tet4_to_tet10(wl.mesh.nodes, wl.mesh.elements)
    u = 0.005 * np.random.default_rng(seed).standard_normal(X10.shape)
    return Workload(f"tet10_fan{n_elem}_box5", TetMesh(X10, c10), u, wl.mu, wl.lam)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("n_elem", [65, 128, 200])
def test_wide_tet10_matches_oracle(precision, n_elem):
    """Tet10: a corner node shared by 65, 128 or 200 elements makes the plan
    wide; its general path (element rows into scratch, every CSR value and f
    entry summed over its contributions in element order) against the
    oracle, the f-only path, and the same bits on a second call."""
    import torch

    wl = _tet10_wide(n_elem, seed=19)
    banner(f"edge: wide Tet10 plan, fan of {n_elem} elements {precision}")
    plan, f, k, f_ref = _check(wl, precision)
    assert plan.info["wide"] == 1 and plan.info["max_incidences"] == n_elem and plan.npe == 10
    dev = plan.device
    u = torch.as_tensor(wl.u.reshape(-1), device=dev)
    mu = torch.as_tensor(wl.mu, device=dev)
    lam = torch.as_tensor(wl.lam, device=dev)
    f2, k2 = plan.assemble(u, mu, lam, precision=precision)
    assert np.array_equal(f2.cpu().numpy(), f) and np.array_equal(k2.cpu().numpy(), k)
    f_only, _ = plan.assemble(u, mu, lam, precision=precision, want_k=False)
    assert np.array_equal(f_only.cpu().numpy(), f)


def test_wide_tet10_mooney_rivlin_matches_oracle():
    """Another model through the same general path: Mooney-Rivlin (c10, c01,
    K per element) against the oracle's Tet10 element summed by the
    reference's scatter."""
    import scipy.sparse as spsp
    import torch

    from oracle import fem
    from psierra_amd.engine import AssemblyPlan

    wl = _tet10_wide(80, seed=20)
    m = wl.mesh
    E = m.n_elements
    c10 = 0.4 * wl.mu
    c01 = 0.1 * wl.mu
    K = wl.lam + (2.0 / 3.0) * wl.mu
    banner("edge: wide Tet10 plan, Mooney-Rivlin")
    plan = AssemblyPlan(m.nodes, m.elements)
    assert plan.info["wide"] == 1
    dev = plan.device
    prm = [torch.as_tensor(a, device=dev) for a in (c10, c01, K)]
    f, k = plan.assemble_hyperelastic(torch.as_tensor(wl.u.reshape(-1), device=dev), "mooney_rivlin", prm,
                                      precision="fp64")
    n = 3 * m.n_nodes
    f_ref = np.zeros(n)
    rows, cols, vals = [], [], []
    for e in range(E):
        conn = m.elements[e]
        ke, fe = fem.tet10_force_and_tangent(m.nodes[conn], wl.u[conn],
                                             fem.material_fn(1, (c10[e], c01[e], K[e])))
        dofs = (3 * conn[:, None] + np.arange(3)).ravel()
        np.add.at(f_ref, dofs, fe)
        rows.append(np.repeat(dofs, 30))
        cols.append(np.tile(dofs, 30))
        vals.append(ke.ravel())
    K_ref = spsp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n)).tocsr()
    indptr, indices = plan.pattern()
    assert np.array_equal(indptr, K_ref.indptr) and np.array_equal(indices, K_ref.indices)
    assert rel_err(k.cpu().numpy(), K_ref.data) < 1e-12
    assert rel_err(f.cpu().numpy(), f_ref) < 1e-11


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_zero_displacement(precision):
    from psierra_amd import workloads as W

    wl = W.jittered_box_tet4(n=6, seed=12)
    wl.u = np.zeros_like(wl.u)
    banner(f"edge: zero displacement {precision}")
    # The oracle forms F = (X + u)^T H, so its f is rounding noise (~1e-14
    # of the stiffness scale) rather than 0; the engine forms grad u directly
    # and returns exact zeros.  K is compared as usual.
    f_ref, K_ref = A.assemble(wl.mesh.nodes, wl.mesh.elements, wl.u, wl.mu, wl.lam)
    plan, f, k = gpu_assemble(wl, precision)
    assert np.all(f == 0.0)
    assert np.abs(f_ref).max() < 1e-9 * np.abs(K_ref.data).max()
    assert rel_err(k, K_ref.data) < (1e-12 if precision == "fp64" else 1e-5)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_all_tiles_bit_identical(precision):
    from psierra_amd import workloads as W

    wl = W.jittered_box_tet4(n=9, seed=13)
    banner(f"edge: tiles 64/128/256 {precision}")
    ref = None
    for tile in (64, 128, 256):
        plan, f, k = gpu_assemble(wl, precision, tile=tile)
        assert plan.info["tile"] == tile
        if ref is None:
            ref = (f, k)
        else:
            assert np.array_equal(f, ref[0]) and np.array_equal(k, ref[1])
