"""Free-DOF CSC, residual and the Newton solve on the GPU path.

Cantilever workload (20x4x4, 1920 Tet4) clamped at x = 0.  The device CSC of
K[free][:, free] must equal scipy's (K + empty)[free][:, free].tocsc() bit for
bit in pattern and within 1e-12 in values (FP64); the Newton increment
within 1e-10 (FP64) / 1e-4 (FP32 elements); a two-step loaded solve must
converge to the oracle's solution within 1e-8.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import assembly as A
from oracle import newton as ON
from psierra_amd import workloads as W
from tests.gpu_helpers import banner, rel_err

pytestmark = pytest.mark.gpu


def _clamped(wl):
    from psierra_amd.boundary import DirichletConstraint, DirichletManager, select_nodes_by_plane

    nodes = select_nodes_by_plane(wl.mesh, "yz", "min")
    mgr = DirichletManager(wl.mesh.n_dofs)
    mgr.add_constraint(DirichletConstraint(nodes, (0, 1, 2), 0.0))
    return mgr


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_free_csc_residual_and_newton_step(precision):
    import torch

    from psierra_amd.engine import AssemblyPlan
    from psierra_amd.solver import DirectSolver, FreeDofSystem

    wl = W.cantilever_tet4()
    mgr = _clamped(wl)
    free = mgr.free_dofs
    assert len(mgr.prescribed_dofs) == 75 and len(free) == 1500
    du_ref, r_ref, Kff_ref = ON.newton_step(wl.mesh.nodes, wl.mesh.elements, wl.u, wl.mu, wl.lam, free)
    plan = AssemblyPlan(wl.mesh.nodes, wl.mesh.elements)
    dev = plan.device
    banner(f"newton step {precision}")
    f, k = plan.assemble(torch.as_tensor(wl.u.ravel(), device=dev), torch.as_tensor(wl.mu, device=dev),
                         torch.as_tensor(wl.lam, device=dev), precision=precision)
    sysm = FreeDofSystem(plan, free)
    Kff = sysm.csc(k)
    r, rn = sysm.residual(f)
    assert np.array_equal(Kff.indptr, Kff_ref.indptr) and np.array_equal(Kff.indices, Kff_ref.indices)
    tol = 1e-12 if precision == "fp64" else 1e-5
    assert rel_err(Kff.data, Kff_ref.data) < tol
    assert rel_err(r.cpu().numpy(), r_ref) < (1e-11 if precision == "fp64" else 1e-5)
    assert abs(rn - np.linalg.norm(r_ref)) <= (1e-11 if precision == "fp64" else 1e-5) * np.linalg.norm(r_ref)
    du = DirectSolver().solve(Kff, r.cpu().numpy())
    if precision == "fp64":
        assert rel_err(du, du_ref) < 1e-10
    else:
        # FP32 element errors (~1e-7 of K) are amplified by cond(K_ff) in the
        # forward error; the backward error against the oracle system is the
        # precision-independent check.
        back = np.linalg.norm(Kff_ref @ du - r_ref) / (
            abs(Kff_ref).max() * np.linalg.norm(du) + np.linalg.norm(r_ref))
        assert back < 1e-6
        assert rel_err(du, du_ref) < 1e-3


def test_csc_drops_exact_zeros_like_scipy():
    """Zero a few K values: the device CSC drops them as scipy's CSR '+' does."""
    import torch

    from psierra_amd.engine import AssemblyPlan
    from psierra_amd.solver import FreeDofSystem

    wl = W.cantilever_tet4()
    free = _clamped(wl).free_dofs
    plan = AssemblyPlan(wl.mesh.nodes, wl.mesh.elements)
    dev = plan.device
    banner("csc zero drop")
    _, k = plan.assemble(torch.as_tensor(wl.u.ravel(), device=dev), torch.as_tensor(wl.mu, device=dev),
                         torch.as_tensor(wl.lam, device=dev))
    rng = np.random.default_rng(9)
    zap = rng.choice(plan.nnz, 500, replace=False)
    k[torch.as_tensor(zap, device=dev)] = 0.0
    k[torch.as_tensor(zap[:50], device=dev)] = -0.0
    indptr, indices = plan.pattern()
    import scipy.sparse as spsp

    K_host = spsp.csr_matrix((k.cpu().numpy(), indices, indptr), shape=(plan.n_dofs, plan.n_dofs))
    ref = A.free_dof_csc(K_host, free)
    got = FreeDofSystem(plan, free).csc(k)
    assert np.array_equal(got.indptr, ref.indptr) and np.array_equal(got.indices, ref.indices)
    assert np.array_equal(got.data, ref.data)
    kept = FreeDofSystem(plan, free).csc(k, drop_zeros=False)
    assert kept.nnz > got.nnz


def test_csc_structural_pass_and_compaction_in_turn():
    """The extraction writes every structural entry in one pass and rewrites
    only when an exact zero was dropped: with no zeros, with zeros (dropped,
    and kept with drop_zeros=False: scipy's indexing keeps explicit zeros),
    and with no zeros again on the same object, each output equals scipy's."""
    import scipy.sparse as spsp
    import torch

    from psierra_amd.engine import AssemblyPlan
    from psierra_amd.solver import FreeDofSystem

    wl = W.jittered_box_tet4(n=5, seed=8)
    plan = AssemblyPlan(wl.mesh.nodes, wl.mesh.elements)
    dev = plan.device
    n = plan.n_dofs
    free = _clamped(wl).free_dofs
    u, mu, lam = (torch.as_tensor(x, device=dev) for x in (wl.u.ravel(), wl.mu, wl.lam))
    banner("csc structural pass and compaction")
    _, k = plan.assemble(u, mu, lam)
    indptr, indices = plan.pattern()
    sysm = FreeDofSystem(plan, free)

    def host(kv):
        return spsp.csr_matrix((kv.cpu().numpy(), indices, indptr), shape=(n, n))

    def same(got, ref):
        return (np.array_equal(got.indptr, ref.indptr) and np.array_equal(got.indices, ref.indices)
                and np.array_equal(got.data, ref.data))

    assert int((k == 0).sum()) == 0
    assert same(sysm.csc(k), A.free_dof_csc(host(k), free))
    kz = k.clone()
    kz[torch.as_tensor(np.random.default_rng(4).choice(plan.nnz, 300, replace=False), device=dev)] = 0.0
    assert same(sysm.csc(kz), A.free_dof_csc(host(kz), free))
    assert same(sysm.csc(kz, drop_zeros=False), host(kz)[free, :][:, free].tocsc())
    assert same(sysm.csc(k), A.free_dof_csc(host(k), free))


def test_loaded_cantilever_solve_matches_oracle():
    from psierra_amd.analysis import NonlinearStaticAnalysis
    from psierra_amd.boundary import select_nodes_by_plane
    from psierra_amd.domain import SolidDomain
    from psierra_amd.element import GpuTet4TotalLagrangian
    from psierra_amd.material import NeoHookeanField

    wl = W.cantilever_tet4()
    mgr = _clamped(wl)
    tip = select_nodes_by_plane(wl.mesh, "yz", "max")
    f_ext = np.zeros(wl.mesh.n_dofs)
    f_ext[3 * tip + 2] = -0.05 / len(tip)
    mat = NeoHookeanField(wl.mu, wl.lam)
    dom = SolidDomain(wl.mesh, mat, GpuTet4TotalLagrangian(mat))
    banner("newton solve")
    res = NonlinearStaticAnalysis(dom).solve(mgr, f_ext=f_ext, n_steps=2, nr_iterations=15, nr_tol=1e-8)
    assert res.converged and res.steps_completed == 2
    u_ref, ok, iters = ON.solve(wl.mesh.nodes, wl.mesh.elements, wl.mu, wl.lam, mgr.free_dofs,
                                mgr.prescribed_dofs, mgr.prescribed_values, f_ext, 2, 15, 1e-8)
    assert ok
    assert rel_err(res.displacement, u_ref) < 1e-8
    assert res.iteration_history == iters
    assert abs(res.displacement[3 * tip + 2].mean()) > 1e-3


@pytest.mark.parametrize("element_name", ["tl_fp32", "fbar"])
def test_graph_replayed_newton_solve_equals_eager(element_name):
    # every f_int / K_T assembly of the loop replayed from a HIP graph, with
    # the loop's own eager work (residuals, CSC extraction, line search) in
    # between: the iterates must be the eager solve's bit for bit
    from psierra_amd.analysis import NonlinearStaticAnalysis
    from psierra_amd.boundary import DirichletConstraint, DirichletManager, select_nodes_by_plane
    from psierra_amd.domain import SolidDomain
    from psierra_amd.element import GpuTet4Fbar, GpuTet4TotalLagrangian
    from psierra_amd.material import IsotropicElasticParameters, NeoHookean, NeoHookeanField

    if element_name == "tl_fp32":
        wl = W.cantilever_tet4()
        mesh = wl.mesh
        mgr = _clamped(wl)
        tip = select_nodes_by_plane(mesh, "yz", "max")
        f_ext = np.zeros(mesh.n_dofs)
        f_ext[3 * tip + 2] = -0.05 / len(tip)
        mat = NeoHookeanField(wl.mu, wl.lam)
        make = lambda: GpuTet4TotalLagrangian(mat, precision="fp32")
    else:
        # the nearly incompressible beam of the F-bar locking test
        from tests.test_gpu_fbar import _beam

        mesh = _beam()
        params = IsotropicElasticParameters()
        params.set_from_youngs_poisson(youngs_modulus=1e5, poissons_ratio=0.499, mass_density=1000.0)
        mat = NeoHookean(params)
        mgr = DirichletManager(mesh.n_dofs)
        mgr.add_constraint(DirichletConstraint(select_nodes_by_plane(mesh, "yz", "min"), (0, 1, 2), 0.0))
        tip = select_nodes_by_plane(mesh, "yz", "max")
        f_ext = np.zeros(mesh.n_dofs)
        f_ext[3 * tip + 1] = 5.0 / len(tip)
        make = lambda: GpuTet4Fbar(mat)
    results = []
    for use_graphs in (False, True):
        dom = SolidDomain(mesh, mat, make())
        banner(f"newton {element_name} graphs={use_graphs}")
        res = NonlinearStaticAnalysis(dom).solve(mgr, f_ext=f_ext, n_steps=2, nr_iterations=15, nr_tol=1e-4,
                                                 use_graphs=use_graphs)
        assert res.converged and res.steps_completed == 2
        results.append(res)
    assert np.array_equal(results[0].displacement, results[1].displacement)
    assert results[0].iteration_history == results[1].iteration_history
    assert results[0].residual_history == results[1].residual_history


def test_csc_random_free_set_element_and_fbar_patterns():
    """A random fifth of the DOFs prescribed (so nodes have partly free rows and
    columns) and a few K values zeroed: the device CSC of the element pattern and
    of the F-bar pattern equal scipy's (K + empty)[free][:, free].tocsc() bit for
    bit."""
    import scipy.sparse as spsp
    import torch

    from psierra_amd.engine import AssemblyPlan, FbarAssembly
    from psierra_amd.solver import FreeDofSystem

    wl = W.jittered_box_tet4(n=6, seed=5)
    plan = AssemblyPlan(wl.mesh.nodes, wl.mesh.elements)
    dev = plan.device
    rng = np.random.default_rng(11)
    n = plan.n_dofs
    free = np.setdiff1d(np.arange(n), rng.choice(n, n // 5, replace=False))
    u, mu, lam = (torch.as_tensor(x, device=dev) for x in (wl.u.ravel(), wl.mu, wl.lam))
    banner("csc random free set")
    _, k = plan.assemble(u, mu, lam)
    k[torch.as_tensor(rng.choice(plan.nnz, 200, replace=False), device=dev)] = 0.0
    indptr, indices = plan.pattern()
    ref = A.free_dof_csc(spsp.csr_matrix((k.cpu().numpy(), indices, indptr), shape=(n, n)), free)
    got = FreeDofSystem(plan, free).csc(k)
    assert np.array_equal(got.indptr, ref.indptr) and np.array_equal(got.indices, ref.indices)
    assert np.array_equal(got.data, ref.data)

    fb = FbarAssembly(plan)
    _, k2 = fb.assemble(u, mu, lam, float(np.mean(wl.lam + 2.0 / 3.0 * wl.mu)))
    k2[torch.as_tensor(rng.choice(fb.nnz, 200, replace=False), device=dev)] = 0.0
    indptr2, indices2 = fb.pattern()
    ref2 = A.free_dof_csc(spsp.csr_matrix((k2.cpu().numpy(), indices2, indptr2), shape=(n, n)), free)
    got2 = FreeDofSystem(plan, free, fbar=fb).csc(k2)
    assert np.array_equal(got2.indptr, ref2.indptr) and np.array_equal(got2.indices, ref2.indices)
    assert np.array_equal(got2.data, ref2.data)
