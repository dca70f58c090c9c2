"""Batched ECSW projection on FP32 matrix cores vs the FP64 oracle.

Six seeded instances of uneven size (boxes with 4..9 hexes per side, rank 64
orthonormal bases, 5% RID with weights U(0.5, 2)) in one launch, and ranks
8, 30 and 64.  K_r and f_r within 1e-4 relative (Frobenius) of the FP64
restatement of HyperReducedECSWModel.evaluate_forces_and_stiffness; K_r
exactly symmetric (model.py:317); the model built with the reference
constructor (model.py:50-59) equals the batch and projects M, K_lin and the
base load as the reference does.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import ecsw as OE
from psierra_amd import workloads as W
from tests.gpu_helpers import banner, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def instances():
    return W.ecsw_instances(n_instances=6, seed=44, side_range=(4, 9), rid_fraction=0.05)


def test_batched_ecsw_matches_oracle(instances):
    from psierra_amd.ecsw import BatchedECSW

    banner("ecsw batched")
    b = BatchedECSW(instances)
    ur = np.stack([inst["u_r"] for inst in instances])
    f, K = b.evaluate(ur)
    f, K = f.cpu().numpy(), K.cpu().numpy()
    for i, inst in enumerate(instances):
        m = inst["mesh"]
        fr, Kr = OE.evaluate_forces_and_stiffness(m.nodes, m.elements, inst["basis"], inst["rid"],
                                                  inst["weights"], inst["mu"], inst["lam"], inst["u_r"])
        ek, ef = rel_err(K[i], Kr), rel_err(f[i], fr)
        print(f"instance {i} sides {inst['sides']}: K_r rel {ek:.2e}, f_r rel {ef:.2e}")
        assert ek < 1e-4 and ef < 1e-4
        assert np.array_equal(K[i], K[i].T)


@pytest.mark.parametrize("rank", [8, 30, 64])
def test_basis_ranks_match_oracle(rank):
    """Ranks below the 64-wide matrix-core tiles (zero-padded columns) and the
    full 64, three uneven instances each, against the oracle."""
    from psierra_amd.ecsw import BatchedECSW

    insts = W.ecsw_instances(n_instances=3, seed=45 + rank, side_range=(4, 8), rank=rank, rid_fraction=0.05)
    banner(f"ecsw rank {rank}")
    f, K = BatchedECSW(insts).evaluate(np.stack([inst["u_r"] for inst in insts]))
    assert f.shape == (3, rank) and K.shape == (3, rank, rank)
    f, K = f.cpu().numpy(), K.cpu().numpy()
    for i, inst in enumerate(insts):
        m = inst["mesh"]
        fr, Kr = OE.evaluate_forces_and_stiffness(m.nodes, m.elements, inst["basis"], inst["rid"],
                                                  inst["weights"], inst["mu"], inst["lam"], inst["u_r"])
        ek, ef = rel_err(K[i], Kr), rel_err(f[i], fr)
        print(f"rank {rank} instance {i}: K_r rel {ek:.2e}, f_r rel {ef:.2e}")
        assert ek < 1e-4 and ef < 1e-4
        assert np.array_equal(K[i], K[i].T)


def test_model_with_reference_constructor():
    """HyperReducedECSWModel(domain, basis, rid, dirichlet, damping, f_ext_base)
    as model.py:50-59 builds it: M_r, K_r_0, C_r and the reduced base load
    against the oracle's assembled mass and linear stiffness projected on the
    free-DOF basis; evaluate_forces_and_stiffness against the oracle and equal
    to the one-instance batch."""
    from types import SimpleNamespace

    from oracle import assembly as A
    from psierra_amd.boundary import DirichletConstraint, DirichletManager, select_nodes_by_plane
    from psierra_amd.domain import SolidDomain
    from psierra_amd.ecsw import BatchedECSW, HyperReducedECSWModel
    from psierra_amd.element import GpuTet4TotalLagrangian
    from psierra_amd.material import NeoHookeanField

    banner("ecsw model constructor")
    wl = W.cantilever_tet4()
    mesh = wl.mesh
    rng = np.random.default_rng(46)
    r = 24
    Phi = np.linalg.qr(rng.standard_normal((mesh.n_dofs, r)))[0]
    basis = SimpleNamespace(basis_matrix=Phi, n_modes=r)
    idx = np.sort(rng.choice(mesh.n_elements, 90, replace=False))
    rid = SimpleNamespace(element_indices=idx, weights=rng.uniform(0.5, 2.0, idx.size))
    rho = 7.5
    dom = SolidDomain(mesh, NeoHookeanField(wl.mu, wl.lam, mass_density=rho), GpuTet4TotalLagrangian(None))
    mgr = DirichletManager(mesh.n_dofs)
    mgr.add_constraint(DirichletConstraint(select_nodes_by_plane(mesh, "yz", "min"), (0, 1, 2), 0.0))
    f_ext = np.zeros(mesh.n_dofs)
    tip = select_nodes_by_plane(mesh, "yz", "max")
    f_ext[3 * tip + 2] = -0.01 / len(tip)
    damping = SimpleNamespace(alpha=0.3, beta=2e-3)
    model = HyperReducedECSWModel(dom, basis, rid, dirichlet=mgr, damping=damping, f_ext_base=f_ext)
    free = mgr.free_dofs
    Pf = Phi[free]
    M = A.mass_matrix(mesh.nodes, mesh.elements, rho)
    Kl = A.linear_stiffness(mesh.nodes, mesh.elements, wl.mu, wl.lam)
    M_r = Pf.T @ (M[np.ix_(free, free)] @ Pf)
    K_r0 = Pf.T @ (Kl[np.ix_(free, free)] @ Pf)
    assert model.n_modes == r and model.n_elements_reduced == idx.size
    assert rel_err(model.M_r, M_r) < 1e-12 and np.array_equal(model.M_r, model.M_r.T)
    assert rel_err(model.K_r_0, K_r0) < 1e-11 and np.array_equal(model.K_r_0, model.K_r_0.T)
    assert rel_err(model.C_r, 0.3 * M_r + 2e-3 * K_r0) < 1e-11
    assert rel_err(model.f_r_base, Pf.T @ f_ext[free]) < 1e-14
    u_r = rng.standard_normal(r)
    u_r *= 0.01 / np.abs(Phi @ u_r).max()
    f1, K1 = model.evaluate_forces_and_stiffness(u_r)
    fr, Kr = OE.evaluate_forces_and_stiffness(mesh.nodes, mesh.elements, Phi, idx, rid.weights,
                                              wl.mu[idx], wl.lam[idx], u_r)
    assert rel_err(K1, Kr) < 1e-4 and rel_err(f1, fr) < 1e-4
    inst = {"mesh": mesh, "basis": Phi, "rid": idx, "weights": rid.weights, "mu": wl.mu[idx], "lam": wl.lam[idx]}
    f2, K2 = BatchedECSW([inst]).evaluate(u_r[None])
    assert np.array_equal(f1, f2[0].cpu().numpy()) and np.array_equal(K1, K2[0].cpu().numpy())
    res, _ = model.evaluate_residual(u_r)
    assert np.array_equal(res, f1 - model.f_r_base)
    u = model.reconstruct_displacement(u_r)
    assert np.array_equal(u[free], Pf @ u_r) and np.all(u[mgr.prescribed_dofs] == 0.0)
    with pytest.raises(ValueError, match="must match reduced dimension"):
        model.evaluate_forces_and_stiffness(np.zeros(r + 1))


def test_rid_sizes_at_batch_and_chunk_edges(instances):
    """RIDs of 1, 5, 8, 17, 128 and 129 elements: below one matrix-core batch,
    exactly one, across a 16-element row batch, exactly one 128-element work
    item and one past it; each equal to the oracle."""
    from psierra_amd.ecsw import BatchedECSW

    banner("ecsw rid edges")
    big = max(instances, key=lambda inst: inst["mesh"].n_elements)
    rng = np.random.default_rng(7)
    cases = []
    for n in (1, 5, 8, 17, 128, 129):
        assert n <= big["mesh"].n_elements
        inst = dict(big)
        inst["rid"] = np.sort(rng.choice(big["mesh"].n_elements, n, replace=False)).astype(np.int64)
        inst["weights"] = rng.uniform(0.5, 2.0, n)
        cases.append(inst)
    b = BatchedECSW(cases)
    f, K = b.evaluate(np.stack([c["u_r"] for c in cases]))
    f, K = f.cpu().numpy(), K.cpu().numpy()
    for i, c in enumerate(cases):
        m = c["mesh"]
        fr, Kr = OE.evaluate_forces_and_stiffness(m.nodes, m.elements, c["basis"], c["rid"], c["weights"],
                                                  c["mu"], c["lam"], c["u_r"])
        ek, ef = rel_err(K[i], Kr), rel_err(f[i], fr)
        print(f"rid {len(c['rid'])}: K_r rel {ek:.2e}, f_r rel {ef:.2e}")
        assert ek < 1e-4 and ef < 1e-4
        assert np.array_equal(K[i], K[i].T)
