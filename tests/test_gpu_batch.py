"""Batched assembly (psx_assemble_batch): many Tet4 (or many Tet10) meshes
of uneven size, or one mesh repeated for parameter samples, in one launch.

Every instance's f and K must be bit-identical to its own psx_assemble (the
same tile routine runs on the same tiles), in both launch modes (persistent
grid, one work-group per tile), for f + K and f only, FP64 and FP32, with
per-call parameters and with bound materials; one instance is checked
against the oracle's scipy assembly (src/psierra/domain/solid.py:230-325,
src/psierra/utility/assembly.py:37-108) at the stated tolerances; inversion
is reported per instance with the reference's message; a HIP graph of the
batched launch replays to the eager bits.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import assembly as A
from psierra_amd import workloads as W
from psierra_amd._lib import PsxError
from tests.gpu_helpers import banner, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def meshes():
    from psierra_amd.engine import AssemblyPlan

    # uneven sizes: sides 2..7 per axis (252 .. 840 elements), and the
    # second mesh again as a seventh instance (one plan twice in a batch)
    wls = W.batch_boxes_tet4(n_instances=6, seed=21, side_range=(2, 7))
    plans = [AssemblyPlan(w.mesh.nodes, w.mesh.elements) for w in wls]
    return wls + [wls[1]], plans + [plans[1]]


def _inputs(wls, plans, seed=0):
    import torch

    dev = plans[0].device
    rng = np.random.default_rng(seed)
    us, mus, lams = [], [], []
    for w in wls:
        us.append(torch.as_tensor(w.u.reshape(-1), device=dev))
        mu = w.mu * np.exp(0.3 * rng.standard_normal(w.mu.size))   # per-element values
        mus.append(torch.as_tensor(mu, device=dev))
        lams.append(torch.as_tensor(5.0 * mu, device=dev))
    return us, mus, lams


@pytest.mark.parametrize("persistent", [True, False])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_bit_identical_to_single_plans(meshes, precision, persistent):
    import torch

    from psierra_amd.engine import AssemblyBatch

    wls, plans = meshes
    us, mus, lams = _inputs(wls, plans)
    banner(f"batch: {len(plans)} uneven meshes {precision} persistent={persistent}")
    batch = AssemblyBatch(plans, persistent=persistent)
    assert batch.info["n_instances"] == len(plans) and batch.info["persistent"] == int(persistent)
    assert batch.info["n_tiles"] == sum(p.info["n_tiles"] for p in plans)
    fs, ks = batch.assemble(us, mus, lams, precision=precision)
    fo, _ = batch.assemble(us, mus, lams, precision=precision, want_k=False)
    torch.cuda.synchronize()
    for i, p in enumerate(plans):
        f1, k1 = p.assemble(us[i], mus[i], lams[i], precision=precision)
        f2, _ = p.assemble(us[i], mus[i], lams[i], precision=precision, want_k=False)
        assert torch.equal(fs[i], f1) and torch.equal(ks[i], k1), i
        assert torch.equal(fo[i], f2), i
    # instance 0 against the oracle
    w = wls[0]
    f_ref, K_ref = A.assemble(w.mesh.nodes, w.mesh.elements, w.u, mus[0].cpu().numpy(), lams[0].cpu().numpy())
    K = K_ref.copy()
    K.data = ks[0].cpu().numpy()
    indptr, indices = plans[0].pattern()
    assert np.array_equal(indptr, K_ref.indptr) and np.array_equal(indices, K_ref.indices)
    tol_k, tol_f = (1e-12, 1e-11) if precision == "fp64" else (1e-5, 1e-5)
    assert rel_err(K.data, K_ref.data) < tol_k
    assert rel_err(fs[0].cpu().numpy(), f_ref) < tol_f


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_parameter_samples_on_one_plan(meshes, precision):
    """One plan repeated: each sample its own displacement and parameters."""
    import torch

    from psierra_amd.engine import AssemblyBatch

    wls, plans = meshes
    plan, w = plans[3], wls[3]
    dev = plan.device
    rng = np.random.default_rng(5)
    S = 5
    us = [torch.as_tensor((w.u * rng.uniform(0.2, 1.0)).reshape(-1), device=dev) for _ in range(S)]
    mus = [torch.as_tensor(w.mu * rng.uniform(0.5, 2.0), device=dev) for _ in range(S)]
    lams = [5.0 * m for m in mus]
    banner(f"batch: {S} parameter samples on one plan {precision}")
    batch = AssemblyBatch([plan] * S)
    fs, ks = batch.assemble(us, mus, lams, precision=precision)
    torch.cuda.synchronize()
    for s in range(S):
        f1, k1 = plan.assemble(us[s], mus[s], lams[s], precision=precision)
        assert torch.equal(fs[s], f1) and torch.equal(ks[s], k1), s
    assert not torch.equal(ks[0], ks[1])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_bound_materials(precision):
    import torch

    from psierra_amd.engine import AssemblyBatch, AssemblyPlan

    wls = W.batch_boxes_tet4(n_instances=4, seed=22, side_range=(3, 6))
    plans = [AssemblyPlan(w.mesh.nodes, w.mesh.elements) for w in wls]
    us, mus, lams = _inputs(wls, plans, seed=1)
    banner(f"batch: bound materials {precision}")
    batch = AssemblyBatch(plans)
    with pytest.raises(ValueError, match="psx_plan_bind_material on every plan"):
        batch.assemble(us, precision=precision)
    for p, m, l in zip(plans, mus, lams):
        p.bind_material(m, l)
    fs, ks = batch.assemble(us, precision=precision)
    torch.cuda.synchronize()
    for i, p in enumerate(plans):
        f1, k1 = p.assemble(us[i], None, None, precision=precision)
        assert torch.equal(fs[i], f1) and torch.equal(ks[i], k1), i
    # rebinding one plan refreshes the batch's table before the next launch
    plans[2].bind_material(2.0 * mus[2], 2.0 * lams[2])
    fs, ks = batch.assemble(us, precision=precision)
    f1, k1 = plans[2].assemble(us[2], None, None, precision=precision)
    assert torch.equal(fs[2], f1) and torch.equal(ks[2], k1)


def test_batch_retile_refresh_and_refusals(meshes):
    import torch

    from psierra_amd.engine import AssemblyBatch, AssemblyPlan

    wls, plans = meshes
    sub = [AssemblyPlan(w.mesh.nodes, w.mesh.elements) for w in wls[:3]]
    us, mus, lams = _inputs(wls[:3], sub, seed=2)
    banner("batch: retile refresh")
    batch = AssemblyBatch(sub)
    for p in sub:
        p.set_tile(128)
    fs, ks = batch.assemble(us, mus, lams, precision="fp64")
    batch.refresh_info()
    assert batch.info["tile"] == 128 and batch.info["n_tiles"] == sum(p.info["n_tiles"] for p in sub)
    for i, p in enumerate(sub):
        f1, k1 = p.assemble(us[i], mus[i], lams[i], precision="fp64")
        assert torch.equal(fs[i], f1) and torch.equal(ks[i], k1), i
    sub[0].set_tile(256)
    with pytest.raises(ValueError, match="another tile size"):
        batch.assemble(us, mus, lams, precision="fp64")
    # mixed element types and wide plans are refused at creation
    from psierra_amd.mesh import tet4_to_tet10

    X10, c10 = tet4_to_tet10(wls[0].mesh.nodes, wls[0].mesh.elements)
    with pytest.raises(ValueError, match="another element type"):
        AssemblyBatch([plans[0], AssemblyPlan(X10, c10)])
    from tests.test_gpu_edge_cases import _fan_and_box

    fan = _fan_and_box(80, seed=3)
    with pytest.raises(PsxError, match="wide"):
        AssemblyBatch([plans[0], AssemblyPlan(fan.mesh.nodes, fan.mesh.elements)])


def test_batch_inversion_per_instance(meshes):
    import torch

    from psierra_amd.engine import AssemblyBatch

    wls, plans = meshes
    us, mus, lams = _inputs(wls, plans, seed=3)
    w = wls[4]
    u = w.u.copy()
    conn = w.mesh.elements[11]
    X = w.mesh.nodes[conn]
    u[conn[0]] = 2.5 * (X[1:].mean(axis=0) - X[0])   # node 0 through the opposite face
    us[4] = torch.as_tensor(u.reshape(-1), device=plans[0].device)
    banner("batch: inversion in instance 4")
    batch = AssemblyBatch(plans)
    with pytest.raises(ValueError, match=r"Element inversion detected: det\(F\) = -?[0-9.]+e[-+][0-9]+ <= 0\."):
        batch.assemble(us, mus, lams, precision="fp64")
    st = batch.status.cpu().numpy().reshape(-1, 2)
    assert st[4, 0] == 1 and all(st[i, 0] == 0 for i in range(len(plans)) if i != 4)
    plans[4].assemble(us[4], mus[4], lams[4], precision="fp64", check_status=False)
    torch.cuda.synchronize()
    assert st[4, 1] == plans[4].status.cpu().numpy()[1]


def test_batch_graph_replay(meshes):
    import torch

    from psierra_amd.engine import AssemblyBatch
    from psierra_amd.graph import GraphedCall

    wls, plans = meshes
    us, mus, lams = _inputs(wls, plans, seed=4)
    batch = AssemblyBatch(plans)
    fs, ks = batch.alloc_outputs(True)
    banner("batch: graph capture and replay")
    eager_f, eager_k = batch.assemble(us, mus, lams, precision="fp32")
    eager_f = [f.clone() for f in eager_f]
    eager_k = [k.clone() for k in eager_k]
    launch = batch.launcher(us, mus, lams, fs, ks)   # pointer tables built outside the capture
    g = GraphedCall(lambda: launch("fp32"), us[0], (fs, ks))
    for rep in range(3):
        for f, k in zip(fs, ks):
            f.zero_()
            k.zero_()
        torch.cuda.synchronize()
        g.graph.replay()
        torch.cuda.synchronize()
        for i in range(len(plans)):
            assert torch.equal(fs[i], eager_f[i]) and torch.equal(ks[i], eager_k[i]), (rep, i)


def test_batch_autotune_keeps_results(meshes):
    import torch

    from psierra_amd.engine import AssemblyBatch
    from psierra_amd.graph import autotune_batch

    wls, plans = meshes
    us, mus, lams = _inputs(wls, plans, seed=6)
    batch = AssemblyBatch(plans)
    banner("batch: autotune of the launch mode")
    f0, k0 = batch.assemble(us, mus, lams, precision="fp32")
    f0 = [f.clone() for f in f0]
    k0 = [k.clone() for k in k0]
    best, timings = autotune_batch(batch, us, mus, lams, precision="fp32", reps=2)
    assert set(timings) == {"persistent", "tile_per_group"} and batch.info["persistent"] == int(best)
    fs, ks = batch.assemble(us, mus, lams, precision="fp32")
    torch.cuda.synchronize()
    for i in range(len(plans)):
        assert torch.equal(fs[i], f0[i]) and torch.equal(ks[i], k0[i]), i


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_tet10_batch_bit_identical_to_single_plans(precision):
    """Tet10 meshes of uneven size (the Tet4 boxes with mid-edge nodes), one
    of them twice: f + K and f only equal each plan's own assembly bit for
    bit; instance 0 against the oracle's scipy assembly."""
    import torch

    from psierra_amd.engine import AssemblyBatch, AssemblyPlan
    from psierra_amd.mesh import tet4_to_tet10

    wls = W.batch_boxes_tet4(n_instances=4, seed=23, side_range=(2, 5))
    dev = None
    meshes, plans, us, mus, lams = [], [], [], [], []
    rng = np.random.default_rng(7)
    for w in wls:
        X10, c10 = tet4_to_tet10(w.mesh.nodes, w.mesh.elements)
        u = 0.02 * np.sin(3.0 * X10) * rng.uniform(0.5, 1.0)
        plan = AssemblyPlan(X10, c10)
        dev = plan.device
        mu = w.mu[0] * np.exp(0.2 * rng.standard_normal(c10.shape[0]))
        meshes.append((X10, c10, u, mu))
        plans.append(plan)
        us.append(torch.as_tensor(u.reshape(-1), device=dev))
        mus.append(torch.as_tensor(mu, device=dev))
        lams.append(torch.as_tensor(5.0 * mu, device=dev))
    plans.append(plans[1])
    us.append(0.5 * us[1])
    mus.append(mus[1])
    lams.append(lams[1])
    banner(f"batch: {len(plans)} Tet10 meshes {precision}")
    batch = AssemblyBatch(plans)
    assert batch.info["n_tiles"] == sum(p.info["n_tiles"] for p in plans)
    fs, ks = batch.assemble(us, mus, lams, precision=precision)
    fo, _ = batch.assemble(us, mus, lams, precision=precision, want_k=False)
    torch.cuda.synchronize()
    for i, p in enumerate(plans):
        f1, k1 = p.assemble(us[i], mus[i], lams[i], precision=precision)
        f2, _ = p.assemble(us[i], mus[i], lams[i], precision=precision, want_k=False)
        assert torch.equal(fs[i], f1) and torch.equal(ks[i], k1), i
        assert torch.equal(fo[i], f2), i
    X10, c10, u, mu = meshes[0]
    f_ref, K_ref = A.assemble(X10, c10, u, mu, 5.0 * mu)
    tol_k, tol_f = (1e-12, 1e-11) if precision == "fp64" else (1e-5, 1e-5)
    assert rel_err(ks[0].cpu().numpy(), K_ref.data) < tol_k
    assert rel_err(fs[0].cpu().numpy(), f_ref) < tol_f
    with pytest.raises(ValueError, match="Tet10 plans bind no material"):
        batch.assemble(us, precision=precision)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_vjp_bit_identical_to_single_plans(meshes, precision):
    """The adjoint of every instance in one launch: the bits of each plan's
    own vjp (one plan twice, with different inputs); instance 0's K_T w
    against the oracle's assembled K_T times w."""
    import torch

    from psierra_amd.engine import AssemblyBatch

    wls, plans = meshes
    us, mus, lams = _inputs(wls, plans, seed=8)
    rng = np.random.default_rng(12)
    ws = [torch.as_tensor(rng.standard_normal(p.n_dofs), device=p.device) for p in plans]
    banner(f"batch: adjoint of {len(plans)} instances {precision}")
    batch = AssemblyBatch(plans)
    gus, gmus, glams = batch.vjp(us, mus, lams, ws, precision=precision)
    gus2, none_mu, none_lam = batch.vjp(us, mus, lams, ws, precision=precision, want_params=False)
    assert none_mu is None and none_lam is None
    torch.cuda.synchronize()
    for i, p in enumerate(plans):
        gu, gmu, glam = p.vjp(us[i], mus[i], lams[i], ws[i], precision=precision)
        assert torch.equal(gus[i], gu) and torch.equal(gmus[i], gmu) and torch.equal(glams[i], glam), i
        assert torch.equal(gus2[i], gu), i
    w = wls[0]
    _, K_ref = A.assemble(w.mesh.nodes, w.mesh.elements, w.u, mus[0].cpu().numpy(), lams[0].cpu().numpy())
    ref = K_ref @ ws[0].cpu().numpy()
    assert rel_err(gus[0].cpu().numpy(), ref) < (1e-12 if precision == "fp64" else 1e-5)
