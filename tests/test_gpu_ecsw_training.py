"""GPU ECSW training matrix against the oracle (FP64).

psx_ecsw_training / ECSWTrainingMatrix.build against
oracle.ecsw.training_matrix, the restatement of ECSWTrainingMatrix.build
(src/psierra/hyperreduction/ecsw.py:171-365), which test_oracle_ecsw_training
pins to the assembled force and stiffness.  Tolerance 1e-12 relative
(Frobenius) for A and b; rank 8, 13 (odd: the stiffness triangle's middle
row) and 64 (the kernel's lane-per-mode limit), 70 snapshots to cross the 16-snapshot passes, the stiffness block with and
without a weight, a ragged last group of elements, the sums of squares the
norm balancing uses, and the reference's exception for an inverting
snapshot.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import ecsw as OE
from psierra_amd import workloads as W
from tests.gpu_helpers import banner, rel_err

pytestmark = pytest.mark.gpu


def _domain(wl):
    from psierra_amd.domain import SolidDomain
    from psierra_amd.element import GpuTet4TotalLagrangian
    from psierra_amd.material import NeoHookeanField

    mat = NeoHookeanField(wl.mu, wl.lam)
    return SolidDomain(wl.mesh, mat, GpuTet4TotalLagrangian(mat))


def _inputs(wl, r, m, seed):
    rng = np.random.default_rng(seed)
    n = wl.mesh.n_dofs
    basis = np.linalg.qr(rng.standard_normal((n, r)))[0]
    snaps = 0.002 * rng.standard_normal((n, m)) + rng.uniform(0.2, 1.0, m) * wl.u.reshape(-1, 1)
    return basis, snaps


@pytest.mark.parametrize("r,m,stiff,weight", [(8, 3, False, None), (64, 4, True, None), (8, 70, True, 0.5),
                                               (13, 5, True, None)])
def test_training_matrix_matches_oracle(r, m, stiff, weight):
    from psierra_amd.ecsw import ECSWTrainingMatrix

    wl = W.cantilever_tet4()
    basis, snaps = _inputs(wl, r, m, seed=31 + r + m)
    banner(f"ecsw training r={r} m={m} stiffness={stiff} weight={weight}")
    A, b = ECSWTrainingMatrix.build(_domain(wl), basis, snaps, include_stiffness=stiff, stiffness_weight=weight)
    A_ref, b_ref = OE.training_matrix(wl.mesh.nodes, wl.mesh.elements, basis, snaps, wl.mu, wl.lam,
                                      include_stiffness=stiff, stiffness_weight=weight)
    assert A.shape == A_ref.shape
    assert rel_err(A[: m * r], A_ref[: m * r]) < 1e-12
    if stiff:
        assert rel_err(A[m * r:], A_ref[m * r:]) < 1e-12
    assert rel_err(b, b_ref) < 1e-12


def test_training_matrix_jittered_box():
    from psierra_amd.ecsw import ECSWTrainingMatrix

    wl = W.jittered_box_tet4(n=12, seed=17)
    basis, snaps = _inputs(wl, 32, 6, seed=5)
    banner("ecsw training jittered box 12^3")
    A, b = ECSWTrainingMatrix.build(_domain(wl), basis, snaps, include_stiffness=True)
    A_ref, b_ref = OE.training_matrix(wl.mesh.nodes, wl.mesh.elements, basis, snaps, wl.mu, wl.lam,
                                      include_stiffness=True)
    assert rel_err(A, A_ref) < 1e-12 and rel_err(b, b_ref) < 1e-12


def test_inverting_snapshot_raises_reference_message():
    from psierra_amd.ecsw import ECSWTrainingMatrix

    wl = W.cantilever_tet4()
    basis, snaps = _inputs(wl, 8, 2, seed=3)
    conn = wl.mesh.elements[5]
    X = wl.mesh.nodes[conn]
    bad = snaps[:, 1].reshape(-1, 3)
    bad[conn[0]] = 2.5 * (X[1:].mean(axis=0) - X[0])
    banner("ecsw training inversion")
    with pytest.raises(ValueError, match=r"Element inversion detected: det\(F\) = -?[0-9.]+e[-+][0-9]+ <= 0\."):
        ECSWTrainingMatrix.build(_domain(wl), basis, snaps)


def test_snapshot_argument_checks():
    from psierra_amd.ecsw import ECSWTrainingMatrix

    wl = W.cantilever_tet4()
    basis, _ = _inputs(wl, 8, 1, seed=4)
    dom = _domain(wl)
    with pytest.raises(ValueError, match="Snapshots list cannot be empty"):
        ECSWTrainingMatrix.build(dom, basis, [])
    with pytest.raises(ValueError, match="does not match domain DOF count"):
        ECSWTrainingMatrix.build(dom, basis, np.zeros(7))


def _raw_training(plan, st, basis, snaps, stiff):
    """psx_ecsw_training through the C ABI: unscaled A and the sums of squares."""
    import torch

    from psierra_amd._lib import check, lib, ptr
    from psierra_amd.engine import current_stream_handle

    dev = plan.device
    r, m = basis.shape[1], snaps.shape[1]
    rows = m * r + (r * (r + 1) // 2 if stiff else 0)
    A = torch.full((rows, plan.n_elements), float("nan"), dtype=torch.float64, device=dev)
    nwork = int(lib().psx_ecsw_training_work_doubles(plan._h))
    work = torch.empty(nwork, dtype=torch.float64, device=dev)
    sumsq = torch.zeros(2, dtype=torch.float64, device=dev)
    status = torch.zeros(2, dtype=torch.int32, device=dev)
    phi_t = torch.as_tensor(np.ascontiguousarray(basis), device=dev)
    snap_t = torch.as_tensor(np.ascontiguousarray(snaps), device=dev)
    check(lib().psx_ecsw_training(plan._h, ptr(phi_t), r, ptr(snap_t), m, ptr(st["mu"]), ptr(st["lam"]), int(stiff),
                                  ptr(A), ptr(work), nwork, ptr(sumsq), current_stream_handle(), ptr(status)))
    torch.cuda.synchronize()
    assert status.cpu().tolist()[0] == 0
    return A, sumsq


def test_ragged_last_group_sums_of_squares_and_determinism():
    """E = 2,058 (the last work-group holds 10 of its 16 elements): A equals
    the oracle's unscaled blocks, every entry written; the kernel's sums of
    squares equal the squared Frobenius norms of the two blocks; a second
    call gives the same bits."""
    import torch

    wl = W.jittered_box_tet4(n=7, seed=18)
    E = wl.mesh.n_elements
    assert E % 16 == 10
    r, m = 24, 5
    basis, snaps = _inputs(wl, r, m, seed=7)
    dom = _domain(wl)
    plan = dom.element.plan(wl.mesh)
    st = dom.element._state(wl.mesh)
    banner(f"ecsw training ragged group E={E} r={r} m={m}")
    A, sumsq = _raw_training(plan, st, basis, snaps, True)
    assert not bool(torch.isnan(A).any())
    A_ref, _ = OE.training_matrix(wl.mesh.nodes, wl.mesh.elements, basis, snaps, wl.mu, wl.lam)
    Ah = A.cpu().numpy()
    assert rel_err(Ah[: m * r], A_ref) < 1e-12
    # unscaled stiffness rows: the oracle's blocks before its norm balancing
    Kst = OE.training_matrix(wl.mesh.nodes, wl.mesh.elements, basis, snaps[:, :1], wl.mu, wl.lam,
                             include_stiffness=True, stiffness_weight=None)[0][r:]
    s_f1 = np.linalg.norm(OE.training_matrix(wl.mesh.nodes, wl.mesh.elements, basis, snaps[:, :1], wl.mu,
                                             wl.lam)[0]) / np.sqrt(r)
    s_s = np.linalg.norm(Ah[m * r:]) / np.sqrt(r * (r + 1) // 2)
    assert rel_err(Ah[m * r:] * (s_f1 / s_s), Kst) < 1e-12
    sq = sumsq.cpu().numpy()
    assert abs(sq[0] - np.sum(Ah[: m * r] ** 2)) <= 1e-13 * sq[0]
    assert abs(sq[1] - np.sum(Ah[m * r:] ** 2)) <= 1e-13 * sq[1]
    A2, sumsq2 = _raw_training(plan, st, basis, snaps, True)
    assert torch.equal(A, A2) and torch.equal(sumsq, sumsq2)


def test_inversion_in_the_last_group_raises():
    from psierra_amd.ecsw import ECSWTrainingMatrix

    wl = W.cantilever_tet4()
    basis, snaps = _inputs(wl, 8, 2, seed=3)
    e_bad = wl.mesh.n_elements - 3
    conn = wl.mesh.elements[e_bad]
    X = wl.mesh.nodes[conn]
    bad = snaps[:, 1].reshape(-1, 3)
    bad[conn[0]] = 2.5 * (X[1:].mean(axis=0) - X[0])
    banner(f"ecsw training inversion of element {e_bad} in the last group")
    with pytest.raises(ValueError, match=r"Element inversion detected"):
        ECSWTrainingMatrix.build_device(_domain(wl), basis, snaps)
