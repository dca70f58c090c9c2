"""GPU ECSW training matrix against the oracle (FP64).

psx_ecsw_training / ECSWTrainingMatrix.build against
oracle.ecsw.training_matrix, the restatement of ECSWTrainingMatrix.build
(src/psierra/hyperreduction/ecsw.py:171-365), which test_oracle_ecsw_training
pins to the assembled force and stiffness.  Tolerance 1e-12 relative
(Frobenius) for A and b; rank 8 and 64 (the kernel's lane-per-mode limit),
70 snapshots to cross the 64-snapshot pass, the stiffness block with and
without a weight, and the reference's exception for an inverting snapshot.
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


@pytest.mark.parametrize("r,m,stiff,weight", [(8, 3, False, None), (64, 4, True, None), (8, 70, True, 0.5)])
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


@pytest.mark.parametrize("chunk", [1, 37, 100, 4096])
def test_chunked_work_block_gives_identical_matrix(chunk):
    """The elements in chunks of `chunk` columns (not dividing E = 10,368;
    100 is rounded down to 64) give the same bits as one chunk of all
    elements, and the oracle's A."""
    import torch

    from psierra_amd.ecsw import ECSWTrainingMatrix

    wl = W.jittered_box_tet4(n=12, seed=18)
    basis, snaps = _inputs(wl, 16, 3, seed=7)
    dom = _domain(wl)
    E = wl.mesh.n_elements
    banner(f"ecsw training chunk={chunk} of E={E}")
    A_one, b_one = ECSWTrainingMatrix.build_device(dom, basis, snaps, include_stiffness=True, work_elements=E)
    A_chk, b_chk = ECSWTrainingMatrix.build_device(dom, basis, snaps, include_stiffness=True, work_elements=chunk)
    torch.cuda.synchronize()
    assert torch.equal(A_one, A_chk) and torch.equal(b_one, b_chk)
    A_ref, _ = OE.training_matrix(wl.mesh.nodes, wl.mesh.elements, basis, snaps, wl.mu, wl.lam,
                                  include_stiffness=True)
    assert rel_err(A_chk.cpu().numpy(), A_ref) < 1e-12


def test_inversion_in_a_later_chunk_raises():
    from psierra_amd.ecsw import ECSWTrainingMatrix

    wl = W.cantilever_tet4()
    basis, snaps = _inputs(wl, 8, 2, seed=3)
    e_bad = wl.mesh.n_elements - 3
    conn = wl.mesh.elements[e_bad]
    X = wl.mesh.nodes[conn]
    bad = snaps[:, 1].reshape(-1, 3)
    bad[conn[0]] = 2.5 * (X[1:].mean(axis=0) - X[0])
    banner(f"ecsw training inversion of element {e_bad} in the last chunk of 64")
    with pytest.raises(ValueError, match=r"Element inversion detected"):
        ECSWTrainingMatrix.build_device(_domain(wl), basis, snaps, work_elements=64)
    with pytest.raises(ValueError, match="work_elements"):
        ECSWTrainingMatrix.build_device(_domain(wl), basis, snaps, work_elements=0)
