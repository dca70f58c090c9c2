"""psx_ecsw_training_noalloc (the allocation-free variant of
psx_ecsw_training) against psx_ecsw_training through the C ABI.

Two calls in a row, the second with a smaller input (fewer elements, lower
rank, fewer snapshots, no stiffness block), each compared with the existing
entry point on the same inputs: the FP64 training matrix A and the sums of
squares must be equal bit for bit (both run the same kernels; the partial
sums sit in the library's static array instead of the caller's work
array).  The first call leaves its partials in the shared array, so the
second call's sums show that each call overwrites what it reads.
"""
from __future__ import annotations

import numpy as np
import pytest

from psierra_amd import workloads as W
from tests.gpu_helpers import banner

pytestmark = pytest.mark.gpu


def _inputs(n, rank, m, seed):
    import torch

    from psierra_amd.engine import AssemblyPlan

    wl = W.jittered_box_tet4(n=n, seed=seed)
    plan = AssemblyPlan(wl.mesh.nodes, wl.mesh.elements)
    dev = plan.device
    rng = np.random.default_rng(seed)
    N = wl.mesh.n_dofs
    phi = np.linalg.qr(rng.standard_normal((N, rank)))[0]
    snaps = 0.3 * wl.u.reshape(-1, 1) * rng.uniform(0.5, 1.5, m) + 1e-4 * rng.standard_normal((N, m))
    t = {
        "phi": torch.as_tensor(np.ascontiguousarray(phi), device=dev),
        "snaps": torch.as_tensor(np.ascontiguousarray(snaps), device=dev),
        "mu": torch.as_tensor(wl.mu, device=dev),
        "lam": torch.as_tensor(wl.lam, device=dev),
    }
    return wl, plan, t


def _train(plan, t, rank, m, stiff, noalloc):
    import torch

    from psierra_amd._lib import check, lib, ptr
    from psierra_amd.engine import current_stream_handle

    rows = m * rank + (rank * (rank + 1) // 2 if stiff else 0)
    dev = plan.device
    A = torch.full((rows, plan.n_elements), float("nan"), dtype=torch.float64, device=dev)
    sumsq = torch.full((2,), float("nan"), dtype=torch.float64, device=dev)
    status = torch.zeros(2, dtype=torch.int32, device=dev)
    if noalloc:
        check(lib().psx_ecsw_training_noalloc(plan._h, ptr(t["phi"]), rank, ptr(t["snaps"]), m, ptr(t["mu"]),
                                              ptr(t["lam"]), int(stiff), ptr(A), ptr(sumsq), current_stream_handle(),
                                              ptr(status)))
    else:
        nwork = int(lib().psx_ecsw_training_work_doubles(plan._h))
        work = torch.empty(nwork, dtype=torch.float64, device=dev)
        check(lib().psx_ecsw_training(plan._h, ptr(t["phi"]), rank, ptr(t["snaps"]), m, ptr(t["mu"]),
                                      ptr(t["lam"]), int(stiff), ptr(A), ptr(work), nwork, ptr(sumsq),
                                      current_stream_handle(), ptr(status)))
    torch.cuda.synchronize()
    assert status.cpu().tolist()[0] == 0
    return A, sumsq


def test_noalloc_training_equals_training_on_two_calls():
    import torch

    cases = [(10, 48, 4, True, 31), (5, 12, 2, False, 32)]   # the second input is smaller
    for n, rank, m, stiff, seed in cases:
        banner(f"ecsw training noalloc box {n}^3 rank {rank} snapshots {m} stiffness {stiff}")
        wl, plan, t = _inputs(n, rank, m, seed)
        A_ref, sq_ref = _train(plan, t, rank, m, stiff, noalloc=False)
        A, sq = _train(plan, t, rank, m, stiff, noalloc=True)
        assert not bool(torch.isnan(A).any()) and not bool(torch.isnan(sq).any())
        assert torch.equal(A, A_ref) and torch.equal(sq, sq_ref), (n, rank, m, stiff)


def test_noalloc_training_argument_checks():
    """Null and out-of-range arguments are refused on the host, before any
    launch."""
    import torch

    from psierra_amd._lib import check, lib, ptr
    from psierra_amd.engine import current_stream_handle

    banner("ecsw training noalloc argument checks")
    wl, plan, t = _inputs(3, 8, 1, 33)
    status = torch.zeros(2, dtype=torch.int32, device=plan.device)
    A = torch.zeros((8, plan.n_elements), dtype=torch.float64, device=plan.device)
    with pytest.raises(Exception, match="ranks 1..64"):
        check(lib().psx_ecsw_training_noalloc(plan._h, ptr(t["phi"]), 65, ptr(t["snaps"]), 1, ptr(t["mu"]),
                                              ptr(t["lam"]), 0, ptr(A), None, current_stream_handle(),
                                              ptr(status)))
    with pytest.raises(Exception, match="null argument"):
        check(lib().psx_ecsw_training_noalloc(plan._h, ptr(t["phi"]), 8, ptr(t["snaps"]), 1, ptr(t["mu"]),
                                              ptr(t["lam"]), 0, None, None, current_stream_handle(),
                                              ptr(status)))
