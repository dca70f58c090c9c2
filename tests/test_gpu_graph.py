"""hipGraph replay of the assembly and the tile autotuner.

The graph is replayed several times with new displacements and eager work in
between; every replay must equal the eager assembly bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

from psierra_amd import workloads as W
from tests.gpu_helpers import banner

pytestmark = pytest.mark.gpu


def test_graph_replays_match_eager():
    import torch

    from psierra_amd.engine import AssemblyPlan
    from psierra_amd.graph import GraphedAssembly

    wl = W.jittered_box_tet4(n=8, seed=21)
    plan = AssemblyPlan(wl.mesh.nodes, wl.mesh.elements)
    banner("graph capture")
    g = GraphedAssembly(plan, wl.mu, wl.lam, precision="fp32")
    dev = plan.device
    mu = torch.as_tensor(wl.mu, device=dev)
    lam = torch.as_tensor(wl.lam, device=dev)
    for s in (1.0, 0.5, -0.7, 1.0):
        u = torch.as_tensor(s * wl.u.ravel(), device=dev)
        junk = torch.randn(1000, device=dev).sum()  # eager work between replays
        f_g, k_g = g.replay(u, check_status=True)
        f_g, k_g = f_g.clone(), k_g.clone()
        torch.cuda.synchronize()
        f_e, k_e = plan.assemble(u, mu, lam, precision="fp32")
        torch.cuda.synchronize()
        assert torch.equal(f_g, f_e) and torch.equal(k_g, k_e), s
        del junk


def test_autotune_picks_a_supported_tile():
    import torch

    from psierra_amd.engine import AssemblyPlan
    from psierra_amd.graph import autotune_tile

    wl = W.jittered_box_tet4(n=8, seed=22)
    plan = AssemblyPlan(wl.mesh.nodes, wl.mesh.elements)
    dev = plan.device
    banner("autotune")
    best, t = autotune_tile(plan, torch.as_tensor(wl.u.ravel(), device=dev), torch.as_tensor(wl.mu, device=dev),
                            torch.as_tensor(wl.lam, device=dev))
    assert best in (64, 128, 256) and plan.info["tile"] == best
    assert all(np.isfinite(v) and v > 0 for v in t.values())


@pytest.mark.parametrize("kind", ["tl", "fbar"])
def test_element_autotuned_tile(kind):
    from psierra_amd.element import GpuTet4Fbar, GpuTet4TotalLagrangian
    from psierra_amd.material import NeoHookeanField

    wl = W.jittered_box_tet4(n=8, seed=22)
    mu = np.full(wl.mesh.n_elements, 1.0e3)
    mat = NeoHookeanField(mu, 5.0 * mu)
    cls = GpuTet4TotalLagrangian if kind == "tl" else GpuTet4Fbar
    banner(f"autotuned tile {kind}")
    auto = cls(mat, precision="fp32", tile="auto")
    fixed = cls(mat, precision="fp32")
    f1, k1 = auto.assemble_device(wl.mesh, wl.u.ravel())
    f2, k2 = fixed.assemble_device(wl.mesh, wl.u.ravel())
    st = auto._state(wl.mesh)
    assert set(st["tile_timings"]) == {64, 128, 256}
    assert st["plan"].info["tile"] == min(st["tile_timings"], key=st["tile_timings"].get)
    # every tile gives the same bits
    assert bool((f1 == f2).all()) and bool((k1 == k2).all())
