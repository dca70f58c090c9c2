"""hipGraph capture of the assembly launches and a tile autotuner.

``GraphedAssembly`` captures one ``psx_assemble`` call (displacement
permutation + assembly kernel, no allocation, no memset, no host sync) into
a HIP graph through torch's graph API; ``replay()`` re-runs it on whatever
the captured input buffers hold, which is what a Newton loop needs: copy the
new iterate into ``u`` and replay.

``GraphedCall`` does the same for any assembly call (the elements'
``graph_assembly`` builds one per mesh for the Newton loop of
``psierra_amd.analysis``).

``autotune_tile`` times the work-group tile sizes the kernels are built for
on the caller's own plan and inputs and keeps the fastest; ``autotune_batch``
does the same for a batch's launch mode (persistent grid or one work-group
per tile).
"""
from __future__ import annotations

import numpy as np

from .engine import AssemblyPlan, _torch


class GraphedCall:
    """A HIP graph of ``launch()``, which reads the device buffer ``u`` and
    writes ``outputs`` (no allocation, memset or host sync inside).
    ``replay(u)`` copies the new iterate into ``u`` and replays; ``check``
    (optional) raises for a bad status after a replay."""

    def __init__(self, launch, u, outputs, check=None, warmup: int = 2):
        torch = _torch()
        self.u = u
        self.outputs = outputs
        self._launch = launch
        self._check = check
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(warmup):
                launch()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            launch()
        torch.cuda.synchronize()

    def replay(self, u=None, check_status: bool = False):
        if u is not None:
            self.u.copy_(u.reshape(-1))
        self.graph.replay()
        if check_status and self._check is not None:
            self._check(self.u)
        return self.outputs


class GraphedAssembly(GraphedCall):
    """``psx_assemble`` of ``plan`` with fixed (mu, lam) captured in a graph;
    ``replay(u)`` returns the device (f, K values) buffers."""

    def __init__(self, plan: AssemblyPlan, mu, lam, precision: str = "fp32", want_k: bool = True, warmup: int = 2):
        torch = _torch()
        dev = plan.device
        self.plan = plan
        self.mu = torch.as_tensor(mu, dtype=torch.float64, device=dev).contiguous()
        self.lam = torch.as_tensor(lam, dtype=torch.float64, device=dev).contiguous()
        self.precision = precision
        self.want_k = want_k
        u = torch.zeros(plan.n_dofs, dtype=torch.float64, device=dev)
        f, k = plan.alloc_outputs(want_k)

        def launch():
            plan.assemble(u, self.mu, self.lam, f=f, kval=k, precision=precision, want_k=want_k,
                          check_status=False)

        super().__init__(launch, u, (f, k), check=plan.raise_on_status, warmup=warmup)

    @property
    def f(self):
        return self.outputs[0]

    @property
    def k(self):
        return self.outputs[1]

    def replay(self, u=None, check_status: bool = False):
        if u is not None:
            u = self.plan._device_f64(u, self.plan.n_dofs, "u")
        return super().replay(u, check_status)


def autotune_tile(plan: AssemblyPlan, u, mu, lam, precision: str = "fp32", reps: int = 5, tiles=None):
    """Pick the fastest work-group tile for this plan; returns {tile: ms}."""
    torch = _torch()
    if tiles is None:
        # a tile holds every node's whole incidence run
        tiles = [t for t in ((64, 128, 256) if plan.npe == 4 else (64, 128)) if t >= plan.info["max_incidences"]]
    f, k = plan.alloc_outputs(True)
    timings = {}
    for tile in tiles:
        plan.set_tile(tile)
        plan.assemble(u, mu, lam, f=f, kval=k, precision=precision)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            plan.assemble(u, mu, lam, f=f, kval=k, precision=precision, check_status=False)
        b.record()
        torch.cuda.synchronize()
        timings[tile] = a.elapsed_time(b) / reps
    best = min(timings, key=timings.get)
    plan.set_tile(best)
    plan.raise_on_status(plan._device_f64(u, plan.n_dofs, "u"))
    return best, timings


def autotune_batch(batch, us, mus=None, lams=None, precision: str = "fp32", reps: int = 5):
    """Pick the faster launch mode of an ``AssemblyBatch`` on the caller's
    inputs (the results are the same bits either way); returns
    (persistent, {mode: ms})."""
    torch = _torch()
    fs, ks = batch.assemble(us, mus, lams, precision=precision)   # checks the inputs and the status
    us = [p._device_f64(u, p.n_dofs, "u") for p, u in zip(batch.plans, us)]
    if mus is not None:
        mus = [p._device_f64(m, p.n_elements, "mu") for p, m in zip(batch.plans, mus)]
        lams = [p._device_f64(m, p.n_elements, "lam") for p, m in zip(batch.plans, lams)]
    launch = batch.launcher(us, mus, lams, fs, ks)
    timings = {}
    for persistent in (False, True):
        batch.set_persistent(persistent)
        launch(precision)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            launch(precision)
        b.record()
        torch.cuda.synchronize()
        timings["persistent" if persistent else "tile_per_group"] = a.elapsed_time(b) / reps
    best = timings["persistent"] < timings["tile_per_group"]
    batch.set_persistent(best)
    batch.raise_on_status(us)
    return best, timings
