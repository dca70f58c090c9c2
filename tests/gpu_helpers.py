"""Helpers shared by the GPU parity tests (oracle comparison, error norms)."""
from __future__ import annotations

import sys

import numpy as np


def banner(msg: str) -> None:
    """Name the case on stdout before a GPU step (flushed, so a crash keeps it)."""
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def rel_err(a, b) -> float:
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def max_rel_err(a, b) -> float:
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


def gpu_assemble(wl, precision="fp64", want_k=True, tile=0):
    """Assemble a workload on the GPU through the public element hook path."""
    import torch

    from psierra_amd.engine import AssemblyPlan

    plan = AssemblyPlan(wl.mesh.nodes, wl.mesh.elements, tile=tile)
    dev = plan.device
    u = torch.as_tensor(wl.u.reshape(-1), device=dev)
    mu = torch.as_tensor(wl.mu, device=dev)
    lam = torch.as_tensor(wl.lam, device=dev)
    f, k = plan.assemble(u, mu, lam, precision=precision, want_k=want_k)
    torch.cuda.synchronize()
    return plan, f.cpu().numpy(), (k.cpu().numpy() if k is not None else None)
