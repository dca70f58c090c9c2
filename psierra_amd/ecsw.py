"""Batched ECSW hyper-reduced evaluation on the GPU.

``BatchedECSW`` evaluates, for many independent instances in one persistent
launch (``psx_ecsw_evaluate``),
  f_r = sum_e w_e Phi_e^T f_e(Phi_e u_r),
  K_r = sym(sum_e w_e Phi_e^T k_e(Phi_e u_r) Phi_e),
the online step of HyperReducedECSWModel.evaluate_forces_and_stiffness
(src/psierra/hyperreduction/model.py:254-318).  Phi_e (the basis rows of an
element's 12 DOFs) is gathered once at construction, as the reference caches
it (model.py:185-213).  Tet4 Neo-Hookean, reduced rank 64.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, lib, ptr
from .engine import _torch, current_stream_handle

RANK = 64


def element_basis_rows(basis: np.ndarray, conn: np.ndarray) -> np.ndarray:
    """Phi_e (m, 12, r) from a basis (3N, r) and Tet4 connectivity (m, 4)."""
    dofs = (3 * conn[:, :, None] + np.arange(3)[None, None, :]).reshape(conn.shape[0], 12)
    return basis[dofs]


class BatchedECSW:
    def __init__(self, instances) -> None:
        torch = _torch()
        Xs, phis, ws, mus, lams, offs = [], [], [], [], [], [0]
        for inst in instances:
            mesh = inst["mesh"]
            rid = np.asarray(inst["rid"], dtype=np.int64)
            w = np.asarray(inst["weights"], dtype=np.float64)
            if rid.shape != w.shape:
                raise ValueError("Element indices length must match weights length.")
            basis = np.asarray(inst["basis"], dtype=np.float64)
            if basis.shape != (mesh.n_dofs, RANK):
                raise ValueError(f"basis must be (n_dofs, {RANK})")
            conn = np.asarray(mesh.elements, dtype=np.int64)[rid]
            if conn.shape[1] != 4:
                raise ValueError("BatchedECSW evaluates Tet4 meshes")
            Xs.append(np.asarray(mesh.nodes)[conn])
            phis.append(element_basis_rows(basis, conn))
            ws.append(w)
            mus.append(np.broadcast_to(np.asarray(inst["mu"], dtype=np.float64), rid.shape))
            lams.append(np.broadcast_to(np.asarray(inst["lam"], dtype=np.float64), rid.shape))
            offs.append(offs[-1] + rid.size)
        self.n_instances = len(instances)
        X = np.ascontiguousarray(np.concatenate(Xs))
        phi = np.ascontiguousarray(np.concatenate(phis))
        w = np.ascontiguousarray(np.concatenate(ws))
        mu = np.ascontiguousarray(np.concatenate(mus))
        lam = np.ascontiguousarray(np.concatenate(lams))
        offs = np.asarray(offs, dtype=np.int64)
        h = ctypes.c_void_p()
        check(lib().psx_ecsw_create(self.n_instances, ptr(offs), ptr(X), ptr(phi), ptr(w), ptr(mu), ptr(lam),
                                    RANK, ctypes.byref(h)))
        self._h = h
        buf = (ctypes.c_int64 * 5)()
        check(lib().psx_ecsw_info(h, buf, 5))
        self.info = dict(zip(("n_instances", "n_rid_elements", "n_work_items", "grid", "device_bytes"),
                             (int(v) for v in buf)))
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.status = torch.zeros(2, dtype=torch.int32, device=self.device)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().psx_ecsw_destroy(h)
            except Exception:
                pass
            self._h = None

    def evaluate(self, u_r, f_r=None, K_r=None, check_status=True):
        torch = _torch()
        u = torch.as_tensor(u_r, dtype=torch.float64, device=self.device).contiguous()
        if u.shape != (self.n_instances, RANK):
            raise ValueError(f"u_r must be ({self.n_instances}, {RANK})")
        if f_r is None:
            f_r = torch.empty((self.n_instances, RANK), dtype=torch.float64, device=self.device)
        if K_r is None:
            K_r = torch.empty((self.n_instances, RANK, RANK), dtype=torch.float64, device=self.device)
        check(lib().psx_ecsw_evaluate(self._h, ptr(u), ptr(f_r), ptr(K_r), current_stream_handle(),
                                      ptr(self.status)))
        if check_status:
            st = self.status.cpu().numpy()
            if st[0] != 0:
                raise ValueError(f"Element inversion detected in RID element {int(st[1])}: det(F) <= 0.")
        return f_r, K_r


class HyperReducedECSWModel:
    """Single-instance view with the reference's evaluate_forces_and_stiffness."""

    def __init__(self, mesh, basis, element_indices, weights, mu, lam) -> None:
        self.r = RANK
        self._batch = BatchedECSW(
            [{"mesh": mesh, "basis": basis, "rid": element_indices, "weights": weights, "mu": mu, "lam": lam}]
        )

    def evaluate_forces_and_stiffness(self, u_r):
        u = np.asarray(u_r, dtype=np.float64).ravel()
        if u.size != self.r:
            raise ValueError(f"u_r size ({u.size}) must match reduced dimension ({self.r}).")
        f, K = self._batch.evaluate(u[None, :])
        return f[0].cpu().numpy(), K[0].cpu().numpy()
