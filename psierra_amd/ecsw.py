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


def _snapshot_matrix(snapshots) -> np.ndarray:
    """(n_dofs, m) snapshot matrix with the reference's argument checks
    (src/psierra/hyperreduction/ecsw.py:208-225)."""
    if isinstance(snapshots, np.ndarray):
        if snapshots.ndim == 2:
            return snapshots
        if snapshots.ndim == 1:
            return snapshots.reshape(-1, 1)
        raise ValueError(f"Snapshot array must be 1D or 2D, got shape {snapshots.shape}")
    if isinstance(snapshots, (list, tuple)):
        if len(snapshots) == 0:
            raise ValueError("Snapshots list cannot be empty.")
        return np.column_stack(snapshots)
    raise ValueError(f"Snapshots must be a 2D array or sequence of 1D vectors, got type {type(snapshots)}")


TRAIN_WORK_BYTES = 1 << 30


def train_work_elements(rows: int, n_elements: int, work_bytes: int = TRAIN_WORK_BYTES) -> int:
    """Elements per chunk of the training matrix's element-major work block:
    ``work_bytes`` of FP64 rows.  On the 48^3 box (663k elements, 2,592
    rows) a 1 GiB block is as fast as one block of all elements (13.8 GB)
    on MI355X; much smaller blocks add launches (3,000 elements: +10%)."""
    per = work_bytes // (8 * max(rows, 1))
    return int(max(1, min(n_elements, per - per % 64 if per >= 64 else per)))


class ECSWTrainingMatrix:
    """Virtual-work training matrix for ECSW on the GPU.

    ECSWTrainingMatrix.build (src/psierra/hyperreduction/ecsw.py:171-365),
    single-domain Tet4 path: column e stacks Phi_e^T f_e(u_k) over the
    snapshots k (``psx_ecsw_training``, FP64), optionally followed by the
    upper triangle of Phi_e^T K0_e Phi_e scaled so that the RMS row norms of
    the two blocks match (times ``stiffness_weight``), and b = A 1.  The
    norm balancing and b are formed from the device array with torch
    reductions.  ``free_dofs`` is accepted and unused, as in the reference.
    """

    @staticmethod
    def build(domain, basis, snapshots, free_dofs=None, include_stiffness: bool = False,
              stiffness_weight: float | None = None):
        A, b = ECSWTrainingMatrix.build_device(domain, basis, snapshots, include_stiffness=include_stiffness,
                                               stiffness_weight=stiffness_weight)
        return A.cpu().numpy(), b.cpu().numpy()

    @staticmethod
    def build_device(domain, basis, snapshots, include_stiffness: bool = False,
                     stiffness_weight: float | None = None, work_elements: int | None = None):
        torch = _torch()
        snap = _snapshot_matrix(snapshots)
        n_dofs, m = snap.shape
        if n_dofs != domain.n_dofs:
            raise ValueError(
                f"Snapshot DOF count ({n_dofs}) does not match domain DOF count ({domain.n_dofs})."
            )
        phi = np.asarray(getattr(basis, "basis_matrix", basis), dtype=np.float64)
        r = phi.shape[1]
        element = domain.element
        if not hasattr(element, "plan"):
            raise TypeError("ECSWTrainingMatrix on the GPU needs a GPU element (GpuTet4TotalLagrangian)")
        mesh = domain.mesh
        plan = element.plan(mesh)
        if plan.npe != 4:
            raise NotImplementedError("the GPU training matrix is built for Tet4 meshes")
        st = element._state(mesh)
        dev = plan.device
        E = plan.n_elements
        phi_t = torch.as_tensor(np.ascontiguousarray(phi), device=dev)
        snap_t = torch.as_tensor(np.ascontiguousarray(snap, dtype=np.float64), device=dev)
        n_triu = r * (r + 1) // 2 if include_stiffness else 0
        rows = m * r + n_triu
        A = torch.empty((rows, E), dtype=torch.float64, device=dev)
        chunk = train_work_elements(rows, E) if work_elements is None else int(work_elements)
        work = torch.empty((max(min(chunk, E), 1), rows), dtype=torch.float64, device=dev)
        status = torch.zeros(2, dtype=torch.int32, device=dev)
        check(lib().psx_ecsw_training(plan._h, ptr(phi_t), r, ptr(snap_t), m, ptr(st["mu"]), ptr(st["lam"]),
                                      int(bool(include_stiffness)), ptr(A), ptr(work), chunk,
                                      current_stream_handle(), ptr(status)))
        del work
        stv = status.cpu().numpy()
        if stv[0] != 0:
            # the reference's exception for the lowest inverted element, with
            # det F of the first snapshot that inverts it
            from .engine import INVERSION_MSG, _det3

            conn = plan.elements[int(stv[1])]
            X = plan.nodes[conn]
            Dinv = np.linalg.inv((X[1:] - X[0]).T)
            H = np.vstack([-Dinv.sum(axis=0), Dinv])
            dets = [_det3((X + snap[:, k].reshape(-1, 3)[conn]).T @ H) for k in range(m)]
            raise ValueError(INVERSION_MSG.format(next((d for d in dets if d <= 0.0), min(dets))))
        if include_stiffness:
            s_force = float(torch.linalg.vector_norm(A[: m * r])) / np.sqrt(max(m * r, 1))
            s_stiff = float(torch.linalg.vector_norm(A[m * r:])) / np.sqrt(max(n_triu, 1))
            if s_force > 1e-12 and s_stiff > 1e-12:
                weight = float(stiffness_weight) if stiffness_weight is not None else 1.0
                A[m * r:] *= (s_force / s_stiff) * weight
        return A, A.sum(dim=1)
