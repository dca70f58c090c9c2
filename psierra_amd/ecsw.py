"""Batched ECSW hyper-reduced evaluation on the GPU.

``BatchedECSW`` evaluates, for many independent instances in one persistent
launch (``psx_ecsw_evaluate``),
  f_r = sum_e w_e Phi_e^T f_e(Phi_e u_r),
  K_r = sym(sum_e w_e Phi_e^T k_e(Phi_e u_r) Phi_e),
the online step of HyperReducedECSWModel.evaluate_forces_and_stiffness
(src/psierra/hyperreduction/model.py:254-318).  Phi_e (the basis rows of an
element's 12 DOFs) is gathered once at construction, as the reference caches
it (model.py:185-213).  Tet4 Neo-Hookean, reduced rank 1..64 (one rank per
batch).  ``HyperReducedECSWModel`` is the reference model with its own
constructor, evaluated through a one-instance batch.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, lib, ptr
from .engine import _torch, current_stream_handle

MAX_RANK = 64


def element_basis_rows(basis: np.ndarray, conn: np.ndarray) -> np.ndarray:
    """Phi_e (m, 12, r) from a basis (3N, r) and Tet4 connectivity (m, 4)."""
    dofs = (3 * conn[:, :, None] + np.arange(3)[None, None, :]).reshape(conn.shape[0], 12)
    return basis[dofs]


class BatchedECSW:
    """instances: dicts with mesh, basis (n_dofs, r), rid, weights, mu, lam
    (scalars or one value per RID element); every basis has the same rank r,
    1 <= r <= 64."""

    def __init__(self, instances) -> None:
        torch = _torch()
        Xs, phis, ws, mus, lams, offs = [], [], [], [], [], [0]
        rank = None
        for inst in instances:
            mesh = inst["mesh"]
            rid = np.asarray(inst["rid"], dtype=np.int64).ravel()
            w = np.asarray(inst["weights"], dtype=np.float64).ravel()
            if rid.shape != w.shape:
                raise ValueError(
                    f"Element indices length ({rid.size}) must match weights length ({w.size})."
                )
            basis = np.asarray(inst["basis"], dtype=np.float64)
            if basis.ndim != 2 or basis.shape[0] != mesh.n_dofs:
                raise ValueError(f"basis must be (n_dofs={mesh.n_dofs}, r), got {basis.shape}")
            if rank is None:
                rank = basis.shape[1]
            if basis.shape[1] != rank:
                raise ValueError("every instance of a batch needs the same basis rank")
            if not 1 <= rank <= MAX_RANK:
                raise ValueError(f"the GPU ECSW projection takes a basis rank of 1..{MAX_RANK}, got {rank}")
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
        self.rank = int(rank)
        X = np.ascontiguousarray(np.concatenate(Xs))
        phi = np.ascontiguousarray(np.concatenate(phis))
        w = np.ascontiguousarray(np.concatenate(ws))
        mu = np.ascontiguousarray(np.concatenate(mus))
        lam = np.ascontiguousarray(np.concatenate(lams))
        offs = np.asarray(offs, dtype=np.int64)
        h = ctypes.c_void_p()
        check(lib().psx_ecsw_create(self.n_instances, ptr(offs), ptr(X), ptr(phi), ptr(w), ptr(mu), ptr(lam),
                                    self.rank, ctypes.byref(h)))
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
        r = self.rank
        u = torch.as_tensor(u_r, dtype=torch.float64, device=self.device).contiguous()
        if u.shape != (self.n_instances, r):
            raise ValueError(f"u_r must be ({self.n_instances}, {r})")
        if f_r is None:
            f_r = torch.empty((self.n_instances, r), dtype=torch.float64, device=self.device)
        if K_r is None:
            K_r = torch.empty((self.n_instances, r, r), dtype=torch.float64, device=self.device)
        check(lib().psx_ecsw_evaluate(self._h, ptr(u), ptr(f_r), ptr(K_r), current_stream_handle(),
                                      ptr(self.status)))
        if check_status:
            st = self.status.cpu().numpy()
            if st[0] != 0:
                raise ValueError(f"Element inversion detected in RID element {int(st[1])}: det(F) <= 0.")
        return f_r, K_r


def _sym(M: np.ndarray) -> np.ndarray:
    return 0.5 * (M + M.T)


class HyperReducedECSWModel:
    """HyperReducedECSWModel (src/psierra/hyperreduction/model.py:35-410) for
    a Tet4 solid domain with a GPU element and a Neo-Hookean material, with
    the reference constructor.

    ``basis``: a PODBasis (``basis_matrix`` (n_dofs, r), ``n_modes``) or a
    plain (n_dofs, r) array; ``rid``: a ReducedIntegrationDomain
    (``element_indices``, ``weights`` > 0).  As the reference does, the
    constructor partitions the DOFs by ``dirichlet`` (model.py:91-108),
    projects the consistent mass and linear stiffness onto the free-DOF basis
    (:113-127), forms the Rayleigh damping matrix (:122-127) and the reduced
    base load (:129-138), and caches Phi_e for every RID element (:185-213).
    ``evaluate_forces_and_stiffness`` runs the FP32 matrix-core projection
    (``psx_ecsw_evaluate``); its per-element material parameters are read
    from the domain's material at construction.
    """

    def __init__(self, domain, basis, rid, dirichlet=None, damping=None, f_ext_base=None,
                 surface_nodes=None) -> None:
        self.domain = domain
        self.basis = basis
        self.rid = rid
        Phi = np.asarray(getattr(basis, "basis_matrix", basis), dtype=np.float64)
        if Phi.ndim != 2 or Phi.shape[0] != domain.n_dofs:
            raise ValueError(f"basis must be (n_dofs={domain.n_dofs}, r), got {Phi.shape}")
        self.r = int(getattr(basis, "n_modes", Phi.shape[1]))
        idx = np.asarray(rid.element_indices, dtype=np.int64).ravel()
        w = np.asarray(rid.weights, dtype=np.float64).ravel()
        if idx.size != w.size:
            raise ValueError(f"Element indices length ({idx.size}) must match weights length ({w.size}).")
        if np.any(w <= 0.0):
            raise ValueError("All ECSW element weights must be strictly positive (w_e > 0).")

        if dirichlet is not None:
            if isinstance(dirichlet, (list, tuple)):
                from .boundary import DirichletManager

                mgr = DirichletManager(domain.n_dofs)
                for c in dirichlet:
                    mgr.add_constraint(c)
                dirichlet = mgr
            self.dirichlet = dirichlet
            self.free_dofs = np.asarray(dirichlet.free_dofs, dtype=np.int64)
            self.prescribed_dofs = np.asarray(dirichlet.prescribed_dofs, dtype=np.int64)
            self.prescribed_values = np.asarray(dirichlet.prescribed_values, dtype=np.float64)
        else:
            self.dirichlet = None
            self.free_dofs = np.arange(domain.n_dofs, dtype=np.int64)
            self.prescribed_dofs = np.array([], dtype=np.int64)
            self.prescribed_values = np.array([], dtype=np.float64)
        self.Phi_f = Phi[self.free_dofs, :]
        ff = np.ix_(self.free_dofs, self.free_dofs)

        M = domain.assemble_mass_matrix().tocsr()
        self.M_r = _sym(self.Phi_f.T @ (M[ff] @ self.Phi_f))
        K_lin = domain.assemble_stiffness_matrix().tocsr()
        self.K_r_0 = _sym(self.Phi_f.T @ (K_lin[ff] @ self.Phi_f))
        self.damping = damping
        alpha = float(getattr(damping, "alpha", 0.0)) if damping is not None else 0.0
        beta = float(getattr(damping, "beta", 0.0)) if damping is not None else 0.0
        self.C_r = _sym(alpha * self.M_r + beta * self.K_r_0)
        if f_ext_base is not None:
            f_f = np.asarray(f_ext_base, dtype=np.float64).ravel()[self.free_dofs]
            if self.prescribed_dofs.size and np.any(self.prescribed_values != 0.0):
                f_f = f_f - K_lin[np.ix_(self.free_dofs, self.prescribed_dofs)] @ self.prescribed_values
            self.f_r_base = self.Phi_f.T @ f_f
        else:
            self.f_r_base = np.zeros(self.r)

        if surface_nodes is not None and hasattr(basis, "extract_surface_basis"):
            self.surface_basis = basis.extract_surface_basis(surface_nodes, is_node_indices=True)
        else:
            self.surface_basis = None

        from .material import lame_arrays

        mesh = domain.mesh
        mu, lam = lame_arrays(domain.element.material, mesh.n_elements)
        self._batch = BatchedECSW([{"mesh": mesh, "basis": Phi, "rid": idx, "weights": w,
                                    "mu": np.asarray(mu)[idx], "lam": np.asarray(lam)[idx]}])

    @property
    def n_modes(self) -> int:
        return self.r

    @property
    def n_elements_reduced(self) -> int:
        return int(np.asarray(self.rid.element_indices).size)

    @property
    def compression_ratio(self) -> float:
        return float(self.n_elements_reduced / self.domain.mesh.n_elements)

    @property
    def speedup_factor(self) -> float:
        return float(self.domain.mesh.n_elements / max(self.n_elements_reduced, 1))

    def evaluate_forces_and_stiffness(self, u_r):
        u = np.asarray(u_r, dtype=np.float64).ravel()
        if u.size != self.r:
            raise ValueError(f"u_r size ({u.size}) must match reduced dimension ({self.r}).")
        f, K = self._batch.evaluate(u[None, :])
        return f[0].cpu().numpy(), K[0].cpu().numpy()

    def evaluate_residual(self, u_r, f_ext_r=None):
        """r_r = f_r_int(u_r) - f_ext_r (default the reduced base load) and
        K_r_t (model.py:320-344)."""
        target = self.f_r_base if f_ext_r is None else np.asarray(f_ext_r, dtype=np.float64).ravel()
        f, K = self.evaluate_forces_and_stiffness(u_r)
        return f - target, K

    def reconstruct_displacement(self, u_r) -> np.ndarray:
        """u = Phi_f u_r on the free DOFs, prescribed values elsewhere (model.py:346-365)."""
        u = np.zeros(self.domain.n_dofs)
        if self.prescribed_dofs.size:
            u[self.prescribed_dofs] = self.prescribed_values
        u[self.free_dofs] = self.Phi_f @ np.asarray(u_r, dtype=np.float64).ravel()
        return u

    def project_load(self, f_ext) -> np.ndarray:
        """Phi_f^T f_ext[free] (model.py:384-387)."""
        return self.Phi_f.T @ np.asarray(f_ext, dtype=np.float64).ravel()[self.free_dofs]


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
                     stiffness_weight: float | None = None):
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
        nwork = int(lib().psx_ecsw_training_work_doubles(plan._h))
        work = torch.empty(max(nwork, 1), dtype=torch.float64, device=dev)
        sumsq = torch.zeros(2, dtype=torch.float64, device=dev)
        status = torch.zeros(2, dtype=torch.int32, device=dev)
        check(lib().psx_ecsw_training(plan._h, ptr(phi_t), r, ptr(snap_t), m, ptr(st["mu"]), ptr(st["lam"]),
                                      int(bool(include_stiffness)), ptr(A), ptr(work), nwork, ptr(sumsq),
                                      current_stream_handle(), ptr(status)))
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
            # the reference's balancing (ecsw.py:350-357) from the kernel's sums of squares
            sq = sumsq.cpu().numpy()
            s_force = float(np.sqrt(sq[0])) / np.sqrt(max(m * r, 1))
            s_stiff = float(np.sqrt(sq[1])) / np.sqrt(max(n_triu, 1))
            if s_force > 1e-12 and s_stiff > 1e-12:
                weight = float(stiffness_weight) if stiffness_weight is not None else 1.0
                A[m * r:] *= (s_force / s_stiff) * weight
        return A, A.sum(dim=1)
