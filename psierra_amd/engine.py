"""Host-side engine objects: the assembly plan and stream-ordered launches.

``AssemblyPlan`` owns a ``psx_plan`` (pattern, incidence map, node order,
reference coordinates on the device) for one mesh.  ``assemble`` evaluates
f_int and the K_T values into caller-visible torch tensors on the current
HIP stream.  No work falls back to the CPU: if the library or a GPU is
missing the calls raise.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib, ptr

PRECISIONS = {"fp64": _lib.PSX_FP64, "fp32": _lib.PSX_FP32}

INVERSION_MSG = "Element inversion detected: det(F) = {:.4e} <= 0."
MODELS = {"neohookean": 0, "mooney_rivlin": 1, "yeoh": 2}


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("psierra_amd needs a HIP GPU (torch.cuda.is_available() is False)")
    return torch


def current_stream_handle():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


# Host copies above this many bytes go through page-locked memory: a pageable
# device-to-host copy runs at ~4 GB/s on the MI355X boxes, a page-locked one
# at ~56 GB/s (1.1 GB of K values: 253 ms against 20 ms).
PINNED_MIN_BYTES = 1 << 20


def host_array(t) -> np.ndarray:
    """numpy copy of a device tensor, through page-locked memory for large
    ones; the page-locked block returns to torch's host cache when the array
    (which holds the tensor) is freed."""
    torch = _torch()
    if t.numel() * t.element_size() < PINNED_MIN_BYTES:
        return t.cpu().numpy()
    h = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    h.copy_(t, non_blocking=True)
    torch.cuda.current_stream().synchronize()
    return h.numpy()


def device_f64(x, device):
    """float64 device copy of a host array, staged through page-locked memory
    when large (the copy is ordered on the current stream)."""
    torch = _torch()
    a = np.ascontiguousarray(x, dtype=np.float64)
    if a.nbytes < PINNED_MIN_BYTES:
        return torch.as_tensor(a, device=device)
    return torch.from_numpy(a).pin_memory().to(device, non_blocking=True)


def precision_code(precision: str) -> int:
    try:
        return PRECISIONS[precision]
    except KeyError:
        raise ValueError(f"precision must be one of {sorted(PRECISIONS)}, got {precision!r}") from None


def pattern_host(elements: np.ndarray, n_nodes: int) -> tuple[np.ndarray, np.ndarray]:
    """Structural CSR (indptr, indices) of the global matrix, built natively on the host."""
    conn = np.ascontiguousarray(elements, dtype=np.int64)
    E, npe = conn.shape
    nnz = ctypes.c_int64(0)
    check(lib().psx_pattern_nnz(ptr(conn), E, npe, n_nodes, ctypes.byref(nnz)))
    indptr = np.empty(3 * n_nodes + 1, dtype=np.int32)
    indices = np.empty(nnz.value, dtype=np.int32)
    check(lib().psx_pattern_build(ptr(conn), E, npe, n_nodes, ptr(indptr), ptr(indices)))
    return indptr, indices


def fbar_pattern_host(elements: np.ndarray, n_nodes: int) -> tuple[np.ndarray, np.ndarray]:
    """CSR (indptr, indices) of Tet4Fbar's K_dev + K_vol (node pairs sharing
    an element or a neighbouring node), built natively on the host."""
    conn = np.ascontiguousarray(elements, dtype=np.int64)
    E, npe = conn.shape
    if npe != 4:
        raise ValueError("the F-bar pattern is defined for Tet4 connectivity")
    nnz = ctypes.c_int64(0)
    check(lib().psx_fbar_pattern_nnz(ptr(conn), E, n_nodes, ctypes.byref(nnz)))
    indptr = np.empty(3 * n_nodes + 1, dtype=np.int32)
    indices = np.empty(nnz.value, dtype=np.int32)
    check(lib().psx_fbar_pattern_build(ptr(conn), E, n_nodes, ptr(indptr), ptr(indices)))
    return indptr, indices


def _det3(M):
    return float(np.linalg.det(M))


def inversion_error(nodes: np.ndarray, elements: np.ndarray, u_host_rows, e: int) -> ValueError:
    """The reference's exception for element e (det F of its first failing point)."""
    conn = elements[e]
    X = nodes[conn]
    u = u_host_rows(conn)
    if conn.size == 4:
        D = (X[1:] - X[0]).T
        H = np.vstack([-np.linalg.inv(D).sum(axis=0), np.linalg.inv(D)])
        return ValueError(INVERSION_MSG.format(_det3((X + u).T @ H)))
    from .workloads import _tet10_dN, tet10_quadrature_points

    for qp in tet10_quadrature_points():
        dN = _tet10_dN(*qp)
        H = dN @ np.linalg.inv(X.T @ dN)
        d = _det3((X + u).T @ H)
        if d <= 0.0:
            return ValueError(INVERSION_MSG.format(d))
    return ValueError(INVERSION_MSG.format(float("nan")))


class AssemblyPlan:
    """Device plan for one mesh (pattern + incidence map + node order)."""

    def __init__(self, nodes: np.ndarray, elements: np.ndarray, tile: int = 0):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        self.elements = np.ascontiguousarray(elements, dtype=np.int64)
        if self.nodes.ndim != 2 or self.nodes.shape[1] != 3:
            raise ValueError("nodes must be (N, 3)")
        torch = _torch()
        self.device = torch.device("cuda", torch.cuda.current_device())
        E, npe = self.elements.shape
        handle = ctypes.c_void_p()
        check(
            lib().psx_plan_create(
                ptr(self.elements), E, npe, ptr(self.nodes), self.nodes.shape[0], tile, ctypes.byref(handle)
            )
        )
        self._h = handle
        self._indptr = None
        self._indices = None
        self.status = torch.zeros(2, dtype=torch.int32, device=self.device)
        self.refresh_info()

    def refresh_info(self):
        buf = (ctypes.c_int64 * len(_lib.INFO_KEYS))()
        check(lib().psx_plan_info(self._h, buf, len(_lib.INFO_KEYS)))
        self.info = dict(zip(_lib.INFO_KEYS, (int(v) for v in buf)))
        self.n_elements = self.info["n_elements"]
        self.n_nodes = self.info["n_nodes"]
        self.npe = self.info["npe"]
        self.nnz = self.info["nnz"]
        self.n_dofs = 3 * self.n_nodes

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().psx_plan_destroy(h)
            except Exception:
                pass
            self._h = None

    def set_tile(self, tile: int) -> None:
        check(lib().psx_plan_set_tile(self._h, tile))
        self.refresh_info()

    def pattern(self) -> tuple[np.ndarray, np.ndarray]:
        if self._indptr is None:
            indptr = np.empty(self.n_dofs + 1, dtype=np.int32)
            indices = np.empty(self.nnz, dtype=np.int32)
            check(lib().psx_plan_pattern(self._h, ptr(indptr), ptr(indices)))
            self._indptr, self._indices = indptr, indices
        return self._indptr, self._indices

    def alloc_outputs(self, want_k: bool = True):
        torch = _torch()
        f = torch.empty(self.n_dofs, dtype=torch.float64, device=self.device)
        k = torch.empty(self.nnz, dtype=torch.float64, device=self.device) if want_k else None
        return f, k

    def assemble_mass(self, rho, kval=None):
        """Consistent mass matrix values (nnz, CSR order of ``pattern()``) on the
        device, FP64; rho: per-element density (E,) or a scalar."""
        torch = _torch()
        if np.isscalar(rho) or (isinstance(rho, np.ndarray) and rho.ndim == 0):
            rho = torch.full((self.n_elements,), float(rho), dtype=torch.float64, device=self.device)
        rho = self._device_f64(rho, self.n_elements, "rho")
        if kval is None:
            kval = torch.empty(self.nnz, dtype=torch.float64, device=self.device)
        check(lib().psx_assemble_mass(self._h, ptr(rho), ptr(kval), current_stream_handle()))
        return kval

    def bind_material(self, mu, lam):
        """Binds per-element (mu, lam) to the plan (Tet4): ``assemble`` with
        ``mu = lam = None`` then uses them, the FP32 rows reading (V mu, V K)
        from their geometry records (``psx_plan_bind_material``).  Bind again
        after the parameters change."""
        mu = self._device_f64(mu, self.n_elements, "mu")
        lam = self._device_f64(lam, self.n_elements, "lam")
        check(lib().psx_plan_bind_material(self._h, ptr(mu), ptr(lam), current_stream_handle()))

    def assemble(self, u, mu, lam, f=None, kval=None, precision="fp64", want_k=True, check_status=True):
        """f_int (3N) and K_T values (nnz, CSR order of ``pattern()``) on the device.

        u: (N,3) or (3N,) float64 device tensor; mu, lam: (E,) float64 device
        tensors, or both None for the material of ``bind_material``.  Launches
        on the current stream.  With ``check_status`` the call synchronises
        and raises the reference's ValueError on inversion.
        """
        torch = _torch()
        u = self._device_f64(u, self.n_dofs, "u")
        if mu is not None or lam is not None:
            mu = self._device_f64(mu, self.n_elements, "mu")
            lam = self._device_f64(lam, self.n_elements, "lam")
        if f is None or (want_k and kval is None):
            f0, k0 = self.alloc_outputs(want_k)
            f = f if f is not None else f0
            kval = kval if kval is not None else k0
        self._check_out(f, self.n_dofs, "f")
        if want_k:
            self._check_out(kval, self.nnz, "kval")
        else:
            kval = None
        check(
            lib().psx_assemble(
                self._h,
                ptr(u),
                ptr(mu),
                ptr(lam),
                ptr(f),
                ptr(kval),
                precision_code(precision),
                current_stream_handle(),
                ptr(self.status),
            )
        )
        if check_status:
            self.raise_on_status(u)
        return f, kval

    def assemble_hyperelastic(self, u, model, params, f=None, kval=None, precision="fp64", want_k=True,
                              check_status=True):
        """As ``assemble`` for material ``model`` (MODELS: "neohookean" (mu,
        lambda), "mooney_rivlin" (c10, c01, K), "yeoh" (c10, c20, c30, K)),
        ``params`` one (E,) array or device tensor per parameter."""
        code = MODELS[model] if isinstance(model, str) else int(model)
        npar = (2, 3, 4)[code]
        if len(params) != npar:
            raise ValueError(f"model {model} takes {npar} parameter arrays, got {len(params)}")
        u = self._device_f64(u, self.n_dofs, "u")
        prm = [self._device_f64(a, self.n_elements, f"param{i}") for i, a in enumerate(params)]
        if f is None or (want_k and kval is None):
            f0, k0 = self.alloc_outputs(want_k)
            f = f if f is not None else f0
            kval = kval if kval is not None else k0
        self._check_out(f, self.n_dofs, "f")
        if want_k:
            self._check_out(kval, self.nnz, "kval")
        else:
            kval = None
        arr = (ctypes.c_void_p * 4)(*[t.data_ptr() for t in prm], *([0] * (4 - npar)))
        check(lib().psx_assemble_hyperelastic(self._h, code, arr, ptr(u), ptr(f), ptr(kval), precision_code(precision),
                                              current_stream_handle(), ptr(self.status)))
        if check_status:
            self.raise_on_status(u)
        return f, kval

    def vjp(self, u, mu, lam, w, precision="fp64", want_params=True, check_status=True):
        """(K_T w, d(w^T f)/dmu, d(w^T f)/dlambda) as device tensors (Tet4 plans)."""
        torch = _torch()
        u = self._device_f64(u, self.n_dofs, "u")
        w = self._device_f64(w, self.n_dofs, "w")
        mu = self._device_f64(mu, self.n_elements, "mu")
        lam = self._device_f64(lam, self.n_elements, "lam")
        gu = torch.empty(self.n_dofs, dtype=torch.float64, device=self.device)
        gmu = glam = None
        if want_params:
            gmu = torch.empty(self.n_elements, dtype=torch.float64, device=self.device)
            glam = torch.empty(self.n_elements, dtype=torch.float64, device=self.device)
        check(
            lib().psx_vjp(
                self._h, ptr(u), ptr(mu), ptr(lam), ptr(w), ptr(gu), ptr(gmu), ptr(glam),
                precision_code(precision), current_stream_handle(), ptr(self.status),
            )
        )
        if check_status:
            self.raise_on_status(u)
        return gu, gmu, glam

    def raise_on_status(self, u):
        st = self.status.cpu().numpy()
        if st[0] != 0:
            uh = u.reshape(-1, 3)
            raise inversion_error(
                self.nodes, self.elements, lambda conn: uh[conn].cpu().numpy(), int(st[1])
            )

    def _device_f64(self, x, n, name):
        torch = _torch()
        if not isinstance(x, torch.Tensor):
            x = device_f64(x, self.device)
        if x.device != self.device or x.dtype != torch.float64:
            x = x.to(device=self.device, dtype=torch.float64)
        x = x.contiguous()
        if x.numel() != n:
            raise ValueError(f"{name} has {x.numel()} entries, expected {n}")
        return x

    def _check_out(self, t, n, name):
        torch = _torch()
        if not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or not t.is_contiguous():
            raise ValueError(f"{name} must be a contiguous float64 device tensor")
        if t.device != self.device or t.numel() != n:
            raise ValueError(f"{name} must have {n} entries on {self.device}")


BATCH_INFO_KEYS = ("n_instances", "n_tiles", "n_nodes", "tile", "persistent", "nnz")


class AssemblyBatch:
    """Many Tet4 (or many Tet10) plans assembled by one launch
    (``psx_assemble_batch``).

    ``plans``: ``AssemblyPlan`` objects of one element type, device and tile
    size, meshes of uneven size or one plan repeated for parameter samples
    (each instance has its own displacement buffer).  The batch keeps the plans alive.
    ``assemble`` takes one displacement (and, unless the plans' bound
    materials are used, one mu and lambda) per instance and returns the
    per-instance f and K values, each bit-identical to that plan's own
    ``assemble``.  ``persistent`` selects a persistent grid walking the
    global tile list instead of one work-group per tile (the default, faster
    as measured).  No reference counterpart: it stands for a loop
    of ``SolidDomain`` assemblies over many domains or parameter samples
    (src/psierra/domain/solid.py:230-325).
    """

    def __init__(self, plans, persistent: bool = False):
        torch = _torch()
        self.plans = list(plans)
        if not self.plans:
            raise ValueError("AssemblyBatch needs at least one plan")
        self.device = self.plans[0].device
        handles = (ctypes.c_void_p * len(self.plans))(*[p._h.value for p in self.plans])
        h = ctypes.c_void_p()
        check(lib().psx_batch_create(handles, len(self.plans), ctypes.byref(h)))
        self._h = h
        self.status = torch.zeros(2 * len(self.plans), dtype=torch.int32, device=self.device)
        self.set_persistent(persistent)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().psx_batch_destroy(h)
            except Exception:
                pass
            self._h = None

    def __len__(self):
        return len(self.plans)

    def set_persistent(self, persistent: bool) -> None:
        check(lib().psx_batch_set_persistent(self._h, int(bool(persistent))))
        self.refresh_info()

    def refresh_info(self):
        buf = (ctypes.c_int64 * len(BATCH_INFO_KEYS))()
        check(lib().psx_batch_info(self._h, buf, len(BATCH_INFO_KEYS)))
        self.info = dict(zip(BATCH_INFO_KEYS, (int(v) for v in buf)))

    def alloc_outputs(self, want_k: bool = True):
        outs = [p.alloc_outputs(want_k) for p in self.plans]
        return [o[0] for o in outs], [o[1] for o in outs]

    def _table(self, tensors):
        torch = _torch()
        return torch.tensor([t.data_ptr() for t in tensors], dtype=torch.int64, device=self.device)

    def assemble(self, us, mus=None, lams=None, fs=None, ks=None, precision="fp64", want_k=True,
                 check_status=True):
        """Per-instance lists (f, K values); ``mus = lams = None`` for the
        plans' bound materials.  Launches on the current stream; with
        ``check_status`` synchronises and raises the reference's ValueError
        for the first instance with an inverted element."""
        n = len(self.plans)
        if len(us) != n:
            raise ValueError(f"{len(us)} displacements for {n} instances")
        if (mus is None) != (lams is None):
            raise ValueError("give both mus and lams, or neither (bound materials)")
        us = [p._device_f64(u, p.n_dofs, "u") for p, u in zip(self.plans, us)]
        if mus is not None:
            if len(mus) != n or len(lams) != n:
                raise ValueError(f"mus and lams need {n} entries")
            mus = [p._device_f64(m, p.n_elements, "mu") for p, m in zip(self.plans, mus)]
            lams = [p._device_f64(m, p.n_elements, "lam") for p, m in zip(self.plans, lams)]
        if fs is None or (want_k and ks is None):
            f0, k0 = self.alloc_outputs(want_k)
            fs = fs if fs is not None else f0
            ks = ks if ks is not None else k0
        for p, f in zip(self.plans, fs):
            p._check_out(f, p.n_dofs, "f")
        if want_k:
            for p, k in zip(self.plans, ks):
                p._check_out(k, p.nnz, "kval")
        else:
            ks = None
        self.launcher(us, mus, lams, fs, ks)(precision)
        if check_status:
            self.raise_on_status(us)
        return fs, (ks if want_k else [None] * n)

    def launcher(self, us, mus, lams, fs, ks):
        """A launch of ``psx_assemble_batch`` over fixed device buffers (the
        arguments as ``assemble`` checks them, already on the device): the
        pointer tables are built now, so ``launch(precision)`` makes no copy
        and no allocation and can be captured in a HIP graph."""
        tables = [self._table(us), self._table(mus) if mus is not None else None,
                  self._table(lams) if lams is not None else None, self._table(fs),
                  self._table(ks) if ks is not None else None]
        args = [ptr(t) for t in tables]

        def launch(precision="fp64"):
            # ``tables`` stays referenced by this closure while it can run
            check(lib().psx_assemble_batch(self._h, *args, precision_code(precision), current_stream_handle(),
                                           ptr(self.status)))
            return tables

        return launch

    def vjp(self, us, mus, lams, ws, precision="fp64", want_params=True, check_status=True):
        """Per-instance lists (K_T w, d(w^T f)/dmu, d(w^T f)/dlambda) of a Tet4
        batch in one launch (``psx_vjp_batch``), each the bits of its plan's
        own ``vjp``."""
        torch = _torch()
        n = len(self.plans)
        if not (len(us) == len(mus) == len(lams) == len(ws) == n):
            raise ValueError(f"vjp needs {n} displacements, parameter pairs and cotangents")
        us = [p._device_f64(u, p.n_dofs, "u") for p, u in zip(self.plans, us)]
        ws = [p._device_f64(w, p.n_dofs, "w") for p, w in zip(self.plans, ws)]
        mus = [p._device_f64(m, p.n_elements, "mu") for p, m in zip(self.plans, mus)]
        lams = [p._device_f64(m, p.n_elements, "lam") for p, m in zip(self.plans, lams)]
        gus = [torch.empty(p.n_dofs, dtype=torch.float64, device=self.device) for p in self.plans]
        gmus = glams = None
        if want_params:
            gmus = [torch.empty(p.n_elements, dtype=torch.float64, device=self.device) for p in self.plans]
            glams = [torch.empty(p.n_elements, dtype=torch.float64, device=self.device) for p in self.plans]
        tables = [self._table(x) for x in (us, mus, lams, ws, gus)]
        tables += [self._table(gmus) if gmus else None, self._table(glams) if glams else None]
        check(lib().psx_vjp_batch(self._h, *[ptr(t) for t in tables], precision_code(precision),
                                  current_stream_handle(), ptr(self.status)))
        if check_status:
            self.raise_on_status(us)
        return gus, gmus, glams

    def raise_on_status(self, us):
        st = self.status.cpu().numpy().reshape(-1, 2)
        for i, (flag, e) in enumerate(st):
            if flag != 0:
                p = self.plans[i]
                uh = us[i].reshape(-1, 3)
                raise inversion_error(p.nodes, p.elements, lambda conn: uh[conn].cpu().numpy(), int(e))


FBAR_INVERSION_MSG = "Element inversion detected at element {}: det(F) = {:.4e} <= 0."


class FbarAssembly:
    """Tet4Fbar's domain assembly on the GPU (src/psierra/element/tet4_fbar.py:
    226-459) over a Tet4 ``AssemblyPlan`` (kept alive by this object).

    ``assemble`` returns f (3N) and the values of K_dev + K_vol in the CSR
    order of ``pattern()``: node pairs that share an element or a
    neighbouring node, explicit zeros kept.  Compressible Neo-Hookean with
    per-element (mu, lambda); ``bulk`` is the single K of the volumetric
    part (Tet4Fbar._get_bulk_modulus, tet4_fbar.py:112-118).
    """

    def __init__(self, plan: AssemblyPlan):
        if plan.npe != 4:
            raise ValueError("F-bar needs a Tet4 plan")
        self.plan = plan
        handle = ctypes.c_void_p()
        check(lib().psx_fbar_create(plan._h, ctypes.byref(handle)))
        self._h = handle
        n = ctypes.c_int64()
        check(lib().psx_fbar_nnz(handle, ctypes.byref(n)))
        self.nnz = int(n.value)
        self.n_dofs = plan.n_dofs
        self.device = plan.device
        self._indptr = None
        self._indices = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().psx_fbar_destroy(h)
            except Exception:
                pass
            self._h = None

    def pattern(self) -> tuple[np.ndarray, np.ndarray]:
        if self._indptr is None:
            indptr = np.empty(self.n_dofs + 1, dtype=np.int32)
            indices = np.empty(self.nnz, dtype=np.int32)
            check(lib().psx_fbar_pattern(self._h, ptr(indptr), ptr(indices)))
            self._indptr, self._indices = indptr, indices
        return self._indptr, self._indices

    def assemble(self, u, mu, lam, bulk, f=None, kval=None, precision="fp64", want_k=True, check_status=True):
        """f_int (3N) and, with ``want_k``, the K values (nnz) as device tensors."""
        torch = _torch()
        plan = self.plan
        u = plan._device_f64(u, self.n_dofs, "u")
        mu = plan._device_f64(mu, plan.n_elements, "mu")
        lam = plan._device_f64(lam, plan.n_elements, "lam")
        if f is None:
            f = torch.empty(self.n_dofs, dtype=torch.float64, device=self.device)
        plan._check_out(f, self.n_dofs, "f")
        if want_k:
            if kval is None:
                kval = torch.empty(self.nnz, dtype=torch.float64, device=self.device)
            plan._check_out(kval, self.nnz, "kval")
        else:
            kval = None
        check(
            lib().psx_fbar_assemble(
                self._h, ptr(u), ptr(mu), ptr(lam), float(bulk), ptr(f), ptr(kval), precision_code(precision),
                current_stream_handle(), ptr(plan.status),
            )
        )
        if check_status:
            self.raise_on_status(u)
        return f, kval

    def raise_on_status(self, u):
        """Tet4Fbar's exception for the lowest inverted element of the last call."""
        torch = _torch()
        plan = self.plan
        st = plan.status.cpu().numpy()
        if st[0] != 0:
            e = int(st[1])
            conn = plan.elements[e]
            X = plan.nodes[conn]
            ue = u.reshape(-1, 3)[torch.as_tensor(conn, device=self.device)].cpu().numpy()
            Dinv = np.linalg.inv((X[1:] - X[0]).T)
            H = np.vstack([-Dinv.sum(axis=0), Dinv])
            raise ValueError(FBAR_INVERSION_MSG.format(e, _det3(np.identity(3) + ue.T @ H)))


def element_batch(X, u, mu, lam, want_k=True, precision="fp64"):
    """Per-element k (n, nde, nde) and f (n, nde) on the device (no assembly)."""
    torch = _torch()
    dev = torch.device("cuda", torch.cuda.current_device())
    X = torch.as_tensor(X, dtype=torch.float64, device=dev).contiguous()
    u = torch.as_tensor(u, dtype=torch.float64, device=dev).contiguous()
    mu = torch.as_tensor(mu, dtype=torch.float64, device=dev).contiguous().reshape(-1)
    lam = torch.as_tensor(lam, dtype=torch.float64, device=dev).contiguous().reshape(-1)
    n, npe, _ = X.shape
    if u.shape != X.shape or mu.numel() != n or lam.numel() != n:
        raise ValueError("element_batch: inconsistent shapes")
    nd = 3 * npe
    k = torch.empty((n, nd, nd), dtype=torch.float64, device=dev) if want_k else None
    f = torch.empty((n, nd), dtype=torch.float64, device=dev)
    status = torch.zeros(2, dtype=torch.int32, device=dev)
    check(
        lib().psx_element_batch(
            npe, n, ptr(X), ptr(u), ptr(mu), ptr(lam), ptr(k), ptr(f), precision_code(precision),
            current_stream_handle(), ptr(status),
        )
    )
    st = status.cpu().numpy()
    if st[0] == 2:
        e = int(st[1])
        Xe = X[e].cpu().numpy()
        if npe == 4:
            A = np.ones((4, 4))
            A[:, 1:] = Xe
            raise ValueError(
                f"Negative or zero element volume detected (det(A) = {np.linalg.det(A)}). "
                "Check element vertex ordering."
            )
        raise ValueError("Negative or zero Jacobian determinant detected in element batch.")
    if st[0] == 1:
        e = int(st[1])
        Xe = X[e].cpu().numpy()
        ue = u[e].cpu().numpy()
        raise inversion_error(Xe, np.arange(npe)[None, :], lambda conn: ue[conn], 0)
    return k, f


def material_batch(F, mu, bulk, precision="fp64"):
    """Neo-Hookean (C_voigt (n,6,6), S (n,3,3)) for deformation gradients F (n,3,3)."""
    torch = _torch()
    dev = torch.device("cuda", torch.cuda.current_device())
    F = torch.as_tensor(F, dtype=torch.float64, device=dev).contiguous()
    n = F.shape[0]
    mu = torch.as_tensor(mu, dtype=torch.float64, device=dev).reshape(-1).expand(n).contiguous()
    bulk = torch.as_tensor(bulk, dtype=torch.float64, device=dev).reshape(-1).expand(n).contiguous()
    S = torch.empty((n, 3, 3), dtype=torch.float64, device=dev)
    C = torch.empty((n, 6, 6), dtype=torch.float64, device=dev)
    status = torch.zeros(2, dtype=torch.int32, device=dev)
    check(
        lib().psx_nh_material_batch(
            n, ptr(F), ptr(mu), ptr(bulk), ptr(S), ptr(C), precision_code(precision),
            current_stream_handle(), ptr(status),
        )
    )
    st = status.cpu().numpy()
    if st[0] != 0:
        raise ValueError(INVERSION_MSG.format(float(np.linalg.det(F[int(st[1])].cpu().numpy()))))
    return C, S
