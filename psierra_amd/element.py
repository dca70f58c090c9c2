"""GPU Total Lagrangian Tet4/Tet10 elements with pSierra's Element interface.

Drop-ins for ``Tet4TotalLagrangian`` (src/psierra/element/tet4_total_lagrangian.py)
and ``Tet10TotalLagrangian`` (src/psierra/element/tet10_total_lagrangian.py).

pSierra's ``SolidDomain`` hands the whole assembly to an element that has
``assemble_internal_forces(mesh, displacement)`` and
``assemble_tangent_stiffness_matrix(mesh, displacement)``
(src/psierra/domain/solid.py:238-241, 288-291), the hook ``Tet4Fbar`` uses
(src/psierra/element/tet4_fbar.py:217-459).  These classes implement both
hooks on the GPU with a plan cached per mesh (keyed on the mesh object, as
Tet4Fbar keys its geometry cache, tet4_fbar.py:54-57), so an unmodified
``SolidDomain`` and every analysis above it run the HIP kernels.

The per-element ``compute_nonlinear_force_and_tangent_stiffness`` runs the
device element-batch kernel on one element (used by code such as the ECSW
model that calls elements one at a time).

Two things the hooks do beyond a plain call:
* a Neo-Hookean material is bound to a Tet4 plan whenever it changes
  (``psx_plan_bind_material``), so every assembly reads the per-element
  parameters once per bind instead of once per call;
* the hooks keep the last displacement's f (and K when it was formed) on
  the device: the Newton loop asks for f and then K at the same u
  (src/psierra/analysis/nonlinear_static.py:296-297), and its line search
  evaluates f at the step it then accepts (:366-395), so each iteration's
  f request is served from the cache and its K request is one fused f + K
  launch.  The cache is keyed on the displacement's values, the material
  version and the precision.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as spsp

from . import engine
from .material import MooneyRivlinField, YeohField, lame_arrays, material_model


class _GpuTotalLagrangian:
    supports_element_indexing: bool = True
    _npe = 4

    def __init__(self, material, precision: str = "fp64", tile=0) -> None:
        """tile: work-group tile (0 = the plan's default), or "auto" to time
        the supported tiles on the first assembly and keep the fastest."""
        engine.precision_code(precision)
        self._material = material
        self.precision = precision
        self.tile = tile
        self._cache = {}

    # -- Element interface (src/psierra/element/base.py:28-50) --------------
    @property
    def material(self):
        return self._material

    @material.setter
    def material(self, mat) -> None:
        self._material = mat
        for st in self._cache.values():
            st["params_key"] = None

    @property
    def nodes_per_element(self) -> int:
        return self._npe

    @property
    def dofs_per_node(self) -> int:
        return 3

    @property
    def dofs_per_element(self) -> int:
        return 3 * self._npe

    # -- plan / parameter cache ---------------------------------------------
    def _state(self, mesh):
        key = id(mesh)
        st = self._cache.get(key)
        if st is None or st["mesh"] is not mesh:
            elements = np.asarray(mesh.elements, dtype=np.int64)
            if elements.shape[1] != self._npe:
                raise ValueError(f"{type(self).__name__} needs {self._npe}-node elements")
            tile = 0 if self.tile == "auto" else int(self.tile)
            st = {"mesh": mesh, "plan": engine.AssemblyPlan(mesh.nodes, elements, tile), "params_key": None,
                  "tuned": self.tile != "auto", "bound": False, "hook": None}
            self._cache = {key: st}
        pkey = (id(self._material), getattr(self._material, "version", 0))
        if st["params_key"] != pkey:
            import torch

            model, arrays = material_model(self._material, st["plan"].n_elements)
            dev = st["plan"].device
            st["model"] = model
            st["params"] = [torch.as_tensor(a, dtype=torch.float64, device=dev) for a in arrays]
            st["bound"] = False
            if model == "neohookean":
                st["mu"], st["lam"] = st["params"]
                if st["plan"].npe == 4 and self._binds_material:
                    st["plan"].bind_material(st["mu"], st["lam"])
                    st["bound"] = True
            st["params_key"] = pkey
            st["hook"] = None
        return st

    def plan(self, mesh) -> engine.AssemblyPlan:
        return self._state(mesh)["plan"]

    def free_dof_system(self, mesh, free_dofs):
        """Device free-dof CSC / residual helper over this element's K pattern."""
        from .solver import FreeDofSystem

        return FreeDofSystem(self.plan(mesh), free_dofs)

    # Element name carried by the inversion message (None: the Total
    # Lagrangian text "Element inversion detected: det(F) = ...").
    _inversion_name: str | None = None
    # Bind a Neo-Hookean material to Tet4 plans (the F-bar element runs its
    # own kernels with per-call parameters).
    _binds_material: bool = True

    def _relabel(self, err: ValueError) -> ValueError:
        msg = str(err)
        if self._inversion_name and msg.startswith("Element inversion detected: "):
            return ValueError(msg.replace("Element inversion detected: ",
                                          f"Element inversion detected in {self._inversion_name}: ", 1))
        return err

    def _autotune(self, mesh, u, precision, reps: int = 3) -> None:
        """Time each supported tile on this mesh and input and keep the
        fastest (every tile gives the same bits)."""
        import torch

        st = self._state(mesh)
        st["tuned"] = True
        plan = st["plan"]
        # a tile holds every node's whole incidence run
        tiles = [t for t in ((64, 128, 256) if plan.npe == 4 else (64, 128)) if t >= plan.info["max_incidences"]]
        timings = {}
        for tile in tiles:
            plan.set_tile(tile)
            self.assemble_device(mesh, u, want_k=True, precision=precision)
            torch.cuda.synchronize()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(reps):
                self.assemble_device(mesh, u, want_k=True, check_status=False, precision=precision)
            b.record()
            torch.cuda.synchronize()
            timings[tile] = a.elapsed_time(b) / reps
        st["tile_timings"] = timings
        plan.set_tile(min(timings, key=timings.get))

    def assemble_device(self, mesh, u, want_k: bool = True, f=None, kval=None, check_status: bool = True,
                        precision: str | None = None):
        """f_int and K values as device tensors (u may be a device tensor)."""
        st = self._state(mesh)
        if not st["tuned"]:
            self._autotune(mesh, u, precision or self.precision)
        prec = precision or self.precision
        try:
            if st["model"] == "neohookean":
                mu, lam = (None, None) if st["bound"] else (st["mu"], st["lam"])
                return st["plan"].assemble(
                    u, mu, lam, f=f, kval=kval, precision=prec, want_k=want_k, check_status=check_status,
                )
            return st["plan"].assemble_hyperelastic(
                u, st["model"], st["params"], f=f, kval=kval, precision=prec, want_k=want_k,
                check_status=check_status,
            )
        except ValueError as err:
            raise self._relabel(err) from None

    def _nnz(self, mesh) -> int:
        return self._state(mesh)["plan"].nnz

    def _raise_status(self, mesh, u) -> None:
        try:
            self._state(mesh)["plan"].raise_on_status(u)
        except ValueError as err:
            raise self._relabel(err) from None

    def graph_assembly(self, mesh, want_k: bool = True, precision: str | None = None):
        """This element's assembly on ``mesh`` captured in a HIP graph
        (``graph.GraphedCall``): ``replay(u, check_status)`` returns the device
        (f, kval) buffers, overwritten by the next replay."""
        import torch

        from .graph import GraphedCall

        plan = self._state(mesh)["plan"]
        dev = plan.device
        u = torch.zeros(plan.n_dofs, dtype=torch.float64, device=dev)
        f = torch.empty(plan.n_dofs, dtype=torch.float64, device=dev)
        k = torch.empty(self._nnz(mesh), dtype=torch.float64, device=dev) if want_k else None
        prec = precision or self.precision
        self.assemble_device(mesh, u, want_k=want_k, f=f, kval=k, check_status=False, precision=prec)

        def launch():
            self.assemble_device(mesh, u, want_k=want_k, f=f, kval=k, check_status=False, precision=prec)

        return GraphedCall(launch, u, (f, k), check=lambda uu: self._raise_status(mesh, uu))

    # -- SolidDomain hooks (src/psierra/domain/solid.py:238-241, 288-291) ---
    def _hook_eval(self, mesh, displacement, want_k: bool):
        """(f host array, K values device tensor or None) at ``displacement``,
        from the cache when the last hook call had the same displacement,
        material version and precision (and K when K is asked for)."""
        st = self._state(mesh)
        u = np.ascontiguousarray(displacement, dtype=np.float64).reshape(-1)
        key = (self.precision, self._hook_extra())
        c = st["hook"]
        if c is not None and c["key"] == key and (c["k"] is not None or not want_k) and np.array_equal(c["u"], u):
            return c["f"], c["k"]
        f, k = self.assemble_device(mesh, u, want_k=want_k)
        fh = engine.host_array(f)
        st["hook"] = {"u": u.copy(), "f": fh, "k": k, "key": key}
        return fh, k

    def _hook_extra(self):
        """Settings besides precision and material that change the result."""
        return None

    def assemble_internal_forces(self, mesh, displacement) -> np.ndarray:
        f, _ = self._hook_eval(mesh, displacement, want_k=False)
        return f.copy()

    def assemble_tangent_stiffness_matrix(self, mesh, displacement) -> spsp.csr_matrix:
        _, k = self._hook_eval(mesh, displacement, want_k=True)
        return self._csr(mesh, k)

    def _csr(self, mesh, kval) -> spsp.csr_matrix:
        plan = self._state(mesh)["plan"]
        indptr, indices = plan.pattern()
        n = plan.n_dofs
        return spsp.csr_matrix((engine.host_array(kval), indices, indptr), shape=(n, n))

    # SolidDomain.assemble_stiffness_matrix delegates here when the element
    # defines it (src/psierra/domain/solid.py:180-181): the linear stiffness
    # B^T C B V with C Hooke's law at the reference state is the material's
    # tangent at u = 0 (S = 0, C4 = Hooke for each supported model),
    # assembled in FP64.
    def assemble_stiffness_matrix(self, mesh) -> spsp.csr_matrix:
        plan = self._state(mesh)["plan"]
        import torch

        u0 = torch.zeros(plan.n_dofs, dtype=torch.float64, device=plan.device)
        _, k = self.assemble_device(mesh, u0, want_k=True, precision="fp64")
        return self._csr(mesh, k)

    def assemble_mass_matrix(self, mesh) -> spsp.csr_matrix:
        """Consistent mass on the GPU (SolidDomain.assemble_mass_matrix,
        src/psierra/domain/solid.py:208-228) with rho = material.mass_density."""
        plan = self._state(mesh)["plan"]
        return self._csr(mesh, plan.assemble_mass(float(self._material.mass_density)))

    # -- per-element API ------------------------------------------------------
    def compute_mass_matrix(self, volume: float) -> np.ndarray:
        """Element consistent mass (tet4_base.py:138-151, tet10_base.py:344-372)."""
        c = reference_mass_tensor(self._npe)
        return float(self._material.mass_density) * float(volume) * np.kron(c, np.eye(3))

    def compute_linear_stiffness_matrix(self, node_coords, elem_index=None) -> np.ndarray:
        """Element B^T C B V at the reference state (tet4_base.py:129-136)."""
        k, _ = self.compute_nonlinear_force_and_tangent_stiffness(node_coords, np.zeros((self._npe, 3)),
                                                                   elem_index=elem_index)
        return k

    def _element_params(self, elem_index):
        mat = self._material
        p = getattr(mat, "parameters", None)
        if isinstance(mat, (MooneyRivlinField, YeohField)) or hasattr(p, "c01") or hasattr(p, "c20"):
            raise NotImplementedError("the per-element GPU API evaluates the Neo-Hookean model; "
                                      "other models run through the assembly hooks")
        if getattr(mat, "supports_element_indexing", False) and hasattr(mat, "mu"):
            e = 0 if elem_index is None else int(elem_index)
            return float(mat.mu[e]), float(mat.lam[e])
        mu, lam = lame_arrays(mat, 1)
        return float(mu[0]), float(lam[0])

    def compute_nonlinear_force_and_tangent_stiffness(self, node_coords, u_nodes, elem_index=None, **kw):
        quadrature_order = kw.pop("quadrature_order", None)
        if kw:
            raise TypeError(f"unexpected arguments {sorted(kw)}")
        if quadrature_order not in (None, 4):
            raise ValueError("the GPU Tet10 element implements the 4-point rule only")
        mu, lam = self._element_params(elem_index)
        X = np.asarray(node_coords, dtype=np.float64).reshape(1, self._npe, 3)
        u = np.asarray(u_nodes, dtype=np.float64).reshape(1, self._npe, 3)
        try:
            k, f = engine.element_batch(X, u, [mu], [lam], want_k=True, precision=self.precision)
        except ValueError as err:
            raise self._relabel(err) from None
        return k[0].cpu().numpy(), f[0].cpu().numpy()


def reference_mass_tensor(npe: int) -> np.ndarray:
    """c[p][q] = (1/V) * integral of N_p N_q over the element, exact: the
    shape functions as barycentric polynomials (Tet4 L_p; Tet10 corners
    L_p (2 L_p - 1), mid-edge 4 L_i L_j on edges 01, 12, 20, 03, 13, 23) and
    (1/V) int L^alpha = 6 alpha! / (|alpha| + 3)!."""
    from math import factorial

    def lin(i, pw):
        m = [0, 0, 0, 0]
        m[i] = pw
        return tuple(m)

    if npe == 4:
        N = [[(lin(i, 1), 1.0)] for i in range(4)]
    else:
        N = [[(lin(i, 2), 2.0), (lin(i, 1), -1.0)] for i in range(4)]
        for i, j in ((0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)):
            m = [0, 0, 0, 0]
            m[i] = m[j] = 1
            N.append([(tuple(m), 4.0)])
    c = np.zeros((npe, npe))
    for a in range(npe):
        for b in range(npe):
            tot = 0.0
            for ma, ca in N[a]:
                for mb, cb in N[b]:
                    ex = [x + y for x, y in zip(ma, mb)]
                    num = 6.0
                    for x in ex:
                        num *= factorial(x)
                    tot += ca * cb * num / factorial(sum(ex) + 3)
            c[a, b] = tot
    return c


class GpuTet4TotalLagrangian(_GpuTotalLagrangian):
    _npe = 4


class GpuTet10TotalLagrangian(_GpuTotalLagrangian):
    _npe = 10

    @property
    def default_quadrature_order(self) -> int:
        return 4


class GpuTet4UpdatedLagrangian(GpuTet4TotalLagrangian):
    """Tet4UpdatedLagrangian (src/psierra/element/tet4_updated_lagrangian.py:
    31-109) on the GPU.  Its spatial form (h = H F^-1, Cauchy stress, pushed-
    forward tangent over the current volume) gives the same f_int and K_T as
    the Total Lagrangian form (the reference's own statement, :18-20, checked
    on the oracle's restatements of both), so the same kernels run; only the
    inversion message names the element."""

    _inversion_name = "Tet4UpdatedLagrangian"


class GpuTet10UpdatedLagrangian(GpuTet10TotalLagrangian):
    """Tet10UpdatedLagrangian (src/psierra/element/tet10_updated_lagrangian.py)
    on the GPU: pointwise the same integrand as the Total Lagrangian form at
    each of the 4 quadrature points, so the same kernels run."""

    _inversion_name = "Tet10UpdatedLagrangian"


class GpuTet4Fbar(GpuTet4TotalLagrangian):
    """Tet4Fbar (src/psierra/element/tet4_fbar.py:17-459) on the GPU for the
    compressible Neo-Hookean material with per-element (mu, lambda).

    The domain hooks run ``engine.FbarAssembly``: nodal volume averaging
    J_bar, F_bar = (J_bar / J)^(1/3) F, f = B_bar^T S(F_bar) V and the
    tangent K_dev + K_vol of tet4_fbar.py:332-459, which couples nodes up to
    two element edges apart.  ``bulk_modulus`` is the single K of the
    volumetric part; by default it is taken as Tet4Fbar._get_bulk_modulus
    does (tet4_fbar.py:112-118): the material's ``parameters.bulk_modulus``,
    else from its small-strain constitutive matrix, else the common
    lambda + 2 mu / 3 of a uniform per-element material.

    The per-element API with J_bar = J (no smoothed value given) is the
    Total Lagrangian element (ratio 1, B_bar = B), as in the reference; an
    explicit smoothed J_bar is served through the domain hooks only.
    """

    supports_element_indexing = True
    _binds_material = False

    def __init__(self, material, precision: str = "fp64", tile=0, bulk_modulus: float | None = None) -> None:
        super().__init__(material, precision=precision, tile=tile)
        self.bulk_modulus = bulk_modulus

    def _bulk(self, st) -> float:
        if self.bulk_modulus is not None:
            return float(self.bulk_modulus)
        mat = self._material
        for attr in ("_parameters", "parameters"):
            prm = getattr(mat, attr, None)
            if prm is not None and hasattr(prm, "bulk_modulus"):
                return float(prm.bulk_modulus)
        if hasattr(mat, "compute_constitutive_matrix"):
            C = np.asarray(mat.compute_constitutive_matrix(), dtype=np.float64)
            return float(np.sum(C[:3, :3])) / 9.0
        mu = st["mu"].cpu().numpy()
        lam = st["lam"].cpu().numpy()
        K = lam + (2.0 / 3.0) * mu
        if K.size and np.all(K == K[0]):
            return float(K[0])
        raise ValueError("GpuTet4Fbar needs bulk_modulus= for a material with per-element bulk moduli")

    def _fbar(self, mesh):
        st = self._state(mesh)
        if st["model"] != "neohookean":
            raise NotImplementedError("the GPU F-bar element evaluates the Neo-Hookean model only")
        if "fbar" not in st:
            st["fbar"] = engine.FbarAssembly(st["plan"])
        key = (st["params_key"], self.bulk_modulus)
        if st.get("bulk_key") != key:
            st["bulk"] = self._bulk(st)
            st["bulk_key"] = key
        return st

    def _nnz(self, mesh) -> int:
        return self._fbar(mesh)["fbar"].nnz

    def _hook_extra(self):
        return self.bulk_modulus

    def _raise_status(self, mesh, u) -> None:
        self._fbar(mesh)["fbar"].raise_on_status(u)

    def assemble_device(self, mesh, u, want_k: bool = True, f=None, kval=None, check_status: bool = True,
                        precision: str | None = None):
        st = self._fbar(mesh)
        if not st["tuned"]:
            self._autotune(mesh, u, precision or self.precision)
        return st["fbar"].assemble(u, st["mu"], st["lam"], st["bulk"], f=f, kval=kval,
                                   precision=precision or self.precision, want_k=want_k, check_status=check_status)

    def free_dof_system(self, mesh, free_dofs):
        from .solver import FreeDofSystem

        st = self._fbar(mesh)
        return FreeDofSystem(st["plan"], free_dofs, fbar=st["fbar"])

    def prepare_for_assembly(self, mesh, displacement) -> None:
        """The smoothed J_bar is recomputed on the device by every assembly."""
        self._fbar(mesh)

    def prepare_for_linear_assembly(self, mesh) -> None:
        self._fbar(mesh)

    def _csr(self, mesh, kval) -> spsp.csr_matrix:
        fb = self._fbar(mesh)["fbar"]
        indptr, indices = fb.pattern()
        n = fb.n_dofs
        K = spsp.csr_matrix((engine.host_array(kval), indices.copy(), indptr.copy()), shape=(n, n))
        # the reference returns the CSR sum K_dev + K_vol, which drops exact zeros
        K.eliminate_zeros()
        return K

    def assemble_mass_matrix(self, mesh) -> spsp.csr_matrix:
        return GpuTet4TotalLagrangian._csr(self, mesh, self._state(mesh)["plan"].assemble_mass(
            float(self._material.mass_density)))

    def compute_nonlinear_force_and_tangent_stiffness(self, node_coords, u_nodes, elem_index=None, **kw):
        J_bar = kw.pop("J_bar", None)
        if J_bar is not None:
            raise NotImplementedError("an explicit smoothed J_bar is evaluated through the F-bar assembly hooks")
        return super().compute_nonlinear_force_and_tangent_stiffness(node_coords, u_nodes, elem_index, **kw)
