"""Free-DOF solver matrix and residual on the GPU, SuperLU on the host.

``FreeDofSystem`` wraps ``psx_csc`` for one plan and one set of free DOFs:
``csc(kval)`` returns K[free][:, free] as a scipy CSC matrix with exact zeros
dropped (src/psierra/analysis/nonlinear_static.py:354-355), extracted on the
device and copied to the host once; ``residual(f_int, f_ext_free)`` returns
r = f_ext_free - f_int[free] and ||r|| accumulated in FP64 on the device
(nonlinear_static.py:322-323).  ``DirectSolver`` is scipy's SuperLU, as in
src/psierra/solver/sparse_linear_solver.py:10-19.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import scipy.sparse as spsp
import scipy.sparse.linalg as spla

from ._lib import check, lib, ptr
from .engine import AssemblyPlan, _torch, current_stream_handle, host_array


class DirectSolver:
    def solve(self, A, b):
        if not isinstance(A, spsp.csc_matrix):
            A = A.tocsc()
        return spla.spsolve(A, b)


class FreeDofSystem:
    def __init__(self, plan: AssemblyPlan, free_dofs: np.ndarray, fbar=None) -> None:
        """``fbar``: an ``engine.FbarAssembly`` whose (wider) pattern the K
        values follow, instead of the plan's element pattern."""
        torch = _torch()
        self.plan = plan
        self.free_dofs = np.ascontiguousarray(free_dofs, dtype=np.int64)
        h = ctypes.c_void_p()
        if fbar is None:
            check(lib().psx_csc_create(plan._h, ptr(self.free_dofs), self.free_dofs.size, ctypes.byref(h)))
        else:
            check(lib().psx_fbar_csc_create(fbar._h, ptr(self.free_dofs), self.free_dofs.size, ctypes.byref(h)))
        self._fbar = fbar   # keeps the pattern owner alive
        self._h = h
        nf, nnz = ctypes.c_int64(), ctypes.c_int64()
        check(lib().psx_csc_info(h, ctypes.byref(nf), ctypes.byref(nnz)))
        self.n_free = nf.value
        self.nnz_struct = nnz.value
        dev = plan.device
        self.colptr = torch.empty(self.n_free + 1, dtype=torch.int32, device=dev)
        self.rowind = torch.empty(max(self.nnz_struct, 1), dtype=torch.int32, device=dev)
        self.vals = torch.empty(max(self.nnz_struct, 1), dtype=torch.float64, device=dev)
        self.r = torch.empty(max(self.n_free, 1), dtype=torch.float64, device=dev)
        self.sumsq = torch.zeros(1, dtype=torch.float64, device=dev)
        self.free_dev = torch.as_tensor(self.free_dofs, device=dev)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().psx_csc_destroy(h)
            except Exception:
                pass
            self._h = None

    def csc(self, kval, drop_zeros: bool = True) -> spsp.csc_matrix:
        check(
            lib().psx_csc_extract(
                self._h, ptr(kval), int(drop_zeros), ptr(self.colptr), ptr(self.rowind), ptr(self.vals),
                current_stream_handle(),
            )
        )
        colptr = host_array(self.colptr)
        nnz = int(colptr[-1])
        return spsp.csc_matrix(
            (host_array(self.vals[:nnz]), host_array(self.rowind[:nnz]), colptr),
            shape=(self.n_free, self.n_free),
        )

    def residual(self, f_int, f_ext_free=None):
        check(
            lib().psx_free_residual(
                self._h, ptr(f_int), ptr(f_ext_free), ptr(self.r), ptr(self.sumsq), current_stream_handle()
            )
        )
        return self.r[: self.n_free], math.sqrt(float(self.sumsq.item()))
