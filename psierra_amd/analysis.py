"""Incremental Newton-Raphson on the GPU assembly path.

Restates the displacement/force-controlled part of
NonlinearStaticAnalysis.solve (src/psierra/analysis/nonlinear_static.py:73-549)
without contact or output writers: per load step the prescribed values are
scaled by the load factor, then Newton iterations

  f_int(u)       -> psx_assemble, force only (GPU)
  r, ||r||       -> psx_free_residual (GPU, FP64)
  K_T(u)         -> psx_assemble with values (GPU), only when not converged
  K_ff (CSC)     -> psx_csc_extract (GPU) -> host
  du             -> SuperLU (host, scipy)
  line search    -> ContactAwareLineSearch rules, each trial one force-only
                    GPU assembly and residual.

Convergence and reference-force rules are those of nonlinear_static.py:
250-474 (f_ref = ||f_ext_step,free|| if >= 1e-12, else
max(||r||, ||f_int,free||, 1) at the step's first iteration; converge when
||r|| < nr_tol * f_ref).  On element inversion during assembly the state is
restored and the ValueError propagates; on a failed line search or solve the
step is reported unconverged.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .boundary import DirichletManager
from .engine import device_f64, host_array
from .line_search import ContactAwareLineSearch
from .solver import DirectSolver, FreeDofSystem


@dataclass
class NonlinearStaticResult:
    displacement: np.ndarray
    reactions: np.ndarray
    converged: bool
    steps_completed: int
    iteration_history: list = field(default_factory=list)
    residual_history: list = field(default_factory=list)


class NonlinearStaticAnalysis:
    def __init__(self, domain, linear_solver=None) -> None:
        self.domain = domain
        self.linear_solver = linear_solver if linear_solver is not None else DirectSolver()

    def solve(self, dirichlet, f_ext=None, n_steps=10, nr_iterations=25, nr_tol=1e-4, verbose=False,
              line_search=None, use_graphs: bool = False) -> NonlinearStaticResult:
        """``use_graphs``: run every f_int and K_T assembly of the loop as a
        replay of a HIP graph captured once per solve (same kernels, same
        results)."""
        import torch

        dom = self.domain
        element = dom.element
        mesh = dom.mesh
        n = dom.n_dofs
        plan = element.plan(mesh)
        dev = plan.device
        if isinstance(dirichlet, list):
            mgr = DirichletManager(n)
            for c in dirichlet:
                mgr.add_constraint(c)
            dirichlet = mgr
        free = dirichlet.free_dofs
        p_dofs = dirichlet.prescribed_dofs
        u_p = dirichlet.prescribed_values
        if len(free) == 0:
            raise ValueError("All degrees of freedom are constrained. Nothing to solve.")
        f_ext = np.zeros(n) if f_ext is None else np.asarray(f_ext, dtype=np.float64)
        ls = line_search if line_search is not None else ContactAwareLineSearch()
        system = element.free_dof_system(mesh, free) if hasattr(element, "free_dof_system") else FreeDofSystem(plan, free)
        free_t = system.free_dev
        u = torch.as_tensor(dom.u.copy(), device=dev)
        f_ext_free = torch.as_tensor(f_ext[free], device=dev)
        result = NonlinearStaticResult(dom.u.copy(), np.zeros(n), False, 0)

        if use_graphs:
            g_force = element.graph_assembly(mesh, want_k=False)
            g_tangent = element.graph_assembly(mesh, want_k=True)

            def force(u_vec):
                return g_force.replay(u_vec, check_status=True)[0]

            def tangent(u_vec):
                return g_tangent.replay(u_vec, check_status=True)[1]
        else:

            def force(u_vec):
                f, _ = element.assemble_device(mesh, u_vec, want_k=False)
                return f

            def tangent(u_vec):
                return element.assemble_device(mesh, u_vec, want_k=True)[1]

        for step in range(1, n_steps + 1):
            lf = step / n_steps
            accepted = u.clone()
            if len(p_dofs):
                u[torch.as_tensor(p_dofs, device=dev)] = torch.as_tensor(lf * u_p, device=dev)
            f_step_free = lf * f_ext_free
            f_ext_norm = float(torch.linalg.vector_norm(f_step_free))
            f_ref = f_ext_norm if f_ext_norm >= 1e-12 else None
            converged = False
            iters = 0
            for it in range(1, nr_iterations + 1):
                iters = it
                try:
                    f_int = force(u)
                except ValueError:
                    dom.update_state(u=host_array(accepted))
                    raise
                r, rnorm = system.residual(f_int, f_step_free)
                if not np.isfinite(rnorm):
                    break
                if f_ref is None:
                    f_ref = max(rnorm, float(torch.linalg.vector_norm(f_int[free_t])), 1.0)
                tol = nr_tol * f_ref
                result.residual_history.append(rnorm)
                if verbose:
                    print(f"  Iter {it:2d} | ||R|| = {rnorm:.4e} | Rel = {rnorm / f_ref:.4e}")
                if rnorm < tol:
                    converged = True
                    break
                K_ff = system.csc(tangent(u))
                try:
                    du = self.linear_solver.solve(K_ff, host_array(r))
                except (RuntimeError, np.linalg.LinAlgError):
                    break
                if not np.all(np.isfinite(du)):
                    break
                du_t = device_f64(du, dev)

                def trial(alpha):
                    uc = u.clone()
                    uc[free_t] += alpha * du_t
                    _, nrm = system.residual(force(uc), f_step_free)
                    return nrm

                res = ls.search(trial, rnorm, tol, contact_active=False, reference_norm=f_ref)
                if not res.success:
                    converged = False
                    break
                u[free_t] += res.alpha * du_t
            if not converged:
                u = accepted
                dom.update_state(u=host_array(u))
                result.displacement = host_array(u)
                result.reactions = host_array(force(u)) - ((step - 1) / n_steps) * f_ext
                result.steps_completed = step - 1
                return result
            result.iteration_history.append(iters)
            result.steps_completed = step
        dom.update_state(u=host_array(u))
        result.displacement = host_array(u)
        result.reactions = host_array(force(u)) - f_ext
        result.converged = True
        return result
