"""FP64 Newton restatement (test infrastructure only).

Follows NonlinearStaticAnalysis.solve without contact
(src/psierra/analysis/nonlinear_static.py:250-474): per load step prescribed
values scaled by the load factor, residual r = f_ext_step,free - f_int,free,
f_ref rule, convergence ||r|| < nr_tol f_ref, K_ff = (K + empty)[free][:,free]
.tocsc(), du = spsolve(K_ff, r), and the backtracking rules of
ContactAwareLineSearch.search (src/psierra/analysis/line_search.py:120-260)
without contact.  Built only on oracle.assembly.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse.linalg as spla

from . import assembly as A


def newton_step(nodes, elements, u, mu, lam, free, f_ext_free=None):
    """One Newton increment du on the free DOFs (config: cantilever, x=0 clamped)."""
    f, K = A.assemble(nodes, elements, u.reshape(-1, 3), mu, lam)
    r = (0.0 if f_ext_free is None else f_ext_free) - f[free]
    Kff = A.free_dof_csc(K, free)
    return spla.spsolve(Kff, r), r, Kff


def _line_search(eval_fn, r0, tol, c1=2.0e-4, max_it=16):
    alpha, best, best_a, inverted = 1.0, np.inf, 1.0, False
    for _ in range(max_it):
        try:
            c = float(eval_fn(alpha))
        except ValueError as exc:
            if "inversion" not in str(exc).lower() and "det(f)" not in str(exc).lower():
                raise
            c = np.inf
        if not np.isfinite(c):
            inverted = True
            alpha *= 0.5
            continue
        if c < best:
            best, best_a = c, alpha
        if c <= tol or c <= r0 or c * c <= (1.0 - 2.0 * c1 * alpha) * r0 * r0:
            return True, alpha
        alpha *= 0.5
    if np.isfinite(best) and best < 5.0 * r0 and not inverted:
        return True, best_a
    return False, best_a


def solve(nodes, elements, mu, lam, free, p_dofs, p_vals, f_ext, n_steps, nr_iterations=25, nr_tol=1e-4,
          history=None):
    """history: optional list; (load step, iteration, ||r||) is appended for
    every residual evaluated at the top of a Newton iteration."""
    n = 3 * nodes.shape[0]
    u = np.zeros(n)
    iters = []
    for step in range(1, n_steps + 1):
        lf = step / n_steps
        u[p_dofs] = lf * p_vals
        fs = lf * f_ext[free]
        nf = np.linalg.norm(fs)
        f_ref = nf if nf >= 1e-12 else None
        ok = False
        for it in range(1, nr_iterations + 1):
            f, _ = A.assemble(nodes, elements, u.reshape(-1, 3), mu, lam, want_k=False)
            r = fs - f[free]
            rn = np.linalg.norm(r)
            if history is not None:
                history.append((step, it, float(rn)))
            if f_ref is None:
                f_ref = max(rn, np.linalg.norm(f[free]), 1.0)
            tol = nr_tol * f_ref
            if rn < tol:
                ok = True
                break
            _, K = A.assemble(nodes, elements, u.reshape(-1, 3), mu, lam)
            du = spla.spsolve(A.free_dof_csc(K, free), r)

            def trial(a):
                uc = u.copy()
                uc[free] += a * du
                fc, _ = A.assemble(nodes, elements, uc.reshape(-1, 3), mu, lam, want_k=False)
                return np.linalg.norm(fs - fc[free])

            good, a = _line_search(trial, rn, tol)
            if not good:
                break
            u[free] += a * du
        if not ok:
            return u, False, iters
        iters.append(it)
    return u, True, iters
