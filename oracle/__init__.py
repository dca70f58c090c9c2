"""FP64 CPU oracle for the Tet4/Tet10 Neo-Hookean assembly path.

TEST INFRASTRUCTURE ONLY.  This package restates pSierra's numpy/scipy
algorithm (file:line citations in each function) so that the HIP engine in
``psierra_amd`` can be checked against it.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker or the timed CPU baseline.  The engine
never imports it and has no CPU fallback.

Pinning (what fixes each function independently of its own formula):

* ``material``: S(I) = 0 and C(I) = Hookean C; C = finite difference of
  S w.r.t. E via a polar-decomposition perturbation; S = 2 dPsi/dC by finite
  differences of the closed-form strain energy; inversion raises.
* ``tet4`` / ``tet10`` element routines: V = 1/6 and sum H = 0 on the unit
  tet; rigid translation and 30 degree rotation give f = 0; small-strain
  f matches the linear stiffness times u; k_tan = central FD of f; f = FD of
  the element strain energy; Tet10 constant-strain patch test; vectorised
  batch == per-element loop.
* ``assembly``: structural nnz closed forms of the Kuhn box; scipy's COO->CSR
  is the reference's own routine; ``np.add.at`` scatter vs dense sum.
* ``newton``/``linesearch``: Armijo/halving rules restated with constants;
  one Newton step reduces the residual quadratically on a small problem.
* ``ecsw``: exact symmetry and a dense reconstruction check.

No function of this package is "parity unpinned".
"""
