"""FP64 CPU oracle for the Tet4/Tet10 Neo-Hookean assembly path.

TEST INFRASTRUCTURE ONLY.  This package restates pSierra's numpy/scipy
algorithm (file:line citations in each function) so that the HIP engine in
``psierra_amd`` can be checked against it.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker or the timed CPU baseline.  The engine
never imports it and has no CPU fallback.

What pins each function independently of its own formula (test file in
brackets; every pin has a planted mistake that must fail it):

* ``fem`` materials (Neo-Hookean, Hooke, Mooney-Rivlin, Yeoh): S(I) = 0,
  C(I) = Hooke's C, C = finite difference of S w.r.t. E, S = 2 dPsi/dC by
  finite differences of the closed-form strain energy, inversion raises
  [test_oracle.py].
* ``fem`` Tet4 routines: V = 1/6 and sum H = 0 on the unit tet; rigid
  translation and 30 degree rotation give f = 0; small strain matches the
  linear stiffness; k = central FD of f; f = FD of the element strain energy;
  Updated = Total Lagrangian; vectorised batch == per-element loop
  [test_oracle.py].
* ``fem`` Tet10 routines: the same invariants plus quadratic completeness of
  ``tet10_shape_gradients_ref`` (a quadratic field's nodal values give its
  exact gradient; a compensated mid-edge coefficient fails it)
  [test_oracle.py, test_oracle_pins.py].
* ``assembly``: closed-form structural nnz of the Kuhn box, scipy's
  COO->CSR semantics, ``np.add.at`` vs a dense sum, mass tensors against the
  exact simplex integrals and total mass, linear stiffness = tangent at rest
  [test_oracle.py, test_oracle_mass_stiffness.py].
* ``ecsw.evaluate_forces_and_stiffness``: equals Phi^T f_w and
  Phi^T K_w Phi of the global assembly of the RID with weighted Lame
  parameters (f and K are linear in them); exact symmetry; planted dropped
  weights, wrong DOF map and mismatched projection fail
  [test_oracle_pins.py].
* ``ecsw.training_matrix``: force rows sum to Phi^T f_int(u_k), one column
  equals the per-element routine, stiffness rows sum to Phi^T K0 Phi
  [test_oracle_ecsw_training.py].
* ``newton._line_search``: the acceptance rules on synthetic residual curves
  with known answers [test_oracle.py]; ``newton.solve``: quadratic residual
  decrease near the solution and equilibrium of the converged u by an
  independent assembly; a 0.8-scaled tangent converges linearly and fails
  [test_oracle_pins.py].
* ``fbar``: at u = 0 the tangent is the exact Jacobian of the force;
  homogeneous deformation gives the Total Lagrangian force; rigid motions give
  f = 0; symmetry; the vectorised K_vol equals the per-node loop
  [test_oracle_fbar.py].  Parity partially unpinned: the deviatoric tangent
  K_dev away from u = 0 under inhomogeneous deformation is not the Jacobian
  of f in the reference, so it is pinned only through its building blocks
  (B, G, C, S, each pinned above).
"""
