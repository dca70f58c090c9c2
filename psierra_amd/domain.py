"""Solid domain with pSierra's SolidDomain interface, assembled on the GPU.

Mirrors the parts of ``SolidDomain`` (src/psierra/domain/solid.py:20-339)
that the analyses use: state vectors, ``update_state``, the history-state
pass-throughs to the material, ``assemble_internal_forces``,
``assemble_tangent_stiffness_matrix``, the linear stiffness and consistent
mass and the two energies -- without the reference's per-element host
buffers (E x nde^2 for K and M, the COO indices: ~28 GB at 6.0M Tet4).
Assembly is delegated to the element's hooks, exactly as the reference does
when the element provides them (solid.py:238-241, 288-291).
"""
from __future__ import annotations

import numpy as np


class SolidDomain:
    def __init__(self, mesh, material, element) -> None:
        self.mesh = mesh
        self.material = material
        self.element = element
        self.element.material = self.material
        if not (hasattr(element, "assemble_internal_forces") and hasattr(element, "assemble_tangent_stiffness_matrix")):
            raise TypeError("psierra_amd.SolidDomain needs a GPU element with assembly hooks")
        n = mesh.n_dofs
        self.u = np.zeros(n)
        self.v = np.zeros(n)
        self.a = np.zeros(n)

    @property
    def n_dofs(self) -> int:
        return self.mesh.n_dofs

    @property
    def displacement(self) -> np.ndarray:
        return self.u.reshape(-1, 3)

    @property
    def velocity(self) -> np.ndarray:
        return self.v.reshape(-1, 3)

    @property
    def acceleration(self) -> np.ndarray:
        return self.a.reshape(-1, 3)

    @property
    def current_nodes(self) -> np.ndarray:
        return self.mesh.nodes + self.displacement

    def update_state(self, u=None, v=None, a=None) -> None:
        if u is not None:
            self.u[:] = u
        if v is not None:
            self.v[:] = v
        if a is not None:
            self.a[:] = a

    def reset_state(self) -> None:
        self.u.fill(0.0)
        self.v.fill(0.0)
        self.a.fill(0.0)
        if hasattr(self.material, "reset_state"):
            self.material.reset_state()

    # History-dependent materials: the analyses call these around steps and
    # cutbacks (solid.py:142-166); they pass through to the material.
    def set_time_step(self, dt: float) -> None:
        if hasattr(self.material, "set_time_step"):
            self.material.set_time_step(dt)

    def commit_state(self) -> None:
        if hasattr(self.material, "commit_state"):
            self.material.commit_state()

    def revert_state(self) -> None:
        if hasattr(self.material, "revert_state"):
            self.material.revert_state()

    def get_history_variables(self) -> dict:
        if hasattr(self.material, "get_history_variables"):
            return self.material.get_history_variables()
        return {}

    def set_history_variables(self, vars: dict) -> None:
        if hasattr(self.material, "set_history_variables"):
            self.material.set_history_variables(vars)

    def compute_kinetic_energy(self) -> float:
        """0.5 v^T M v with the consistent mass (solid.py:331-334)."""
        M = self.assemble_mass_matrix()
        return 0.5 * float(self.v @ (M @ self.v))

    def compute_strain_energy(self) -> float:
        """0.5 u^T K u with the linear stiffness (solid.py:336-339)."""
        K = self.assemble_stiffness_matrix()
        return 0.5 * float(self.u @ (K @ self.u))

    def assemble_internal_forces(self) -> np.ndarray:
        return self.element.assemble_internal_forces(self.mesh, self.displacement)

    def assemble_internal_force_vector(self) -> np.ndarray:
        return self.assemble_internal_forces()

    def assemble_tangent_stiffness_matrix(self):
        return self.element.assemble_tangent_stiffness_matrix(self.mesh, self.displacement)

    def assemble_stiffness_matrix(self):
        """Linear stiffness from the reference geometry (solid.py:172-206)."""
        return self.element.assemble_stiffness_matrix(self.mesh)

    def assemble_mass_matrix(self):
        """Consistent mass matrix (solid.py:208-228)."""
        return self.element.assemble_mass_matrix(self.mesh)
