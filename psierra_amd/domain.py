"""Solid domain with pSierra's SolidDomain interface, assembled on the GPU.

Mirrors the parts of ``SolidDomain`` (src/psierra/domain/solid.py:20-325)
that the nonlinear analyses use: state vectors, ``update_state``,
``assemble_internal_forces`` and ``assemble_tangent_stiffness_matrix``.
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
