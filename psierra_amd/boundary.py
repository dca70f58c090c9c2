"""Dirichlet bookkeeping with pSierra's interface (host side).

Restates DirichletConstraint / DOFConstraint / DirichletManager
(src/psierra/boundary/dirichlet.py:8-147) and select_nodes_by_plane
(src/psierra/boundary/selector.py:8-40): sorted prescribed DOFs, sorted free
DOFs and the prescribed values aligned with them.
"""
from __future__ import annotations

import numpy as np


class DirichletConstraint:
    def __init__(self, node_indices, components=(0, 1, 2), value=0.0, dofs_per_node=3, dof_offset=0):
        self.node_indices = np.asarray(node_indices, dtype=np.int64)
        self.components = tuple(components)
        self.dofs_per_node = dofs_per_node
        self.dof_offset = int(dof_offset)
        parts = [self.dof_offset + self.node_indices * dofs_per_node + c for c in self.components]
        self.dofs = np.sort(np.concatenate(parts)) if parts else np.array([], dtype=np.int64)
        if np.isscalar(value):
            self.values = np.full(len(self.dofs), float(value))
        else:
            self.values = np.asarray(value, dtype=np.float64)
            if len(self.values) != len(self.dofs):
                raise ValueError(
                    f"Length of values ({len(self.values)}) does not match number of constrained DOFs ({len(self.dofs)})."
                )


class DOFConstraint:
    def __init__(self, dofs, values=0.0):
        self.dofs = np.asarray(dofs, dtype=np.int64)
        if np.isscalar(values):
            self.values = np.full(len(self.dofs), float(values))
        else:
            self.values = np.asarray(values, dtype=np.float64)
            if len(self.values) != len(self.dofs):
                raise ValueError(
                    f"Length of values ({len(self.values)}) does not match number of DOFs ({len(self.dofs)})."
                )


class DirichletManager:
    def __init__(self, total_dofs: int) -> None:
        self.total_dofs = int(total_dofs)
        self._constraints = []

    def add_constraint(self, constraint) -> None:
        self._constraints.append(constraint)

    def prescribe_dofs(self, dofs, value=0.0) -> None:
        self.add_constraint(DOFConstraint(dofs, value))

    def clear(self) -> None:
        self._constraints.clear()

    @property
    def prescribed_dofs(self) -> np.ndarray:
        if not self._constraints:
            return np.array([], dtype=np.int64)
        return np.unique(np.concatenate([c.dofs for c in self._constraints]))

    @property
    def free_dofs(self) -> np.ndarray:
        mask = np.ones(self.total_dofs, dtype=bool)
        mask[self.prescribed_dofs] = False
        return np.nonzero(mask)[0]

    @property
    def prescribed_values(self) -> np.ndarray:
        p = self.prescribed_dofs
        vals = np.zeros(len(p))
        for c in self._constraints:
            vals[np.searchsorted(p, c.dofs)] = c.values
        return vals


def select_nodes_by_plane(mesh, plane: str, bound: str, tolerance: float = 1e-4) -> np.ndarray:
    axis = {"yz": 0, "xz": 1, "xy": 2}[plane]
    coords = np.asarray(mesh.nodes)[:, axis]
    target = coords.min() if bound == "min" else coords.max()
    return np.nonzero(np.abs(coords - target) <= tolerance)[0].astype(np.int64)
