"""Host logic of psierra_amd.domain.SolidDomain (no GPU): the Domain API the
reference analyses call (src/psierra/domain/solid.py:75-166, 331-339) with a
stand-in element that defines the assembly hooks and a stand-in material
that records the history-state calls."""
from __future__ import annotations

import numpy as np
import scipy.sparse as spsp

from psierra_amd.domain import SolidDomain
from psierra_amd.mesh import TetMesh


class _Material:
    def __init__(self):
        self.calls = []

    def set_time_step(self, dt):
        self.calls.append(("dt", dt))

    def commit_state(self):
        self.calls.append("commit")

    def revert_state(self):
        self.calls.append("revert")

    def reset_state(self):
        self.calls.append("reset")

    def get_history_variables(self):
        return {"h": np.arange(3.0)}

    def set_history_variables(self, vars):
        self.calls.append(("set", sorted(vars)))


class _Element:
    """Hooks returning fixed matrices: K = 2 I, M = 3 I, f = K u."""

    material = None

    def assemble_internal_forces(self, mesh, displacement):
        return 2.0 * np.asarray(displacement).reshape(-1)

    def assemble_tangent_stiffness_matrix(self, mesh, displacement):
        return 2.0 * spsp.identity(mesh.n_dofs, format="csr")

    def assemble_stiffness_matrix(self, mesh):
        return 2.0 * spsp.identity(mesh.n_dofs, format="csr")

    def assemble_mass_matrix(self, mesh):
        return 3.0 * spsp.identity(mesh.n_dofs, format="csr")


def _domain():
    X = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    mesh = TetMesh(X, np.array([[0, 1, 2, 3]]))
    mat = _Material()
    return SolidDomain(mesh, mat, _Element()), mat


def test_history_state_passes_through_to_the_material():
    dom, mat = _domain()
    dom.set_time_step(0.5)
    dom.commit_state()
    dom.revert_state()
    dom.set_history_variables({"b": 1, "a": 2})
    dom.reset_state()
    assert mat.calls == [("dt", 0.5), "commit", "revert", ("set", ["a", "b"]), "reset"]
    assert np.array_equal(dom.get_history_variables()["h"], np.arange(3.0))


def test_state_vectors_and_energies():
    dom, _ = _domain()
    u = np.linspace(-1.0, 1.0, dom.n_dofs)
    v = np.linspace(0.0, 2.0, dom.n_dofs)
    dom.update_state(u=u, v=v)
    assert np.array_equal(dom.displacement.reshape(-1), u)
    assert np.array_equal(dom.current_nodes, dom.mesh.nodes + u.reshape(-1, 3))
    assert np.allclose(dom.assemble_internal_forces(), 2.0 * u)
    assert dom.compute_strain_energy() == 0.5 * float(u @ (2.0 * u))
    assert dom.compute_kinetic_energy() == 0.5 * float(v @ (3.0 * v))
    dom.reset_state()
    assert not dom.u.any() and not dom.v.any() and not dom.a.any()
