"""Shared fixtures; the ``gpu`` marker selects tests that need an MI355X."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and the built engine library")


@pytest.fixture
def regular_tet_coords() -> np.ndarray:
    """Unit right tetrahedron, V = 1/6 (as the reference's tests/conftest.py:21-36)."""
    return np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])


@pytest.fixture
def inverted_tet_coords() -> np.ndarray:
    return np.array([[0.0, 0.0, 0.0], [0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])


@pytest.fixture
def regular_tet10_coords(regular_tet_coords) -> np.ndarray:
    v = regular_tet_coords
    c = np.zeros((10, 3))
    c[:4] = v
    for k, (a, b) in enumerate(((0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3))):
        c[4 + k] = 0.5 * (v[a] + v[b])
    return c


@pytest.fixture
def steel_lame():
    """Hookean fixture of the reference: E = 200 GPa, nu = 0.3 -> (lambda, mu)."""
    from oracle.fem import lame_from_young_poisson

    return lame_from_young_poisson(200e9, 0.3)


@pytest.fixture
def soft_lame():
    """Neo-Hookean fixture of the reference: E = 10 kPa, nu = 0.45 -> (lambda, mu)."""
    from oracle.fem import lame_from_young_poisson

    return lame_from_young_poisson(10e3, 0.45)


def rotation_30deg_111() -> np.ndarray:
    th = np.deg2rad(30.0)
    ax = np.ones(3) / np.sqrt(3.0)
    K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    return np.eye(3) + np.sin(th) * K + (1.0 - np.cos(th)) * (K @ K)
