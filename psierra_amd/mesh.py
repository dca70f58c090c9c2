"""Host-side tetrahedral meshes for the assembly engine.

The engine only needs what pSierra's ``VolumetricMesh`` exposes to the solid
domain (``src/psierra/mesh/volumetric.py:97-124,212-224``): ``nodes`` (N, 3)
float64, ``elements`` (E, npe) int64 and the counts.  ``TetMesh`` is that
duck type without the pyvista dependency.

``box_tet4`` builds a structured box split into Kuhn tetrahedra (six per hex,
all sharing the hex's main diagonal, positively oriented as required by
``src/psierra/element/tet4_base.py:56-61``).  ``tet4_to_tet10`` adds mid-edge
nodes with the numbering of ``src/psierra/mesher/order.py:11-81``.
"""
from __future__ import annotations

import numpy as np

# The six axis orders of a Kuhn split.  Tet k of a hex walks from corner
# (0,0,0) along axes order[0], order[1], order[2] to corner (1,1,1).
_KUHN_ORDERS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))
_KUHN_PARITY = (1, -1, -1, 1, 1, -1)

# VTK quadratic tetra edge order: 01, 12, 20, 03, 13, 23
# (src/psierra/element/tet10_base.py:32-40, src/psierra/mesher/order.py:52-57).
TET10_EDGES = ((0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3))


class TetMesh:
    """Minimal volumetric mesh: reference coordinates plus connectivity."""

    def __init__(self, nodes: np.ndarray, elements: np.ndarray) -> None:
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        self.elements = np.ascontiguousarray(elements, dtype=np.int64)
        if self.nodes.ndim != 2 or self.nodes.shape[1] != 3:
            raise ValueError(f"nodes must have shape (N, 3), got {self.nodes.shape}")
        if self.elements.ndim != 2 or self.elements.shape[1] not in (4, 10):
            raise ValueError(
                f"elements must have shape (E, 4) or (E, 10), got {self.elements.shape}"
            )

    @property
    def n_nodes(self) -> int:
        return int(self.nodes.shape[0])

    @property
    def n_elements(self) -> int:
        return int(self.elements.shape[0])

    @property
    def n_dofs(self) -> int:
        return 3 * self.n_nodes

    @property
    def nodes_per_element(self) -> int:
        return int(self.elements.shape[1])


def box_grid_nodes(nx: int, ny: int, nz: int, lx: float, ly: float, lz: float) -> np.ndarray:
    """Nodes of an (nx+1)(ny+1)(nz+1) grid; node (i,j,k) has id i + (nx+1)(j + (ny+1)k)."""
    xs = np.linspace(0.0, lx, nx + 1)
    ys = np.linspace(0.0, ly, ny + 1)
    zs = np.linspace(0.0, lz, nz + 1)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)


def box_tet4_connectivity(nx: int, ny: int, nz: int) -> np.ndarray:
    """Kuhn split of an nx*ny*nz hex grid: hex (i,j,k) -> elements 6*(i + nx(j + ny k)) + 0..5."""
    sx, sy = 1, nx + 1
    sz = (nx + 1) * (ny + 1)
    step = np.array([sx, sy, sz], dtype=np.int64)
    k, j, i = np.meshgrid(
        np.arange(nz, dtype=np.int64),
        np.arange(ny, dtype=np.int64),
        np.arange(nx, dtype=np.int64),
        indexing="ij",
    )
    base = (i * sx + j * sy + k * sz).ravel()
    diag = int(step.sum())
    tets = np.empty((base.size, 6, 4), dtype=np.int64)
    for t, (order, parity) in enumerate(zip(_KUHN_ORDERS, _KUHN_PARITY)):
        p1 = step[order[0]]
        p2 = p1 + step[order[1]]
        if parity > 0:
            local = (0, p1, p2, diag)
        else:
            local = (0, p2, p1, diag)
        for c in range(4):
            tets[:, t, c] = base + local[c]
    return tets.reshape(-1, 4)


def box_tet4(nx: int, ny: int, nz: int, lx: float = 1.0, ly: float = 1.0, lz: float = 1.0) -> TetMesh:
    return TetMesh(box_grid_nodes(nx, ny, nz, lx, ly, lz), box_tet4_connectivity(nx, ny, nz))


def boundary_node_mask(nodes: np.ndarray, tol: float = 1e-12) -> np.ndarray:
    """Nodes on the bounding box faces."""
    lo = nodes.min(axis=0)
    hi = nodes.max(axis=0)
    return np.any((np.abs(nodes - lo) <= tol) | (np.abs(nodes - hi) <= tol), axis=1)


def tet4_to_tet10(nodes: np.ndarray, elements: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Quadratic upgrade with pSierra's numbering.

    Follows ``compute_tet10_connectivity_and_points``
    (``src/psierra/mesher/order.py:11-81``): mid-edge node of the undirected
    edge (min, max) gets id ``N + rank`` of that edge among the unique edges
    sorted lexicographically; element columns 4..9 follow ``TET10_EDGES``;
    coordinates are edge midpoints.  The lexicographic (min, max) rank is
    obtained here from the scalar key ``min * N + max``, which sorts the same.
    """
    pts = np.asarray(nodes, dtype=np.float64)
    c4 = np.asarray(elements, dtype=np.int64)
    if c4.ndim != 2 or c4.shape[1] != 4:
        raise ValueError(f"Expected elements_tet4 of shape (N, 4), got {c4.shape}.")
    n = pts.shape[0]
    a = np.stack([c4[:, e0] for e0, _ in TET10_EDGES], axis=1)
    b = np.stack([c4[:, e1] for _, e1 in TET10_EDGES], axis=1)
    key = np.minimum(a, b) * n + np.maximum(a, b)
    uniq, inv = np.unique(key.ravel(), return_inverse=True)
    lo = uniq // n
    hi = uniq % n
    mid = 0.5 * (pts[lo] + pts[hi])
    conn10 = np.hstack([c4, n + inv.reshape(-1, 6)])
    return np.vstack([pts, mid]), conn10


def element_node_coords(nodes: np.ndarray, elements: np.ndarray) -> np.ndarray:
    return nodes[elements]
