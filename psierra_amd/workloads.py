"""Seeded synthetic workloads for the Tet4/Tet10 Neo-Hookean assembly path.

Five workloads, each built from ``numpy.random.default_rng(seed)`` with the
draws taken in the order written below:

* ``cantilever_tet4`` (seed 0): 20x4x4 hex box 1.0 x 0.2 x 0.2 m, Kuhn split
  (1920 Tet4, 525 nodes).  Draws: amplitudes a ~ U(-0.01, 0.01, 4), directions
  e ~ N(0,1)^(4x3) normalised per row.  u(X) = sum_k a_k sin(pi k X0 / L) e_k.
  Uniform mu = 1e3 Pa, lambda = 5e3 Pa.
* ``jittered_box_tet4`` (seed 1 for 100^3 hexes): unit cube, interior nodes
  jittered by U(-0.2h, 0.2h) (redrawn while any element has det <= 0), node
  ids permuted by p = rng.permutation(N) with X'[p[i]] = X[i],
  conn' = p[conn]; mu_e = 1e3 exp(0.3 N(0,1)), lambda_e = 5 mu_e; then a, e as
  above and u scaled so max_e ||grad u_e||_F = grad_target (0.1).
* ``jittered_box_tet10`` (seed 2 for 40^3 hexes): Tet4 Kuhn box upgraded to
  Tet10 before jitter, corner nodes jittered, mid-edge nodes placed at the
  midpoints of the jittered corners, all nodes permuted; uniform mu/lambda;
  u scaled so max ||grad u|| at the quadrature points is 0.15.
* ``vjp_box_tet4`` (seed 3, 48^3): as ``jittered_box_tet4`` plus a cotangent
  w ~ N(0,1) per DOF.
* ``ecsw_instances`` (seed 4): 128 boxes with sides U{8..48} per axis.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .mesh import (
    TetMesh,
    box_grid_nodes,
    box_tet4_connectivity,
    boundary_node_mask,
    tet4_to_tet10,
)

_SQ5 = np.sqrt(5.0)
TET10_QP_A = (5.0 - _SQ5) / 20.0
TET10_QP_B = (5.0 + 3.0 * _SQ5) / 20.0


@dataclass
class Workload:
    name: str
    mesh: TetMesh
    u: np.ndarray  # (N, 3) float64
    mu: np.ndarray  # (E,) float64
    lam: np.ndarray  # (E,) float64
    extra: dict = field(default_factory=dict)

    @property
    def n_elements(self) -> int:
        return self.mesh.n_elements


def _smooth_field(rng: np.random.Generator, X: np.ndarray, length: float) -> np.ndarray:
    a = rng.uniform(-0.01, 0.01, 4)
    e = rng.standard_normal((4, 3))
    e /= np.linalg.norm(e, axis=1, keepdims=True)
    k = np.arange(1, 5, dtype=np.float64)
    s = np.sin(np.pi * k[None, :] * X[:, 0:1] / length)  # (N, 4)
    return (s * a[None, :]) @ e


def _tet4_signed_volume6(X: np.ndarray, conn: np.ndarray, chunk: int = 1 << 21) -> np.ndarray:
    out = np.empty(conn.shape[0], dtype=np.float64)
    for s in range(0, conn.shape[0], chunk):
        c = conn[s : s + chunk]
        x0 = X[c[:, 0]]
        d1 = X[c[:, 1]] - x0
        d2 = X[c[:, 2]] - x0
        d3 = X[c[:, 3]] - x0
        out[s : s + chunk] = np.einsum("ij,ij->i", d1, np.cross(d2, d3))
    return out


def tet4_max_grad_norm(X: np.ndarray, conn: np.ndarray, u: np.ndarray, chunk: int = 1 << 20) -> float:
    """max over elements of ||grad u||_F for linear tets."""
    best = 0.0
    for s in range(0, conn.shape[0], chunk):
        c = conn[s : s + chunk]
        D = np.stack([X[c[:, b]] - X[c[:, 0]] for b in (1, 2, 3)], axis=2)  # (m,3,3) columns
        U = np.stack([u[c[:, b]] - u[c[:, 0]] for b in (1, 2, 3)], axis=2)
        G = U @ np.linalg.inv(D)
        best = max(best, float(np.sqrt((G * G).sum(axis=(1, 2))).max()))
    return best


def _tet10_dN(xi: float, eta: float, zeta: float) -> np.ndarray:
    L1, L2, L3 = xi, eta, zeta
    L0 = 1.0 - L1 - L2 - L3
    d = np.zeros((10, 3))
    d[0, :] = -(4.0 * L0 - 1.0)
    d[1, 0] = 4.0 * L1 - 1.0
    d[2, 1] = 4.0 * L2 - 1.0
    d[3, 2] = 4.0 * L3 - 1.0
    d[4] = (4.0 * (L0 - L1), -4.0 * L1, -4.0 * L1)
    d[5] = (4.0 * L2, 4.0 * L1, 0.0)
    d[6] = (-4.0 * L2, 4.0 * (L0 - L2), -4.0 * L2)
    d[7] = (-4.0 * L3, -4.0 * L3, 4.0 * (L0 - L3))
    d[8] = (4.0 * L3, 0.0, 4.0 * L1)
    d[9] = (0.0, 4.0 * L3, 4.0 * L2)
    return d


def tet10_quadrature_points() -> np.ndarray:
    a, b = TET10_QP_A, TET10_QP_B
    return np.array([[a, a, a], [b, a, a], [a, b, a], [a, a, b]])


def tet10_max_grad_norm(X: np.ndarray, conn: np.ndarray, u: np.ndarray, chunk: int = 1 << 18) -> float:
    best = 0.0
    dNs = [_tet10_dN(*q) for q in tet10_quadrature_points()]
    for s in range(0, conn.shape[0], chunk):
        c = conn[s : s + chunk]
        Xe = X[c]
        ue = u[c]
        for dN in dNs:
            J = np.einsum("eai,aj->eij", Xe, dN)
            H = np.einsum("aj,ejk->eak", dN, np.linalg.inv(J))
            G = np.einsum("eai,eaj->eij", ue, H)
            best = max(best, float(np.sqrt((G * G).sum(axis=(1, 2))).max()))
    return best


def cantilever_tet4(seed: int = 0) -> Workload:
    rng = np.random.default_rng(seed)
    nx, ny, nz = 20, 4, 4
    X = box_grid_nodes(nx, ny, nz, 1.0, 0.2, 0.2)
    conn = box_tet4_connectivity(nx, ny, nz)
    u = _smooth_field(rng, X, 1.0)
    E = conn.shape[0]
    mesh = TetMesh(X, conn)
    return Workload(
        "cantilever_tet4_20x4x4",
        mesh,
        u,
        np.full(E, 1.0e3),
        np.full(E, 5.0e3),
    )


def _jitter_and_permute(
    rng: np.random.Generator, X: np.ndarray, conn4: np.ndarray, h: float
) -> tuple[np.ndarray, np.ndarray]:
    interior = ~boundary_node_mask(X)
    while True:
        J = X.copy()
        J[interior] += rng.uniform(-0.2 * h, 0.2 * h, size=(int(interior.sum()), 3))
        if np.all(_tet4_signed_volume6(J, conn4) > 0.0):
            break
    p = rng.permutation(X.shape[0])
    Xp = np.empty_like(J)
    Xp[p] = J
    return Xp, p


def jittered_box_tet4(n: int = 100, seed: int = 1, grad_target: float = 0.1) -> Workload:
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    X0 = box_grid_nodes(n, n, n, 1.0, 1.0, 1.0)
    conn0 = box_tet4_connectivity(n, n, n)
    X, p = _jitter_and_permute(rng, X0, conn0, h)
    conn = p[conn0]
    del conn0
    E = conn.shape[0]
    mu = 1.0e3 * np.exp(0.3 * rng.standard_normal(E))
    lam = 5.0 * mu
    u = _smooth_field(rng, X, 1.0)
    u *= grad_target / tet4_max_grad_norm(X, conn, u)
    return Workload(f"jittered_box_tet4_{n}^3", TetMesh(X, conn), u, mu, lam, {"seed": seed, "h": h})


def jittered_box_tet10(n: int = 40, seed: int = 2, grad_target: float = 0.15) -> Workload:
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    X0 = box_grid_nodes(n, n, n, 1.0, 1.0, 1.0)
    conn4 = box_tet4_connectivity(n, n, n)
    _, conn10 = tet4_to_tet10(X0, conn4)
    interior = ~boundary_node_mask(X0)
    while True:
        Xc = X0.copy()
        Xc[interior] += rng.uniform(-0.2 * h, 0.2 * h, size=(int(interior.sum()), 3))
        if np.all(_tet4_signed_volume6(Xc, conn4) > 0.0):
            break
    X10, conn10b = tet4_to_tet10(Xc, conn4)
    assert np.array_equal(conn10, conn10b)
    p = rng.permutation(X10.shape[0])
    Xp = np.empty_like(X10)
    Xp[p] = X10
    conn = p[conn10]
    E = conn.shape[0]
    u = _smooth_field(rng, Xp, 1.0)
    u *= grad_target / tet10_max_grad_norm(Xp, conn, u)
    return Workload(
        f"jittered_box_tet10_{n}^3",
        TetMesh(Xp, conn),
        u,
        np.full(E, 1.0e3),
        np.full(E, 5.0e3),
        {"seed": seed, "h": h},
    )


def vjp_box_tet4(n: int = 48, seed: int = 3) -> Workload:
    wl = jittered_box_tet4(n=n, seed=seed)
    rng = np.random.default_rng([seed, 1])
    wl.extra["w"] = rng.standard_normal(wl.mesh.n_dofs)
    wl.extra["fd_elements"] = rng.choice(wl.mesh.n_elements, 100, replace=False)
    wl.name = f"vjp_box_tet4_{n}^3"
    return wl


def batch_boxes_tet4(
    n_instances: int = 128, seed: int = 5, side_range: tuple[int, int] = (8, 48), grad_target: float = 0.1
) -> list[Workload]:
    """Seeded meshes of uneven size for batched assembly: box i has sides
    s ~ U{lo..hi} per axis (h = 1/max side), Kuhn Tet4 with jittered interior
    nodes and permuted node ids as in ``jittered_box_tet4``, one mu ~
    logU[3e2, 3e3] per instance with lam = 5 mu, and the smooth displacement
    field scaled so that max ||grad u|| = grad_target."""
    out = []
    for i, child in enumerate(np.random.SeedSequence(seed).spawn(n_instances)):
        rng = np.random.default_rng(child)
        sx, sy, sz = (int(v) for v in rng.integers(side_range[0], side_range[1] + 1, size=3))
        h = 1.0 / max(sx, sy, sz)
        X0 = box_grid_nodes(sx, sy, sz, sx * h, sy * h, sz * h)
        conn0 = box_tet4_connectivity(sx, sy, sz)
        X, p = _jitter_and_permute(rng, X0, conn0, h)
        conn = p[conn0]
        E = conn.shape[0]
        mu = float(np.exp(rng.uniform(np.log(3e2), np.log(3e3))))
        u = _smooth_field(rng, X, sx * h)
        u *= grad_target / tet4_max_grad_norm(X, conn, u)
        out.append(
            Workload(
                f"batch_box_tet4_{sx}x{sy}x{sz}", TetMesh(X, conn), u, np.full(E, mu), np.full(E, 5.0 * mu),
                {"sides": (sx, sy, sz), "instance": i},
            )
        )
    return out


def ecsw_instances(
    n_instances: int = 128,
    seed: int = 4,
    side_range: tuple[int, int] = (8, 48),
    rank: int = 64,
    rid_fraction: float = 0.02,
) -> list[dict]:
    """Seeded instances for the batched ECSW projection.

    Each instance: box with sides s ~ U{lo..hi} per axis (Kuhn Tet4 of the unit
    cube scaled to the side counts, h = 1/max side), mu ~ logU[3e2, 3e3],
    lam = 5 mu, basis V = QR(N(0,1)^(n x rank)), RID = rng.choice(E,
    ceil(rid_fraction E)) with weights U(0.5, 2.0), u_r ~ N(0,1) scaled so that
    max ||grad u|| over the RID is 0.1.
    """
    children = np.random.SeedSequence(seed).spawn(n_instances)
    out = []
    for child in children:
        rng = np.random.default_rng(child)
        sx, sy, sz = (int(v) for v in rng.integers(side_range[0], side_range[1] + 1, size=3))
        h = 1.0 / max(sx, sy, sz)
        X = box_grid_nodes(sx, sy, sz, sx * h, sy * h, sz * h)
        conn = box_tet4_connectivity(sx, sy, sz)
        E = conn.shape[0]
        mu = float(np.exp(rng.uniform(np.log(3e2), np.log(3e3))))
        n_dof = 3 * X.shape[0]
        Q, _ = np.linalg.qr(rng.standard_normal((n_dof, rank)))
        n_rid = int(np.ceil(rid_fraction * E))
        rid = np.sort(rng.choice(E, n_rid, replace=False))
        w = rng.uniform(0.5, 2.0, n_rid)
        ur = rng.standard_normal(rank)
        ue = (Q @ ur).reshape(-1, 3)
        g = tet4_max_grad_norm(X, conn[rid], ue)
        ur *= 0.1 / g
        out.append(
            {
                "sides": (sx, sy, sz),
                "mesh": TetMesh(X, conn),
                "mu": mu,
                "lam": 5.0 * mu,
                "basis": Q,
                "rid": rid,
                "weights": w,
                "u_r": ur,
            }
        )
    return out
