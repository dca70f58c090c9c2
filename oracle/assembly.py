"""FP64 global assembly restated from pSierra (test infrastructure only).

* ``compute_element_dofs`` -- src/psierra/utility/indexing.py:35-59
* ``create_coo_sparsity_indices`` -- src/psierra/utility/assembly.py:6-34
* ``assemble_sparse_matrix`` -- src/psierra/utility/assembly.py:37-76
  (scipy ``coo_matrix(...).tocsr()``: duplicates summed, indices sorted,
  explicit zeros kept)
* ``assemble_vector`` -- src/psierra/utility/assembly.py:79-108 (``np.add.at``)
* ``free_dof_csc`` -- src/psierra/analysis/nonlinear_static.py:354-355
  (``K_tan + K_contact`` with an empty CSR, which drops exact-zero sums,
  then ``[free][:, free].tocsc()``)
* ``assemble`` / ``assemble_loop`` -- SolidDomain.assemble_internal_forces and
  assemble_tangent_stiffness_matrix (src/psierra/domain/solid.py:230-325), the
  first vectorised over element chunks, the second the per-element loop with
  the element evaluated twice, as the reference does.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as spsp

from . import fem


def compute_element_dofs(elements, dofs_per_node=3):
    elements = np.asarray(elements, dtype=np.int64)
    E, npe = elements.shape
    off = np.arange(dofs_per_node)
    return (elements[:, :, None] * dofs_per_node + off[None, None, :]).reshape(E, npe * dofs_per_node)


def create_coo_sparsity_indices(element_dofs):
    E, nd = element_dofs.shape
    rows = np.broadcast_to(element_dofs[:, :, None], (E, nd, nd)).reshape(-1)
    cols = np.broadcast_to(element_dofs[:, None, :], (E, nd, nd)).reshape(-1)
    return rows, cols


def assemble_sparse_matrix(element_matrices, rows, cols, total_dofs):
    data = np.asarray(element_matrices, dtype=np.float64).reshape(-1)
    return spsp.coo_matrix((data, (rows, cols)), shape=(total_dofs, total_dofs)).tocsr()


def assemble_vector(element_vectors, element_dofs, total_dofs):
    g = np.zeros(total_dofs)
    np.add.at(g, element_dofs.ravel(), element_vectors.ravel())
    return g


def structural_pattern(elements, n_nodes):
    """indptr/indices of the structural CSR (all-ones values through scipy)."""
    dofs = compute_element_dofs(elements)
    rows, cols = create_coo_sparsity_indices(dofs)
    n = 3 * n_nodes
    K = spsp.coo_matrix((np.ones(rows.size), (rows, cols)), shape=(n, n)).tocsr()
    return K.indptr, K.indices


def element_arrays(mesh_nodes, elements, u_nodes, idx=None):
    conn = elements if idx is None else elements[idx]
    return mesh_nodes[conn], u_nodes[conn]


def evaluate_elements(nodes, elements, u, mu, lam, want_k=True, chunk=20000):
    """Element f (E, nde) and k (E, nde, nde) in FP64 by the batched restatement."""
    E, npe = elements.shape
    nd = 3 * npe
    f = np.empty((E, nd))
    k = np.empty((E, nd, nd)) if want_k else None
    for s in range(0, E, chunk):
        sl = slice(s, min(E, s + chunk))
        Xe, ue = nodes[elements[sl]], u[elements[sl]]
        ke, fe = fem.element_batch(Xe, ue, mu[sl], lam[sl], want_k)
        f[sl] = fe
        if want_k:
            k[sl] = ke
    return k, f


def assemble(nodes, elements, u, mu, lam, want_k=True, chunk=20000):
    """Global f_int (n,) and K_T (scipy CSR) -- vectorised SolidDomain restatement."""
    k, f = evaluate_elements(nodes, elements, u, mu, lam, want_k, chunk)
    dofs = compute_element_dofs(elements)
    n = 3 * nodes.shape[0]
    fg = assemble_vector(f, dofs, n)
    if not want_k:
        return fg, None
    rows, cols = create_coo_sparsity_indices(dofs)
    return fg, assemble_sparse_matrix(k, rows, cols, n)


def assemble_loop(nodes, elements, u, mu, lam):
    """Per-element loops of SolidDomain (src/psierra/domain/solid.py:252 and :302).

    The element routine is called once in the force loop and again in the
    tangent loop, geometry recomputed each time, as the reference does.
    """
    E, npe = elements.shape
    nd = 3 * npe
    elem_fn = fem.tet4_force_and_tangent if npe == 4 else fem.tet10_force_and_tangent
    f_buf = np.zeros((E, nd))
    k_buf = np.zeros((E, nd, nd))
    for e in range(E):
        c = elements[e]
        _, f_buf[e] = elem_fn(nodes[c], u[c], fem.nh_material(mu[e], lam[e]))
    dofs = compute_element_dofs(elements)
    n = 3 * nodes.shape[0]
    fg = assemble_vector(f_buf, dofs, n)
    for e in range(E):
        c = elements[e]
        k_buf[e], _ = elem_fn(nodes[c], u[c], fem.nh_material(mu[e], lam[e]))
    rows, cols = create_coo_sparsity_indices(dofs)
    return fg, assemble_sparse_matrix(k_buf, rows, cols, n)


def free_dof_csc(K_csr, free_dofs):
    """K_total[free][:, free].tocsc() with K_total = K_tan + empty K_contact."""
    n = K_csr.shape[0]
    K_total = K_csr + spsp.csr_matrix((n, n), dtype=np.float64)
    return K_total[free_dofs, :][:, free_dofs].tocsc()


def node_rows(nodes, elements, u, mu, lam, sample_nodes):
    """Assembled K blocks and f entries for a subset of nodes only.

    Evaluates just the elements incident to ``sample_nodes`` and sums their
    contributions in FP64.  Returns (f_rows (m,3), {a: {b: 3x3 block}}).
    Used to check the engine at sizes where a full scipy assembly is too large.
    """
    sample_nodes = np.asarray(sample_nodes, dtype=np.int64)
    mask = np.isin(elements, sample_nodes).any(axis=1)
    eidx = np.nonzero(mask)[0]
    k, f = evaluate_elements(nodes, elements[eidx], u, mu[eidx], lam[eidx], True)
    conn = elements[eidx]
    out_f = np.zeros((sample_nodes.size, 3))
    pos = {int(a): i for i, a in enumerate(sample_nodes)}
    blocks = {int(a): {} for a in sample_nodes}
    npe = elements.shape[1]
    for j in range(conn.shape[0]):
        for p in range(npe):
            a = int(conn[j, p])
            if a not in blocks:
                continue
            out_f[pos[a]] += f[j, 3 * p : 3 * p + 3]
            for q in range(npe):
                b = int(conn[j, q])
                blk = k[j, 3 * p : 3 * p + 3, 3 * q : 3 * q + 3]
                if b in blocks[a]:
                    blocks[a][b] = blocks[a][b] + blk
                else:
                    blocks[a][b] = blk.copy()
    return out_f, blocks


# ---------------------------------------------------------------------------
# Linear stiffness and consistent mass (src/psierra/domain/solid.py:172-228)
# ---------------------------------------------------------------------------

def _simplex_rule(n1d=6):
    """Quadrature on the reference tetrahedron (volume 1/6) from a collapsed
    (Duffy) tensor product of Gauss-Legendre rules; with n1d = 6 it integrates
    polynomials of total degree <= 2 n1d - 3 exactly."""
    x, w = np.polynomial.legendre.leggauss(n1d)
    x = 0.5 * (x + 1.0)
    w = 0.5 * w
    pts, wts = [], []
    for a, wa in zip(x, w):
        for b, wb in zip(x, w):
            for c, wc in zip(x, w):
                # (a, b, c) in the unit cube -> tetrahedron
                xi = a * (1 - b) * (1 - c)
                eta = b * (1 - c)
                zeta = c
                jac = (1 - b) * (1 - c) ** 2
                pts.append((xi, eta, zeta))
                wts.append(wa * wb * wc * jac)
    return np.array(pts), np.array(wts)


def _shape_values(npe, pts):
    xi, eta, zeta = pts[:, 0], pts[:, 1], pts[:, 2]
    L = np.stack([1.0 - xi - eta - zeta, xi, eta, zeta], axis=1)
    if npe == 4:
        return L
    N = np.empty((len(pts), 10))
    for i in range(4):
        N[:, i] = L[:, i] * (2 * L[:, i] - 1)
    for k, (i, j) in enumerate(((0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3))):
        N[:, 4 + k] = 4 * L[:, i] * L[:, j]
    return N


def reference_mass_tensor(npe):
    """(1/V) * integral of N_p N_q over the element by numerical quadrature
    of the shape functions (Tet4: the linear ones of tet4_base.py; Tet10: the
    quadratic VTK-ordered ones of tet10_base.py:117-179)."""
    pts, wts = _simplex_rule()
    N = _shape_values(npe, pts)
    return 6.0 * np.einsum("q,qa,qb->ab", wts, N, N)


def element_volumes(nodes, elements):
    """Volume of each element's corner tetrahedron (straight edges)."""
    X = nodes[elements[:, :4]]
    return np.linalg.det(X[:, 1:] - X[:, :1]) / 6.0


def mass_matrix(nodes, elements, rho):
    """SolidDomain.assemble_mass_matrix (solid.py:208-228): element matrices
    rho V_e c_ab I3 (tet4_base.py:138-151, tet10_base.py:344-372), summed
    with coo.tocsr in the structural pattern (explicit zeros kept)."""
    npe = elements.shape[1]
    c = reference_mass_tensor(npe)
    V = element_volumes(nodes, elements)
    rho = np.broadcast_to(np.asarray(rho, dtype=np.float64), V.shape)
    me = (rho * V)[:, None, None] * np.kron(c, np.eye(3))[None]
    dofs = compute_element_dofs(elements)
    rows, cols = create_coo_sparsity_indices(dofs)
    return assemble_sparse_matrix(me, rows, cols, 3 * nodes.shape[0])


def linear_stiffness(nodes, elements, mu, lam):
    """SolidDomain.assemble_stiffness_matrix (solid.py:172-206) for Tet4:
    k_e = B_L^T C B_L V with C Hooke's law from (lambda_e, mu_e)
    (tet4_base.py:129-136, neohookean.py:33-60)."""
    E = elements.shape[0]
    mu = np.broadcast_to(np.asarray(mu, dtype=np.float64), (E,))
    lam = np.broadcast_to(np.asarray(lam, dtype=np.float64), (E,))
    ke = np.empty((E, 12, 12))
    for e in range(E):
        V, H = fem.tet4_volume_and_gradients(nodes[elements[e]])
        BL = fem.linear_B(H)
        ke[e] = BL.T @ fem.hookean_C(lam[e], mu[e]) @ BL * V
    dofs = compute_element_dofs(elements)
    rows, cols = create_coo_sparsity_indices(dofs)
    return assemble_sparse_matrix(ke, rows, cols, 3 * nodes.shape[0])
