"""Full-size Tet4 workload (100^3 hexes, 6.0M tets, 1.03M nodes, 3.09M DOF).

A full scipy assembly at this size needs > 35 GB, so parity is checked by
(a) the oracle on a seeded sample of node rows (all elements touching them
evaluated in FP64), (b) size-independent properties: nnz closed form,
sum of f_int = 0 (each element's forces cancel), K_T t = 0 for the three
rigid translations t (f is translation invariant), (c) FP32 vs FP64 on the
whole matrix within the 1e-5 relative tolerance (the FP64 result itself
being pinned to the oracle at 1e-12 on the sampled rows).
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import assembly as A
from psierra_amd import workloads as W
from tests.gpu_helpers import banner, rel_err
from tests.test_oracle import _kuhn_nnz

pytestmark = pytest.mark.gpu

N_HEX = 100


@pytest.fixture(scope="module")
def big():
    import torch

    from psierra_amd.engine import AssemblyPlan

    banner("fullsize workload build")
    wl = W.jittered_box_tet4(n=N_HEX, seed=1)
    plan = AssemblyPlan(wl.mesh.nodes, wl.mesh.elements)
    dev = plan.device
    u = torch.as_tensor(wl.u.reshape(-1), device=dev)
    mu = torch.as_tensor(wl.mu, device=dev)
    lam = torch.as_tensor(wl.lam, device=dev)
    banner("fullsize assemble fp64")
    f64, k64 = plan.assemble(u, mu, lam, precision="fp64")
    banner("fullsize assemble fp32")
    f32, k32 = plan.assemble(u, mu, lam, precision="fp32")
    torch.cuda.synchronize()
    return wl, plan, f64.cpu().numpy(), k64.cpu().numpy(), f32.cpu().numpy(), k32.cpu().numpy()


def test_fullsize_pattern(big):
    wl, plan, *_ = big
    assert wl.mesh.n_elements == 6_000_000 and wl.mesh.n_nodes == 1_030_301
    assert plan.nnz == _kuhn_nnz(N_HEX, N_HEX, N_HEX) == 136_898_109


def test_fullsize_sampled_rows_match_oracle(big):
    wl, plan, f64, k64, f32, k32 = big
    indptr, indices = plan.pattern()
    rng = np.random.default_rng(2024)
    sample = rng.choice(wl.mesh.n_nodes, 40, replace=False)
    f_ref, blocks = A.node_rows(wl.mesh.nodes, wl.mesh.elements, wl.u, wl.mu, wl.lam, sample)
    for i, a in enumerate(sample):
        a = int(a)
        cols = indices[indptr[3 * a] : indptr[3 * a + 1]]
        nb = cols[::3] // 3
        assert sorted(blocks[a].keys()) == list(nb)
        ref = np.zeros((3, cols.size))
        for s, b in enumerate(nb):
            ref[:, 3 * s : 3 * s + 3] = blocks[a][int(b)]
        got64 = np.stack([k64[indptr[3 * a + r] : indptr[3 * a + r + 1]] for r in range(3)])
        got32 = np.stack([k32[indptr[3 * a + r] : indptr[3 * a + r + 1]] for r in range(3)])
        assert rel_err(got64, ref) < 1e-12
        assert rel_err(got32, ref) < 1e-5
        assert rel_err(f64[3 * a : 3 * a + 3], f_ref[i]) < 1e-10
    # f over the sampled rows at the stated 1e-5 (a single node's f is a sum
    # of ~24 element forces that nearly cancel, so it is compared as a whole)
    got32 = np.stack([f32[3 * int(a) : 3 * int(a) + 3] for a in sample])
    assert rel_err(got32, f_ref) < 1e-5


def test_fullsize_invariants(big):
    import scipy.sparse as spsp

    wl, plan, f64, k64, f32, k32 = big
    indptr, indices = plan.pattern()
    n = plan.n_dofs
    fs = f64.reshape(-1, 3).sum(axis=0)
    assert np.abs(fs).max() < 1e-9 * np.abs(f64).max() * np.sqrt(wl.mesh.n_nodes)
    K = spsp.csr_matrix((k64, indices, indptr), shape=(n, n))
    for c in range(3):
        t = np.zeros(n)
        t[c::3] = 1.0
        r = K @ t
        assert np.abs(r).max() < 1e-9 * np.abs(k64).max()


def test_fullsize_fp32_within_tolerance(big):
    wl, plan, f64, k64, f32, k32 = big
    ek, ef = rel_err(k32, k64), rel_err(f32, f64)
    print(f"fullsize fp32 vs fp64: K rel {ek:.3e}, f rel {ef:.3e}")
    assert ek < 1e-5 and ef < 1e-5


def test_fullsize_free_csc(big):
    """Free-DOF CSC of the full matrix clamped at x = 0 (3,060,300 columns, so
    the column-count scan runs over many work-groups and its top level over
    many groups per thread).  The pattern is symmetric, so column j of
    K[free][:, free] holds the free columns of row j: every column's rows and
    the column pointers are checked against that, the values of 300 sampled
    columns against scipy's column slices of the host matrix, and the same
    again with every entry of 64 free rows set to zero (some to -0.0): those
    rows leave every column, which takes the compacting pass."""
    import scipy.sparse as spsp
    import torch

    from psierra_amd.boundary import DirichletConstraint, DirichletManager, select_nodes_by_plane
    from psierra_amd.solver import FreeDofSystem

    wl, plan, f64, k64, f32, k32 = big
    indptr, indices = plan.pattern()
    n = plan.n_dofs
    mgr = DirichletManager(n)
    mgr.add_constraint(DirichletConstraint(select_nodes_by_plane(wl.mesh, "yz", "min"), (0, 1, 2), 0.0))
    free = np.asarray(mgr.free_dofs)
    assert free.size == 3_060_300
    isfree = np.zeros(n, bool)
    isfree[free] = True
    newidx = np.full(n, -1, np.int64)
    newidx[free] = np.arange(free.size)
    row_of = np.repeat(np.arange(n), np.diff(indptr))
    rng = np.random.default_rng(77)
    sample = np.sort(rng.choice(free.size, 300, replace=False))
    zero_rows = np.sort(rng.choice(free, 64, replace=False))
    sysm = FreeDofSystem(plan, free)

    def check(kv, dropped):
        banner(f"fullsize free csc, {int(dropped.sum())} zeroed rows")
        got = sysm.csc(torch.as_tensor(kv, device=plan.device))
        sel = isfree[row_of] & isfree[indices] & ~dropped[indices]
        counts = np.bincount(row_of[sel], minlength=n)[free]
        assert np.array_equal(got.indptr, np.concatenate([[0], np.cumsum(counts)]))
        assert np.array_equal(got.indices, newidx[indices[sel]])
        K = spsp.csr_matrix((kv, indices, indptr), shape=(n, n))
        cols = K[:, free[sample]].tocsc()
        for t, j in enumerate(sample):
            rows = cols.indices[cols.indptr[t] : cols.indptr[t + 1]]
            vals = cols.data[cols.indptr[t] : cols.indptr[t + 1]]
            order = np.argsort(rows)
            rows, vals = rows[order], vals[order]
            keep = isfree[rows] & (vals != 0.0)
            lo, hi = got.indptr[j], got.indptr[j + 1]
            assert np.array_equal(got.indices[lo:hi], newidx[rows[keep]])
            assert np.array_equal(got.data[lo:hi], vals[keep])
        return got.nnz

    none = np.zeros(n, bool)
    nnz_all = check(k64, none)
    kz = k64.copy()
    for i, r in enumerate(zero_rows):
        kz[indptr[r] : indptr[r + 1]] = -0.0 if i % 4 == 0 else 0.0
    dropped = np.zeros(n, bool)
    dropped[zero_rows] = True
    assert check(kz, dropped) < nnz_all
    assert check(k64, none) == nnz_all
