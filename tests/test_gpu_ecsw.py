"""Batched ECSW projection on FP32 matrix cores vs the FP64 oracle.

Six seeded instances of uneven size (boxes with 4..9 hexes per side, rank 64
orthonormal bases, 5% RID with weights U(0.5, 2)) in one launch.  K_r and
f_r within 1e-4 relative (Frobenius) of the FP64 restatement of
HyperReducedECSWModel.evaluate_forces_and_stiffness; K_r exactly symmetric
(model.py:317); the single-instance wrapper equals the batch.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import ecsw as OE
from psierra_amd import workloads as W
from tests.gpu_helpers import banner, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def instances():
    return W.ecsw_instances(n_instances=6, seed=44, side_range=(4, 9), rid_fraction=0.05)


def test_batched_ecsw_matches_oracle(instances):
    from psierra_amd.ecsw import BatchedECSW

    banner("ecsw batched")
    b = BatchedECSW(instances)
    ur = np.stack([inst["u_r"] for inst in instances])
    f, K = b.evaluate(ur)
    f, K = f.cpu().numpy(), K.cpu().numpy()
    for i, inst in enumerate(instances):
        m = inst["mesh"]
        fr, Kr = OE.evaluate_forces_and_stiffness(m.nodes, m.elements, inst["basis"], inst["rid"],
                                                  inst["weights"], inst["mu"], inst["lam"], inst["u_r"])
        ek, ef = rel_err(K[i], Kr), rel_err(f[i], fr)
        print(f"instance {i} sides {inst['sides']}: K_r rel {ek:.2e}, f_r rel {ef:.2e}")
        assert ek < 1e-4 and ef < 1e-4
        assert np.array_equal(K[i], K[i].T)


def test_single_instance_model(instances):
    from psierra_amd.ecsw import BatchedECSW, HyperReducedECSWModel

    inst = instances[2]
    banner("ecsw single")
    model = HyperReducedECSWModel(inst["mesh"], inst["basis"], inst["rid"], inst["weights"], inst["mu"], inst["lam"])
    f1, K1 = model.evaluate_forces_and_stiffness(inst["u_r"])
    f2, K2 = BatchedECSW([inst]).evaluate(inst["u_r"][None])
    assert np.array_equal(f1, f2[0].cpu().numpy()) and np.array_equal(K1, K2[0].cpu().numpy())


def test_rid_sizes_at_batch_and_chunk_edges(instances):
    """RIDs of 1, 5, 8, 17, 128 and 129 elements: below one matrix-core batch,
    exactly one, across a 16-element row batch, exactly one 128-element work
    item and one past it; each equal to the oracle."""
    from psierra_amd.ecsw import BatchedECSW

    banner("ecsw rid edges")
    big = max(instances, key=lambda inst: inst["mesh"].n_elements)
    rng = np.random.default_rng(7)
    cases = []
    for n in (1, 5, 8, 17, 128, 129):
        assert n <= big["mesh"].n_elements
        inst = dict(big)
        inst["rid"] = np.sort(rng.choice(big["mesh"].n_elements, n, replace=False)).astype(np.int64)
        inst["weights"] = rng.uniform(0.5, 2.0, n)
        cases.append(inst)
    b = BatchedECSW(cases)
    f, K = b.evaluate(np.stack([c["u_r"] for c in cases]))
    f, K = f.cpu().numpy(), K.cpu().numpy()
    for i, c in enumerate(cases):
        m = c["mesh"]
        fr, Kr = OE.evaluate_forces_and_stiffness(m.nodes, m.elements, c["basis"], c["rid"], c["weights"],
                                                  c["mu"], c["lam"], c["u_r"])
        ek, ef = rel_err(K[i], Kr), rel_err(f[i], fr)
        print(f"rid {len(c['rid'])}: K_r rel {ek:.2e}, f_r rel {ef:.2e}")
        assert ek < 1e-4 and ef < 1e-4
        assert np.array_equal(K[i], K[i].T)
