"""GPU parity of the Tet10 Neo-Hookean assembly against the FP64 oracle.

Tolerances as for Tet4: FP64 within 1e-12 (K) / 1e-11 (f) relative,
FP32 element arithmetic with FP64 accumulation within 1e-5.  Pattern bit-exact.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import assembly as A
from oracle import fem
from psierra_amd import workloads as W
from tests.conftest import rotation_30deg_111
from tests.gpu_helpers import banner, gpu_assemble, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tet10_box():
    wl = W.jittered_box_tet10(n=4, seed=12)
    f_ref, K_ref = A.assemble(wl.mesh.nodes, wl.mesh.elements, wl.u, wl.mu, wl.lam)
    return wl, f_ref, K_ref


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_tet10_assembly_matches_oracle(tet10_box, precision):
    wl, f_ref, K_ref = tet10_box
    banner(f"tet10 assemble {precision}")
    plan, f, k = gpu_assemble(wl, precision)
    indptr, indices = plan.pattern()
    assert np.array_equal(indptr, K_ref.indptr) and np.array_equal(indices, K_ref.indices)
    ek, ef = rel_err(k, K_ref.data), rel_err(f, f_ref)
    print(f"tet10 {precision}: K rel {ek:.3e}, f rel {ef:.3e}")
    if precision == "fp64":
        assert ek < 1e-12 and ef < 1e-11
    else:
        assert ek < 1e-5 and ef < 1e-5


def test_tet10_deterministic(tet10_box):
    wl = tet10_box[0]
    banner("tet10 determinism")
    _, f1, k1 = gpu_assemble(wl, "fp32", tile=64)
    _, f2, k2 = gpu_assemble(wl, "fp32", tile=128)
    assert np.array_equal(k1, k2) and np.array_equal(f1, f2)


def test_tet10_force_only(tet10_box):
    wl = tet10_box[0]
    banner("tet10 f-only")
    _, f_full, _ = gpu_assemble(wl, "fp64")
    _, f_only, _ = gpu_assemble(wl, "fp64", want_k=False)
    assert rel_err(f_only, f_full) < 1e-13


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_tet10_element_batch(regular_tet10_coords, soft_lame, precision):
    from psierra_amd.engine import element_batch

    lam, mu = soft_lame
    X = regular_tet10_coords
    rng = np.random.default_rng(42)
    us = [0.01 * rng.standard_normal((10, 3)) for _ in range(4)]
    us.append((rotation_30deg_111() @ X.T).T - X)
    Xb = np.stack([X] * len(us))
    banner(f"tet10 element batch {precision}")
    k, f = element_batch(Xb, np.stack(us), [mu] * len(us), [lam] * len(us), precision=precision)
    k, f = k.cpu().numpy(), f.cpu().numpy()
    for j, u in enumerate(us[:-1]):
        kr, fr = fem.tet10_force_and_tangent(X, u, fem.nh_material(mu, lam))
        assert rel_err(k[j], kr) < (1e-12 if precision == "fp64" else 1e-5)
        assert rel_err(f[j], fr) < (1e-10 if precision == "fp64" else 1e-4)
    assert np.allclose(f[-1], 0.0, atol=1e-7 if precision == "fp64" else 1e-3)
