"""CPU checks of the native library: it loads, exports every symbol the public
header declares, and its host pattern builder reproduces scipy's
coo_matrix(...).tocsr() pattern bit for bit (src/psierra/utility/assembly.py:6-76)."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import assembly as A
from psierra_amd import _lib
from psierra_amd.engine import pattern_host
from psierra_amd.mesh import box_tet4, tet4_to_tet10

HEADER = Path(__file__).resolve().parents[1] / "include" / "psierra_amd.h"


def header_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(psx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    L = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 10
    for s in syms:
        assert hasattr(L, s), s
        assert s in _lib.SIGNATURES, f"{s} has no ctypes signature"
    assert L.psx_abi_version() == _lib.ABI_VERSION == 7


def _random_mesh(seed, npe):
    m = box_tet4(4, 3, 3, 1.0, 0.7, 0.6)
    nodes, conn = m.nodes, m.elements
    if npe == 10:
        nodes, conn = tet4_to_tet10(nodes, conn)
    rng = np.random.default_rng(seed)
    p = rng.permutation(nodes.shape[0])
    conn = p[conn]
    # drop a few elements and leave an unused node in the middle of the range
    keep = rng.random(conn.shape[0]) > 0.1
    return nodes.shape[0] + 1, conn[keep]


@pytest.mark.parametrize("npe", [4, 10])
@pytest.mark.parametrize("seed", [0, 1])
def test_pattern_bit_exact_vs_scipy(npe, seed):
    n_nodes, conn = _random_mesh(seed, npe)
    ip_ref, ix_ref = A.structural_pattern(conn, n_nodes)
    ip, ix = pattern_host(conn, n_nodes)
    assert ip.dtype == np.int32 and ix.dtype == np.int32
    assert np.array_equal(ip, ip_ref)
    assert np.array_equal(ix, ix_ref)


def test_pattern_cantilever_counts():
    m = box_tet4(20, 4, 4, 1.0, 0.2, 0.2)
    ip, ix = pattern_host(m.elements, m.n_nodes)
    assert ip[-1] == 55053 and ix.size == 55053


def test_pattern_rejects_bad_connectivity():
    conn = np.array([[0, 1, 2, 7]], dtype=np.int64)
    with pytest.raises(ValueError, match="outside"):
        pattern_host(conn, 4)
    with pytest.raises(ValueError, match="npe"):
        L = _lib.lib()
        nnz = ctypes.c_int64()
        _lib.check(L.psx_pattern_nnz(_lib.ptr(np.zeros((1, 5), np.int64)), 1, 5, 6, ctypes.byref(nnz)))


def test_last_error_is_per_call():
    L = _lib.lib()
    nnz = ctypes.c_int64()
    assert L.psx_pattern_nnz(_lib.ptr(np.array([[0, 1, 2, 9]], np.int64)), 1, 4, 4, ctypes.byref(nnz)) != 0
    assert "outside" in _lib.last_error()
    ok = np.array([[0, 1, 2, 3]], np.int64)
    assert L.psx_pattern_nnz(_lib.ptr(ok), 1, 4, 4, ctypes.byref(nnz)) == 0
    assert _lib.last_error() == "" and nnz.value == 144


@pytest.mark.parametrize("seed", [0, 1])
def test_fbar_pattern_bit_exact_vs_oracle(seed):
    from oracle import fbar as FB
    from psierra_amd.engine import fbar_pattern_host

    n_nodes, conn = _random_mesh(seed, 4)
    ip_ref, ix_ref = FB.distance2_pattern(conn, n_nodes)
    ip, ix = fbar_pattern_host(conn, n_nodes)
    assert ip.dtype == np.int32 and ix.dtype == np.int32
    assert np.array_equal(ip, ip_ref)
    assert np.array_equal(ix, ix_ref)
    # it contains the element pattern
    ip1, ix1 = pattern_host(conn, n_nodes)
    for r in range(0, 3 * n_nodes, 7):
        assert set(ix1[ip1[r]:ip1[r + 1]]) <= set(ix[ip[r]:ip[r + 1]])


def test_batch_entry_points_reject_bad_arguments_before_any_device_call():
    """The batch entry points check their arguments first (no GPU needed):
    null or empty plan lists, null handles and pointer tables, mismatched
    parameter tables, an unknown precision."""
    L = _lib.lib()
    out = ctypes.c_void_p()
    assert L.psx_batch_create(None, 0, ctypes.byref(out)) == _lib.PSX_EINVAL
    assert "n >= 1" in _lib.last_error()
    plans = (ctypes.c_void_p * 2)(None, None)
    assert L.psx_batch_create(plans, 2, ctypes.byref(out)) == _lib.PSX_EINVAL
    assert "null plan" in _lib.last_error()
    assert L.psx_batch_destroy(None) == _lib.PSX_OK
    assert L.psx_batch_set_persistent(None, 1) == _lib.PSX_EINVAL
    info = (ctypes.c_int64 * 6)()
    assert L.psx_batch_info(None, info, 6) == _lib.PSX_EINVAL
    tab = (ctypes.c_void_p * 1)(None)
    assert L.psx_assemble_batch(None, tab, None, None, tab, None, 0, None, tab) == _lib.PSX_EINVAL
    assert "null argument to psx_assemble_batch" in _lib.last_error()
    assert L.psx_vjp_batch(None, tab, tab, tab, tab, tab, None, None, 0, None, tab) == _lib.PSX_EINVAL
    assert "null argument to psx_vjp_batch" in _lib.last_error()
