"""ctypes binding of libpsierra_amd.so (the C ABI declared in include/psierra_amd.h).

The library is built in-tree by ``python -m psierra_amd.build``.  Loading
fails loudly if it is missing: the engine has no CPU fallback.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_int32, c_int64, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libpsierra_amd.so"

PSX_OK = 0
PSX_EINVAL = 1
PSX_EUNSUPPORTED = 2
PSX_EVOLUME = 3
PSX_ENOTREADY = 4
PSX_FP64 = 0
PSX_FP32 = 1

INFO_KEYS = (
    "n_elements",
    "n_nodes",
    "npe",
    "nnz",
    "n_tiles",
    "tile",
    "max_degree",
    "device_bytes",
    "n_incidences",
    "max_incidences",
    "wide",
)

_i64p = POINTER(c_int64)
_i32p = POINTER(c_int32)
_f64p = POINTER(c_double)

# name -> (restype, argtypes)
SIGNATURES = {
    "psx_abi_version": (c_int32, []),
    "psx_plan_bind_material": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "psx_last_error": (c_char_p, []),
    "psx_pattern_nnz": (c_int32, [c_void_p, c_int64, c_int32, c_int64, _i64p]),
    "psx_pattern_build": (c_int32, [c_void_p, c_int64, c_int32, c_int64, c_void_p, c_void_p]),
    "psx_plan_create": (c_int32, [c_void_p, c_int64, c_int32, c_void_p, c_int64, c_int32, POINTER(c_void_p)]),
    "psx_plan_destroy": (c_int32, [c_void_p]),
    "psx_plan_info": (c_int32, [c_void_p, _i64p, c_int32]),
    "psx_plan_pattern": (c_int32, [c_void_p, c_void_p, c_void_p]),
    "psx_plan_set_tile": (c_int32, [c_void_p, c_int32]),
    "psx_assemble": (
        c_int32,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p],
    ),
    "psx_vjp": (
        c_int32,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p],
    ),
    "psx_csc_create": (c_int32, [c_void_p, c_void_p, c_int64, POINTER(c_void_p)]),
    "psx_csc_destroy": (c_int32, [c_void_p]),
    "psx_csc_info": (c_int32, [c_void_p, _i64p, _i64p]),
    "psx_csc_extract": (c_int32, [c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p]),
    "psx_free_residual": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "psx_ecsw_create": (
        c_int32,
        [c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, POINTER(c_void_p)],
    ),
    "psx_ecsw_destroy": (c_int32, [c_void_p]),
    "psx_ecsw_info": (c_int32, [c_void_p, _i64p, c_int32]),
    "psx_ecsw_evaluate": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "psx_assemble_mass": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "psx_fbar_create": (c_int32, [c_void_p, POINTER(c_void_p)]),
    "psx_fbar_destroy": (c_int32, [c_void_p]),
    "psx_fbar_nnz": (c_int32, [c_void_p, _i64p]),
    "psx_fbar_pattern": (c_int32, [c_void_p, c_void_p, c_void_p]),
    "psx_fbar_pattern_nnz": (c_int32, [c_void_p, c_int64, c_int64, _i64p]),
    "psx_fbar_pattern_build": (c_int32, [c_void_p, c_int64, c_int64, c_void_p, c_void_p]),
    "psx_fbar_csc_create": (c_int32, [c_void_p, c_void_p, c_int64, POINTER(c_void_p)]),
    "psx_fbar_assemble": (
        c_int32,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p, c_int32, c_void_p, c_void_p],
    ),
    "psx_assemble_hyperelastic": (
        c_int32,
        [c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p],
    ),
    "psx_ecsw_training": (
        c_int32,
        [c_void_p, c_void_p, c_int32, c_void_p, c_int32, c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_int64,
         c_void_p, c_void_p, c_void_p],
    ),
    "psx_ecsw_training_work_doubles": (c_int64, [c_void_p]),
    "psx_batch_create": (c_int32, [c_void_p, c_int32, POINTER(c_void_p)]),
    "psx_vjp_batch": (
        c_int32,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p],
    ),
    "psx_batch_destroy": (c_int32, [c_void_p]),
    "psx_batch_set_persistent": (c_int32, [c_void_p, c_int32]),
    "psx_batch_info": (c_int32, [c_void_p, _i64p, c_int32]),
    "psx_assemble_batch": (
        c_int32,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p],
    ),
    "psx_ecsw_training_noalloc": (
        c_int32,
        [c_void_p, c_void_p, c_int32, c_void_p, c_int32, c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_void_p,
         c_void_p],
    ),
    "psx_element_batch": (
        c_int32,
        [c_int32, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p],
    ),
    "psx_nh_material_batch": (
        c_int32,
        [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p],
    ),
}

_lib = None

# The C ABI these signatures describe (psx_abi_version() of the library).
ABI_VERSION = 7


class PsxError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"psierra_amd error {code}: {message}")
        self.code = code
        self.message = message


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python -m psierra_amd.build` (the engine has no CPU fallback)"
            )
        L = ctypes.CDLL(str(LIB_PATH))
        # the version first: a library of another version may lack symbols
        L.psx_abi_version.restype = c_int32
        L.psx_abi_version.argtypes = []
        got = L.psx_abi_version()
        if got != ABI_VERSION:
            raise RuntimeError(
                f"{LIB_PATH} has C ABI version {got}, these bindings need {ABI_VERSION}; "
                "rebuild it with `python -m psierra_amd.build`"
            )
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    msg = lib().psx_last_error()
    return msg.decode() if msg else ""


def check(rc: int) -> None:
    if rc == PSX_OK:
        return
    msg = last_error()
    if rc == PSX_EVOLUME:
        raise ValueError(msg)
    if rc == PSX_EINVAL:
        raise ValueError(msg)
    raise PsxError(rc, msg)


def ptr(x) -> c_void_p:
    """Raw pointer of a numpy array or torch tensor (or None)."""
    if x is None:
        return c_void_p(0)
    if hasattr(x, "data_ptr"):
        return c_void_p(x.data_ptr())
    return c_void_p(x.ctypes.data)
