"""Loader for the HIP library nmc_amd/libnmc_hip.so (built by nmc_amd/build.py).

There is no fallback: if the library is missing or cannot be loaded, every
call raises.  The library is looked up in-tree next to this file.
"""
from __future__ import annotations

import ctypes as C
import os

from . import params as P

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libnmc_hip.so")

_lib = None


class NmcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"nmc error {code}: {msg}")
        self.code = code


def lib():
    """The loaded C library (raises NmcError if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NmcError(P.ERR_STATE, f"{LIB_PATH} not built; run nmc_amd/build.py "
                       "(or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    H = C.c_void_p
    rp, tp, ext = C.POINTER(P.RunParams), C.POINTER(P.TissueParams), C.POINTER(P.ExtConfig)
    u8p = C.POINTER(C.c_uint8)
    sigs = {
        "nmc_ext_config_init": (None, [ext]),
        "nmc_validate": (C.c_int, [C.c_int, rp, tp, u8p, ext]),
        "nmc_create": (C.c_int, [C.c_int, C.c_int, C.POINTER(H)]),
        "nmc_prepare": (C.c_int, [H, rp, tp, u8p, ext]),
        "nmc_set_seed": (C.c_int, [H, C.c_uint64, C.c_uint64]),
        "nmc_run": (C.c_int, [H, C.c_uint64]),
        "nmc_run_async": (C.c_int, [H, C.c_uint64]),
        "nmc_synchronize": (C.c_int, [H]),
        "nmc_last_kernel_ms": (C.c_int, [H, C.POINTER(C.c_float)]),
        "nmc_get_results": (C.c_int, [H, C.POINTER(P.Results)]),
        "nmc_add_results": (C.c_int, [H, C.POINTER(P.Results)]),
        "nmc_reset_tallies": (C.c_int, [H]),
        "nmc_destroy": (None, [H]),
        "nmc_last_error": (C.c_char_p, []),
        "nmc_version": (C.c_char_p, []),
        "nmc_set_tuning": (C.c_int, [H, C.POINTER(P.Tuning)]),
        "nmc_get_tuning": (C.c_int, [H, C.POINTER(P.Tuning)]),
        "nmc_last_launch_count": (C.c_int, [H, C.POINTER(C.c_uint32)]),
        "nmc_last_transport_ms": (C.c_int, [H, C.POINTER(C.c_float)]),
        "nmc_mcxyz_sens": (C.c_int, [H, C.POINTER(C.c_int32), C.c_int]),
        "nmc_mcxyz_sens_results": (C.c_int, [H, C.c_void_p, C.c_void_p]),
        "nmc_prepare_batch": (C.c_int, [H, rp, ext, C.POINTER(P.WorkItem), C.c_int]),
        "nmc_run_batch": (C.c_int, [H]),
        "nmc_run_batch_async": (C.c_int, [H]),
        "nmc_get_batch_results": (C.c_int, [H, C.POINTER(C.c_double), C.POINTER(P.Totals)]),
        "nmc_set_graph_mode": (C.c_int, [H, C.c_int]),
        "nmc_layered_prepare": (C.c_int, [H, C.POINTER(P.LayeredConfig), C.POINTER(P.Stack), C.c_int]),
        "nmc_layered_update": (C.c_int, [H, C.POINTER(P.Stack), C.c_int]),
        "nmc_layered_run": (C.c_int, [H]),
        "nmc_layered_results": (C.c_int, [H, C.c_void_p, C.c_void_p, C.POINTER(P.Totals),
                                          C.POINTER(C.c_uint64)]),
        "nmc_layered_backward": (C.c_int, [H, C.c_void_p, C.c_void_p, C.c_void_p]),
        "nmc_layered_records": (C.c_int, [H, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64),
                                          C.POINTER(C.c_int)]),
        "nmc_layered_absorbed_grad": (C.c_int, [H, C.c_void_p, C.c_void_p]),
        "nmc_mcxyz_run_noalloc": (C.c_int, [C.c_int, rp, tp, ext, C.POINTER(P.Tuning), C.c_void_p,
                                            C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.c_void_p]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


EXPORTED = ("nmc_ext_config_init", "nmc_validate", "nmc_create", "nmc_prepare", "nmc_set_seed",
            "nmc_run", "nmc_run_async", "nmc_synchronize", "nmc_last_kernel_ms",
            "nmc_get_results", "nmc_add_results", "nmc_reset_tallies", "nmc_destroy", "nmc_last_error",
            "nmc_version", "nmc_set_tuning", "nmc_get_tuning", "nmc_last_launch_count",
            "nmc_last_transport_ms", "nmc_mcxyz_sens", "nmc_mcxyz_sens_results",
            "nmc_prepare_batch",
            "nmc_run_batch", "nmc_run_batch_async", "nmc_get_batch_results",
            "nmc_set_graph_mode", "nmc_layered_prepare", "nmc_layered_update", "nmc_layered_run",
            "nmc_layered_results", "nmc_layered_backward", "nmc_layered_records",
            "nmc_layered_absorbed_grad", "nmc_mcxyz_run_noalloc")


def check(status: int) -> None:
    if status != P.OK:
        msg = lib().nmc_last_error()
        raise NmcError(status, msg.decode() if msg else "")
