"""Loads libogs_hip.so and declares the C-ABI of include/ogs_hip.h.

There is no fallback: if the library is missing, importing fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libogs_hip.so")

_P = C.c_void_p
_I32 = C.c_int32
_I64P = C.POINTER(C.c_int64)
_I32P = C.POINTER(C.c_int32)


class PlanInfo(C.Structure):
    _fields_ = [(n, C.c_int32) for n in
                ("NrowsN", "NrowsT", "nnzN", "nnzT", "maxRowT", "maxRowN", "nblocksK1", "widthK1",
                 "packedK1", "plainK1", "Ncols", "tileD", "colsT", "colsN", "deltaE", "deltaFallback",
                 "ownerMap", "nMask", "bundleRy", "bundleRz", "bundleTy", "bundleTz", "bundleTileD",
                 "bundleFallback", "rowMapE", "rowMapFallback", "byteItems", "byteFallback", "byteK3")] + \
               [("bytes", C.c_int64)]


class BatchDesc(C.Structure):
    _fields_ = [("plan", C.c_void_p), ("K", C.c_int32), ("d_v", C.c_void_p)]


class BatchInfo(C.Structure):
    _fields_ = [("items", C.c_int32), ("nodes", C.c_int32), ("narrow", C.c_int32), ("order", C.c_int32), ("fused", C.c_int32),
                ("candidates", C.c_int32), ("tune_ms", C.c_float), ("best_us", C.c_float), ("plans", C.c_int32)]


class OgsInfo(C.Structure):
    _fields_ = [("N", C.c_int32), ("Ngather", C.c_int32), ("NlocalT", C.c_int32),
                ("NhaloT", C.c_int32), ("NlocalP", C.c_int32), ("NhaloP", C.c_int32),
                ("NgatherGlobal", C.c_int64),
                ("kind", C.c_int32), ("unique", C.c_int32), ("gather_defined", C.c_int32),
                ("NrowsN", C.c_int32), ("NrowsT", C.c_int32), ("nnzN", C.c_int32), ("nnzT", C.c_int32),
                ("NrowBlocksN", C.c_int32), ("NrowBlocksT", C.c_int32)]


class HaloInfo(C.Structure):
    _fields_ = [("N", C.c_int32), ("Ngather", C.c_int32), ("NlocalT", C.c_int32), ("NhaloT", C.c_int32),
                ("NlocalP", C.c_int32), ("NhaloP", C.c_int32), ("Nhalo", C.c_int32),
                ("NgatherGlobal", C.c_int64), ("gathered_halo", C.c_int32), ("gather_defined", C.c_int32)]


# name -> (restype, argtypes); mirrors include/ogs_hip.h exactly
SIGNATURES = {
    "ogs_hip_error_string": (C.c_char_p, [C.c_int]),
    "ogs_hip_plan_create": (C.c_int, [C.POINTER(_P), _I32, _I32, _I32, _P, _P, _P, _P, _P]),
    "ogs_hip_plan_create_k": (C.c_int, [C.POINTER(_P), _I32, _I32, _I32, _P, _P, _P, _P, C.c_int, _P]),
    "ogs_hip_plan_retain": (C.c_int, [_P]),
    "ogs_hip_plan_destroy": (C.c_int, [_P]),
    "ogs_hip_gather_scatter": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _I32, _P, _P]),
    "ogs_hip_gather": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _I32, _P, _P, _P]),
    "ogs_hip_scatter": (C.c_int, [_P, C.c_int, C.c_int, _I32, _P, _P, _P]),
    "ogs_hip_extract": (C.c_int, [C.c_int, _I32, _I32, _P, _P, _P, _P]),
    "ogs_hip_plan_info": (C.c_int, [_P, C.POINTER(PlanInfo)]),
    "ogs_hip_batch_create": (C.c_int, [C.POINTER(_P), C.c_int, C.c_int, C.POINTER(BatchDesc), _I32, C.c_int, _P]),
    "ogs_hip_batch_run": (C.c_int, [_P, _P]),
    "ogs_hip_batch_destroy": (C.c_int, [_P]),
    "ogs_hip_batch_info": (C.c_int, [_P, C.POINTER(BatchInfo)]),
    "ogs_setup": (C.c_int, [C.POINTER(_P), _I32, _P, C.c_int, C.c_int, C.c_int, C.c_int]),
    "ogs_setup_k": (C.c_int, [C.POINTER(_P), _I32, _P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "ogs_free": (C.c_int, [_P]),
    "ogs_get_info": (C.c_int, [_P, C.POINTER(OgsInfo)]),
    "ogs_get_csr": (C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "ogs_get_plan": (_P, [_P]),
    "ogs_setup_global_to_local": (C.c_int, [_P, _P]),
    "ogs_setup_global_to_local_noalloc": (C.c_int, [_P, _P]),
    "ogs_gather_scatter": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _I32, _P, _P]),
    "ogs_gather": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _I32, _P, _P, _P]),
    "ogs_scatter": (C.c_int, [_P, C.c_int, C.c_int, _I32, _P, _P, _P]),
    "halo_setup": (C.c_int, [C.POINTER(_P), _I32, _P, C.c_int, C.c_int]),
    "halo_setup_from_gather": (C.c_int, [C.POINTER(_P), _P]),
    "halo_free": (C.c_int, [_P]),
    "halo_get_info": (C.c_int, [_P, C.POINTER(HaloInfo)]),
    "halo_exchange": (C.c_int, [_P, C.c_int, _I32, _P, _P]),
    "halo_combine": (C.c_int, [_P, C.c_int, _I32, _P, _P]),
}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m dawgs_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class OgsError(RuntimeError):
    def __init__(self, code: int, where: str = ""):
        msg = lib().ogs_hip_error_string(code).decode()
        super().__init__(f"{where}: {msg} (code {code})" if where else f"{msg} (code {code})")
        self.code = code


def check(rc: int, where: str = "") -> None:
    if rc != 0:
        raise OgsError(rc, where)
