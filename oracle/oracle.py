"""ctypes + numpy wrapper over the CPU oracle (``libogs_oracle.so``) and the
reference's gslib comparator (``_ref/libgs_ref.so``).

TEST INFRASTRUCTURE.  Only the tests, the smoke check and the benchmark's
CPU-baseline leg import this module, and only as the checker.  The product
package (``dawgs_amd``) never imports it.

The C sources restate paranumal/dawgs' host path; see ogs_oracle.c for the
file:line each function follows.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libogs_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libgs_ref.so")

# enum orders: include/ogs.hpp:179-200
FLOAT, DOUBLE, INT32, INT64 = 0, 1, 2, 3
# precision form (the library's own types): bfloat16 storage (numpy uint16
# bit patterns) and float storage with compensated Add (numpy float32)
BF16, FLOAT_C = 4, 5
ADD, MUL, MAX, MIN = 0, 1, 2, 3
SYM, NOTRANS, TRANS = 0, 1, 2
UNSIGNED, SIGNED, HALO = 0, 1, 2

NP_TYPES = {FLOAT: np.float32, DOUBLE: np.float64, INT32: np.int32, INT64: np.int64, BF16: np.uint16}
TYPE_OF = {np.dtype(v): k for k, v in NP_TYPES.items()}
NP_TYPES[FLOAT_C] = np.float32


def _ty(v: np.ndarray, type) -> int:
    """Oracle type code: `type` when given (FLOAT_C for float32 storage with
    compensated Add), else the one the array's dtype names."""
    if type is None:
        return TYPE_OF[v.dtype]
    if np.dtype(NP_TYPES[type]) != v.dtype:
        raise TypeError(f"type {type} stores {np.dtype(NP_TYPES[type])}, got {v.dtype}")
    return type

_P = C.c_void_p
_I32P = C.POINTER(C.c_int32)


class OracleOp(C.Structure):
    """oracle_op_t: ogsOperator_t's host members (ogsOperator.hpp:43-64)."""
    _fields_ = [("Ncols", C.c_int32), ("NrowsN", C.c_int32), ("NrowsT", C.c_int32),
                ("nnzN", C.c_int32), ("nnzT", C.c_int32),
                ("rowStartsN", _I32P), ("rowStartsT", _I32P),
                ("colIdsN", _I32P), ("colIdsT", _I32P),
                ("NrowBlocksN", C.c_int32), ("NrowBlocksT", C.c_int32),
                ("blockRowStartsN", _I32P), ("blockRowStartsT", _I32P),
                ("aliased", C.c_int)]


class OracleOgsStruct(C.Structure):
    """oracle_ogs_t: ogsBase_t counts (ogsBase.hpp:50-62) + two operators."""
    _fields_ = [("N", C.c_int32), ("Ngather", C.c_int32), ("NlocalT", C.c_int32),
                ("NhaloT", C.c_int32), ("NlocalP", C.c_int32), ("NhaloP", C.c_int32),
                ("NgatherGlobal", C.c_int64),
                ("kind", C.c_int), ("unique", C.c_int), ("gather_defined", C.c_int),
                ("local", OracleOp), ("halo", OracleOp)]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.oracle_make_ids.argtypes = [C.c_int] * 4 + [_P]
        L.oracle_setup.argtypes = [C.c_int32, _P, C.c_int, C.c_int, C.POINTER(OracleOgsStruct)]
        L.oracle_free.argtypes = [C.POINTER(OracleOgsStruct)]
        L.oracle_setup_row_blocks.argtypes = [C.POINTER(OracleOp), C.c_int32]
        L.oracle_op_gather.argtypes = [C.POINTER(OracleOp), C.c_int, C.c_int, C.c_int, C.c_int, _P, _P]
        L.oracle_op_scatter.argtypes = [C.POINTER(OracleOp), C.c_int, C.c_int, C.c_int, _P, _P]
        L.oracle_op_gather_scatter.argtypes = [C.POINTER(OracleOp), C.c_int, C.c_int, C.c_int, C.c_int, _P]
        L.oracle_extract.argtypes = [C.c_int, C.c_int32, C.c_int, _P, _P, _P]
        L.oracle_gather_scatter.argtypes = [C.POINTER(OracleOgsStruct), C.c_int, C.c_int, C.c_int, C.c_int, _P]
        L.oracle_gather.argtypes = [C.POINTER(OracleOgsStruct), C.c_int, C.c_int, C.c_int, C.c_int, _P, _P]
        L.oracle_scatter.argtypes = [C.POINTER(OracleOgsStruct), C.c_int, C.c_int, C.c_int, _P, _P]
        L.oracle_reference_q.argtypes = [C.c_uint32, C.c_int32, _P]
        L.oracle_num_threads.restype = C.c_int
        L.oracle_bf16_to_float.argtypes = [C.c_uint16]
        L.oracle_bf16_to_float.restype = C.c_float
        L.oracle_float_to_bf16.argtypes = [C.c_float]
        L.oracle_float_to_bf16.restype = C.c_uint16
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"gslib comparator not built: {REF_PATH}")
        L = C.CDLL(REF_PATH)
        L.gsref_setup.argtypes = [C.c_int, _P, C.c_int]
        L.gsref_setup.restype = _P
        L.gsref_gs_vec.argtypes = [_P, C.c_int, C.c_int, C.c_int, C.c_int, _P]
        L.gsref_free.argtypes = [_P]
        _ref = L
    return _ref


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_P)


def make_ids(nx: int, ny: int, nz: int, p: int) -> np.ndarray:
    """MakeIds at one rank (dawgsMain.cpp:49-143)."""
    ids = np.empty(nx * ny * nz * (p + 1) ** 3, dtype=np.int64)
    lib().oracle_make_ids(nx, ny, nz, p, _ptr(ids))
    return ids


def reference_q(seed: int, n: int) -> np.ndarray:
    """The reference's own input (dawgsMain.cpp:326-330)."""
    q = np.empty(n, dtype=np.float64)
    lib().oracle_reference_q(seed, n, _ptr(q))
    return q


def num_threads() -> int:
    return lib().oracle_num_threads()


class OracleError(RuntimeError):
    def __init__(self, code):
        super().__init__(f"oracle error {code}")
        self.code = code


def _check(rc):
    if rc != 0:
        raise OracleError(rc)


def _arr(p, n):
    if n <= 0:
        return np.zeros(0, dtype=np.int32)
    return np.ctypeslib.as_array(p, shape=(n,)).copy()


class OracleOgs:
    """Single-rank ogs_t restated on the CPU (ogsSetup.cpp:69-190)."""

    def __init__(self, ids: np.ndarray, kind: int = UNSIGNED, unique: bool = False):
        self.ids = np.ascontiguousarray(ids, dtype=np.int64).copy()
        self.s = OracleOgsStruct()
        _check(lib().oracle_setup(len(self.ids), _ptr(self.ids), kind, int(unique), C.byref(self.s)))
        s = self.s
        self.N, self.Ngather, self.NgatherGlobal = s.N, s.Ngather, s.NgatherGlobal
        self.NlocalP, self.NlocalT = s.NlocalP, s.NlocalT
        self.NhaloP, self.NhaloT = s.NhaloP, s.NhaloT
        self.gather_defined = bool(s.gather_defined)
        lo = s.local
        self.NrowsN, self.NrowsT = lo.NrowsN, lo.NrowsT
        self.rowStartsT = _arr(lo.rowStartsT, lo.NrowsT + 1)
        self.rowStartsN = _arr(lo.rowStartsN, lo.NrowsT + 1)
        self.colIdsT = _arr(lo.colIdsT, lo.nnzT)
        self.colIdsN = _arr(lo.colIdsN, lo.nnzN)
        self.NrowBlocksT, self.NrowBlocksN = lo.NrowBlocksT, lo.NrowBlocksN
        self.blockRowStartsT = _arr(lo.blockRowStartsT, lo.NrowBlocksT + 1) if kind != HALO else None
        self.blockRowStartsN = _arr(lo.blockRowStartsN, lo.NrowBlocksN + 1) if kind != HALO else None

    def __del__(self):
        try:
            lib().oracle_free(C.byref(self.s))
        except Exception:
            pass

    def gather_scatter(self, v: np.ndarray, K: int, op: int, trans: int, type=None) -> np.ndarray:
        v = np.ascontiguousarray(v)
        _check(lib().oracle_gather_scatter(C.byref(self.s), _ty(v, type), op, trans, K, _ptr(v)))
        return v

    def gather(self, v: np.ndarray, K: int, op: int, trans: int, type=None) -> np.ndarray:
        v = np.ascontiguousarray(v)
        nrows = self.NrowsN if trans == NOTRANS else self.NrowsT
        gv = np.zeros(max(self.Ngather, self.NlocalT) * K, dtype=v.dtype)
        _check(lib().oracle_gather(C.byref(self.s), _ty(v, type), op, trans, K, _ptr(v), _ptr(gv)))
        return gv[: nrows * K] if trans == NOTRANS else gv[: self.NlocalT * K]

    def scatter(self, v: np.ndarray, gv: np.ndarray, K: int, trans: int, type=None) -> np.ndarray:
        v = np.ascontiguousarray(v)
        gv = np.ascontiguousarray(gv, dtype=v.dtype)
        _check(lib().oracle_scatter(C.byref(self.s), _ty(v, type), trans, K, _ptr(gv), _ptr(v)))
        return v

    # operator-level access (ogsOperator_t host templates), no gather_defined check
    def op_gather(self, v, K, op, trans, type=None):
        v = np.ascontiguousarray(v)
        nrows = self.NrowsN if trans == NOTRANS else self.NrowsT
        gv = np.zeros(max(nrows, 1) * K, dtype=v.dtype)
        _check(lib().oracle_op_gather(C.byref(self.s.local), _ty(v, type), op, trans, K, _ptr(v), _ptr(gv)))
        return gv[: nrows * K]

    def op_scatter(self, v, gv, K, trans, type=None):
        v = np.ascontiguousarray(v)
        gv = np.ascontiguousarray(gv, dtype=v.dtype)
        _check(lib().oracle_op_scatter(C.byref(self.s.local), _ty(v, type), trans, K, _ptr(gv), _ptr(v)))
        return v

    def op_gather_scatter(self, v, K, op, trans, type=None):
        v = np.ascontiguousarray(v)
        _check(lib().oracle_op_gather_scatter(C.byref(self.s.local), _ty(v, type), op, trans, K, _ptr(v)))
        return v


class OracleCsr:
    """ogsOperator_t host templates over caller-given CSR arrays (hand-built
    operators, or another setup's CSR)."""

    def __init__(self, rowStartsT, colIdsT, rowStartsN=None, colIdsN=None, NrowsN=None, Ncols=0):
        self.rsT = np.ascontiguousarray(rowStartsT, dtype=np.int32)
        self.ciT = np.ascontiguousarray(colIdsT, dtype=np.int32)
        self.rsN = self.rsT if rowStartsN is None else np.ascontiguousarray(rowStartsN, dtype=np.int32)
        self.ciN = self.ciT if colIdsN is None else np.ascontiguousarray(colIdsN, dtype=np.int32)
        op = OracleOp()
        op.Ncols = Ncols
        op.NrowsT = len(self.rsT) - 1
        op.NrowsN = op.NrowsT if NrowsN is None else NrowsN
        op.nnzT, op.nnzN = len(self.ciT), len(self.ciN)
        op.rowStartsT = self.rsT.ctypes.data_as(_I32P)
        op.colIdsT = self.ciT.ctypes.data_as(_I32P)
        op.rowStartsN = self.rsN.ctypes.data_as(_I32P)
        op.colIdsN = self.ciN.ctypes.data_as(_I32P)
        self.op = op

    def gather_scatter(self, v, K, op, trans, type=None):
        v = np.ascontiguousarray(v)
        _check(lib().oracle_op_gather_scatter(C.byref(self.op), _ty(v, type), op, trans, K, _ptr(v)))
        return v

    def gather(self, v, K, op, trans, type=None):
        v = np.ascontiguousarray(v)
        nrows = self.op.NrowsN if trans == NOTRANS else self.op.NrowsT
        gv = np.zeros(max(nrows, 1) * K, dtype=v.dtype)
        _check(lib().oracle_op_gather(C.byref(self.op), _ty(v, type), op, trans, K, _ptr(v), _ptr(gv)))
        return gv[: nrows * K]

    def scatter(self, v, gv, K, trans, type=None):
        v = np.ascontiguousarray(v)
        gv = np.ascontiguousarray(gv, dtype=v.dtype)
        _check(lib().oracle_op_scatter(C.byref(self.op), _ty(v, type), trans, K, _ptr(gv), _ptr(v)))
        return v


def csr_gather_scatter(rowStarts, colIds, v, K, op):
    """Sym GatherScatter over an arbitrary CSR."""
    return OracleCsr(rowStarts, colIds).gather_scatter(v, K, op, SYM)


def extract(ids: np.ndarray, q: np.ndarray, K: int, type=None) -> np.ndarray:
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    q = np.ascontiguousarray(q)
    gq = np.empty(len(ids) * K, dtype=q.dtype)
    _check(lib().oracle_extract(_ty(q, type), len(ids), K, _ptr(ids), _ptr(q), _ptr(gq)))
    return gq


def bf16_to_float(b: np.ndarray) -> np.ndarray:
    """bfloat16 bit patterns (uint16) -> float32, exact."""
    return (np.asarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def float_to_bf16(f: np.ndarray) -> np.ndarray:
    """float32 -> bfloat16 bit patterns, round to nearest even, NaN quieted
    (numpy restatement of oracle_float_to_bf16)."""
    u = np.ascontiguousarray(f, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r[nan] = ((u[nan] >> 16) | 0x40).astype(np.uint16)
    return r


class GsRef:
    """The reference's legacy comparator: gslib gs_setup / gs_vec (dawgsGS.cpp:34-71)."""

    DOM = {np.dtype(np.float64): 0, np.dtype(np.float32): 1, np.dtype(np.int32): 2, np.dtype(np.int64): 3}
    # gslib op enum: add, mul, min, max (gs_defs.h:26-31)
    OP = {ADD: 0, MUL: 1, MIN: 2, MAX: 3}

    def __init__(self, ids: np.ndarray, unique: bool = False):
        self.ids = np.ascontiguousarray(ids, dtype=np.int64)
        self.h = ref().gsref_setup(len(self.ids), _ptr(self.ids), int(unique))

    def gs_vec(self, u: np.ndarray, K: int, op: int = ADD, transpose: int = 0) -> np.ndarray:
        u = np.ascontiguousarray(u).copy()
        ref().gsref_gs_vec(_ptr(u), K, self.DOM[u.dtype], self.OP[op], transpose, self.h)
        return u

    def __del__(self):
        try:
            ref().gsref_free(self.h)
        except Exception:
            pass
