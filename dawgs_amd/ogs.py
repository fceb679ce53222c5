"""Python mirror of libp::ogs at one rank, over the C-ABI of libogs_hip.so.

Mirrors the reference's include/ogs.hpp:216-344 (names, argument order and
meaning), with device buffers given as torch tensors on the GPU (torch is
only the allocator and stream provider here) and ids as a numpy int64 array.

    o = ogs_t()
    o.Setup(N, ids, Unsigned, Auto, unique=False, verbose=False)
    o.GatherScatter(q, 1, Add, Sym)          # q: torch tensor on cuda
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib

# enums: include/ogs.hpp:179-200, plus the precision-form storage types:
# BF16 (bfloat16 storage) and FloatC (float storage), both reducing Add in
# float with two-sum compensation (ogs_type_t in include/ogs_hip.h)
Float, Double, Int32, Int64, BF16, FloatC = 0, 1, 2, 3, 4, 5
Add, Mul, Max, Min = 0, 1, 2, 3
Sym, NoTrans, Trans = 0, 1, 2
Auto, Pairwise, CrystalRouter, AllToAll = 0, 1, 2, 3
Unsigned, Signed, Halo = 0, 1, 2
# field counts a plan prepares for (ogs_hip_plan_create_k): every K, or K = 1 only
PlanAllK, PlanK1 = 0, 1


_TYPES = None  # torch dtype -> type code, built on first use (torch imported lazily)


def _types() -> dict:
    global _TYPES
    if _TYPES is None:
        import torch
        _TYPES = {torch.float32: Float, torch.float64: Double, torch.int32: Int32, torch.int64: Int64,
                  torch.bfloat16: BF16}
    return _TYPES


def type_of(t, type=None) -> int:
    """Type code of a buffer: the one its dtype names, or `type` when given
    (FloatC selects the compensated Add on a float32 buffer)."""
    code = _types().get(t.dtype)
    if code is None:
        raise TypeError(f"unsupported dtype {t.dtype}")
    if type is None or type == code:
        return code
    if not (type == FloatC and code == Float):
        raise TypeError(f"type {type} does not store {t.dtype}")
    return type


def _need(t, n: int, name: str) -> None:
    """The kernels index a buffer up to the sizes the operator implies; a
    shorter one is refused before anything is launched."""
    if t.numel() < n:
        raise ValueError(f"{name} holds {t.numel()} values, the operator needs {n}")


def _dev_ptr(t) -> int:
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("device buffer expected (torch tensor on a HIP device)")
    if not t.is_contiguous():
        raise ValueError("contiguous buffer expected")
    return t.data_ptr()


_RAW_STREAM = None  # (current-device getter, raw current-stream getter), found on first use


def _stream(stream) -> int:
    """The given stream, or torch's current stream on the current device."""
    global _RAW_STREAM
    if stream is not None:
        return stream
    if _RAW_STREAM is None:
        import torch
        C = torch._C
        if hasattr(C, "_cuda_getCurrentRawStream") and hasattr(C, "_cuda_getDevice"):
            _RAW_STREAM = (C._cuda_getDevice, C._cuda_getCurrentRawStream)
        else:  # public API (slower: builds a Stream object per call)
            _RAW_STREAM = (torch.cuda.current_device, lambda d: torch.cuda.current_stream(d).cuda_stream)
    dev, raw = _RAW_STREAM
    return raw(dev())


class Plan:
    """ogs_hip_plan_t over a caller-owned device CSR pair (operator layer)."""

    def __init__(self, Ncols, NrowsN, NrowsT, rowStartsN, colIdsN, rowStartsT, colIdsT, stream=None,
                 fields: int = PlanAllK):
        self._keep = (rowStartsN, colIdsN, rowStartsT, colIdsT)
        h = C.c_void_p()
        check(lib().ogs_hip_plan_create_k(C.byref(h), Ncols, NrowsN, NrowsT,
                                          _dev_ptr(rowStartsN), _dev_ptr(colIdsN) if colIdsN.numel() else None,
                                          _dev_ptr(rowStartsT), _dev_ptr(colIdsT) if colIdsT.numel() else None,
                                          fields, _stream(stream)), "ogs_hip_plan_create_k")
        self.h = h
        self.owned = True
        self._sizes()

    @classmethod
    def borrowed(cls, handle):
        """A plan owned elsewhere (an ogs_t's), held by one more reference
        (ogs_hip_plan_retain) for this object's lifetime: it stays usable, and
        keeps its CSR alive, after its owner is freed."""
        p = cls.__new__(cls)
        p.h = C.c_void_p(handle)
        p._keep = None
        p.owned = False
        check(lib().ogs_hip_plan_retain(p.h), "ogs_hip_plan_retain")
        p.owned = True
        p._sizes()
        return p

    def _sizes(self) -> None:
        i = self.info()
        self.Ncols, self.NrowsN, self.NrowsT = i["Ncols"], i["NrowsN"], i["NrowsT"]

    def info(self) -> dict:
        i = _lib.PlanInfo()
        check(lib().ogs_hip_plan_info(self.h, C.byref(i)), "ogs_hip_plan_info")
        return {n: getattr(i, n) for n, _ in i._fields_}

    def GatherScatter(self, v, k, op, trans, stream=None, type=None):
        _need(v, self.Ncols * k, "v")
        check(lib().ogs_hip_gather_scatter(self.h, type_of(v, type), op, trans, k, _dev_ptr(v), _stream(stream)),
              "ogs_hip_gather_scatter")

    def Gather(self, gv, v, k, op, trans, stream=None, type=None):
        _need(v, self.Ncols * k, "v")
        _need(gv, (self.NrowsN if trans == NoTrans else self.NrowsT) * k, "gv")
        check(lib().ogs_hip_gather(self.h, type_of(v, type), op, trans, k, _dev_ptr(gv), _dev_ptr(v), _stream(stream)),
              "ogs_hip_gather")

    def Scatter(self, v, gv, k, trans, stream=None, type=None):
        _need(v, self.Ncols * k, "v")
        _need(gv, (self.NrowsN if trans == Trans else self.NrowsT) * k, "gv")
        check(lib().ogs_hip_scatter(self.h, type_of(v, type), trans, k, _dev_ptr(v), _dev_ptr(gv), _stream(stream)),
              "ogs_hip_scatter")

    def __del__(self):
        try:
            if self.owned and self.h:
                lib().ogs_hip_plan_destroy(self.h)
        except Exception:
            pass


class Batch:
    """ogs_hip_batch_t: GatherScatter Sym of many (plan, K, buffer) triples
    in one persistent launch captured in a HIP graph.  Buffers are bound at
    creation; tune=True times launch shapes on scratch copies first.  The
    batch holds a reference to each plan (and through it the plan's CSR), so
    the plans' owners may be freed first; it keeps the buffers too."""

    def __init__(self, items, op: int = Add, tune: bool = True, stream=None, type=None):
        items = list(items)
        if not items:
            raise ValueError("empty batch")
        ty = type_of(items[0][2], type)
        descs = (_lib.BatchDesc * len(items))()
        self._keep = []
        for i, (plan, K, v) in enumerate(items):
            if type_of(v, type) != ty:
                raise TypeError("all buffers of a batch share one dtype")
            _need(v, plan.Ncols * K, "buffer")
            descs[i].plan = plan.h
            descs[i].K = K
            descs[i].d_v = _dev_ptr(v)
            self._keep.append((plan, v))
        h = C.c_void_p()
        check(lib().ogs_hip_batch_create(C.byref(h), ty, op, descs, len(items), int(tune), _stream(stream)),
              "ogs_hip_batch_create")
        self.h = h

    def run(self, stream=None) -> None:
        check(lib().ogs_hip_batch_run(self.h, _stream(stream)), "ogs_hip_batch_run")

    def info(self) -> dict:
        i = _lib.BatchInfo()
        check(lib().ogs_hip_batch_info(self.h, C.byref(i)), "ogs_hip_batch_info")
        return {n: getattr(i, n) for n, _ in i._fields_}

    def __del__(self):
        try:
            if self.h:
                lib().ogs_hip_batch_destroy(self.h)
        except Exception:
            pass


def extract(gq, q, ids, k, stream=None):
    """OKL extract (ogsKernels.okl:167-177): gq[k + n*K] = q[k + ids[n]*K]."""
    check(lib().ogs_hip_extract(type_of(q), ids.numel(), k, _dev_ptr(ids), _dev_ptr(q), _dev_ptr(gq),
                                _stream(stream)), "ogs_hip_extract")


class ogs_t:
    """ogs::ogs_t (include/ogs.hpp:216-344) at one rank."""

    def __init__(self):
        self.h = None

    def Setup(self, N: int, ids: np.ndarray, kind: int = Unsigned, method: int = Auto,
              unique: bool = False, verbose: bool = False, fields: int = PlanAllK) -> None:
        """ogs_t::Setup (ogsSetup.cpp:44-190).  `ids` (int64) is rewritten
        in place with the chosen signs when unique=True (ogsSetup.cpp:139-140).
        `fields` (no counterpart in the reference): PlanK1 builds the operator
        plan for K = 1 only (other K still run, more slowly)."""
        if not (isinstance(ids, np.ndarray) and ids.dtype == np.int64 and ids.flags.c_contiguous):
            raise TypeError("ids must be a contiguous numpy int64 array")
        if N > len(ids):
            raise ValueError("N exceeds len(ids)")
        self.Free()
        h = C.c_void_p()
        check(lib().ogs_setup_k(C.byref(h), N, ids.ctypes.data_as(C.c_void_p), kind, method,
                                int(unique), int(verbose), fields), "ogs_setup_k")
        self.h = h
        i = _lib.OgsInfo()
        check(lib().ogs_get_info(self.h, C.byref(i)), "ogs_get_info")
        for n, _ in i._fields_:
            setattr(self, n, getattr(i, n))
        self.gather_defined = bool(self.gather_defined)

    def Free(self) -> None:
        if self.h:
            lib().ogs_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.Free()
        except Exception:
            pass

    @property
    def plan(self) -> Plan:
        return Plan.borrowed(lib().ogs_get_plan(self.h))

    def csr(self) -> dict:
        """Host copies of the local operator's CSR and the 512-nnz row blocks."""
        a = {
            "rowStartsN": np.empty(self.NrowsT + 1, np.int32), "colIdsN": np.empty(self.nnzN, np.int32),
            "rowStartsT": np.empty(self.NrowsT + 1, np.int32), "colIdsT": np.empty(self.nnzT, np.int32),
            "blockRowStartsN": np.empty(self.NrowBlocksN + 1, np.int32),
            "blockRowStartsT": np.empty(self.NrowBlocksT + 1, np.int32),
        }
        p = lambda x: x.ctypes.data_as(C.c_void_p)
        check(lib().ogs_get_csr(self.h, p(a["rowStartsN"]), p(a["colIdsN"]), p(a["rowStartsT"]),
                                p(a["colIdsT"]), p(a["blockRowStartsN"]), p(a["blockRowStartsT"])),
              "ogs_get_csr")
        return a

    def SetupGlobalToLocalMapping(self) -> np.ndarray:
        """ogs_t::SetupGlobalToLocalMapping (ogsSetup.cpp:867-886): each node's
        gathered index (row), -1 for nodes with id 0."""
        out = np.empty(self.N, dtype=np.int32)
        check(lib().ogs_setup_global_to_local(self.h, out.ctypes.data_as(C.c_void_p)), "ogs_setup_global_to_local")
        return out

    def GatherScatter(self, v, k: int, op: int, trans: int, stream=None, type=None) -> None:
        self._ready()
        _need(v, self.N * k, "v")
        check(lib().ogs_gather_scatter(self.h, type_of(v, type), op, trans, k, _dev_ptr(v), _stream(stream)),
              "ogs_gather_scatter")

    def Gather(self, gv, v, k: int, op: int, trans: int, stream=None, type=None) -> None:
        self._ready()
        _need(v, self.N * k, "v")
        _need(gv, (self.NrowsN if trans == NoTrans else self.NrowsT) * k, "gv")
        check(lib().ogs_gather(self.h, type_of(v, type), op, trans, k, _dev_ptr(gv), _dev_ptr(v), _stream(stream)),
              "ogs_gather")

    def Scatter(self, v, gv, k: int, trans: int, stream=None, type=None) -> None:
        self._ready()
        _need(v, self.N * k, "v")
        _need(gv, (self.NrowsN if trans == Trans else self.NrowsT) * k, "gv")
        check(lib().ogs_scatter(self.h, type_of(v, type), trans, k, _dev_ptr(v), _dev_ptr(gv), _stream(stream)),
              "ogs_scatter")

    def _ready(self) -> None:
        if not self.h:
            raise RuntimeError("ogs handle is not set up")


class halo_t:
    """ogs::halo_t (include/ogs.hpp:347-397) at one rank: a Halo-kind setup
    (exactly one positive id per group) whose Exchange / Combine move nothing,
    since no node is shared with another rank."""

    def __init__(self):
        self.h = None
        self._ogs = None

    def _info(self):
        i = _lib.HaloInfo()
        check(lib().halo_get_info(self.h, C.byref(i)), "halo_get_info")
        for n, _ in i._fields_:
            setattr(self, n, getattr(i, n))
        self.gathered_halo = bool(self.gathered_halo)
        self.gather_defined = bool(self.gather_defined)

    def Setup(self, N: int, ids: np.ndarray, method: int = Auto, verbose: bool = False) -> None:
        """halo_t::Setup (ogsSetup.cpp:55-64)."""
        if not (isinstance(ids, np.ndarray) and ids.dtype == np.int64 and ids.flags.c_contiguous):
            raise TypeError("ids must be a contiguous numpy int64 array")
        if N > len(ids):
            raise ValueError("N exceeds len(ids)")
        self.Free()
        h = C.c_void_p()
        check(lib().halo_setup(C.byref(h), N, ids.ctypes.data_as(C.c_void_p), method, int(verbose)), "halo_setup")
        self.h = h
        self._info()

    def SetupFromGather(self, ogs: "ogs_t") -> None:
        """halo_t::SetupFromGather (ogsSetup.cpp:888-914); `ogs` must be
        gather-defined and outlive this halo."""
        self.Free()
        h = C.c_void_p()
        check(lib().halo_setup_from_gather(C.byref(h), ogs.h), "halo_setup_from_gather")
        self.h = h
        self._ogs = ogs
        self._info()

    def Exchange(self, v, k: int, stream=None, type=None) -> None:
        check(lib().halo_exchange(self.h, type_of(v, type), k, _dev_ptr(v), _stream(stream)), "halo_exchange")

    def Combine(self, v, k: int, stream=None, type=None) -> None:
        check(lib().halo_combine(self.h, type_of(v, type), k, _dev_ptr(v), _stream(stream)), "halo_combine")

    def Free(self) -> None:
        if self.h:
            lib().halo_free(self.h)
            self.h = None
        self._ogs = None

    def __del__(self):
        try:
            self.Free()
        except Exception:
            pass
