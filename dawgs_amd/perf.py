"""PerformanceTest of the reference (dawgsMain.cpp:146-311) and its degree
sweep (dawgsMain.cpp:507-558) on one MI355X, device side.

For one box it sets up the Unsigned `ogs` and the Signed-unique `sogs` from
the structured numbering, then times, with HIP events on the launch stream:
  Device GatherScatter  ogs.GatherScatter(q, K, Add, Sym)
  Device Gather         sogs.Gather(gq, q, K, Add, Trans)
  Device Scatter        sogs.Scatter(q, gq, K, NoTrans)
and prints the reference's line format (dawgsMain.cpp:221-229) with the
achieved bandwidth appended.  DOFS/s = NgatherGlobal / time per call.
"""
from __future__ import annotations

import time

import numpy as np

from . import ogs as _ogs
from .mesh import make_ids

HBM_PEAK_GBS = 8000.0

# ~56 M local DOFs per degree (48^3 elements at p=7), nx = ny = nz
SWEEP_NX = {1: 192, 2: 128, 3: 96, 4: 77, 5: 64, 6: 55, 7: 48}


def gs_bytes(nnz, nrows, K, s):
    """GatherScatter Sym, gather list == scatter list (CSR formulation)."""
    return 2 * K * nnz * s + 4 * nnz + 4 * (nrows + 1)


def gather_bytes(nnz, nrows, K, s):
    """Gather (Trans) or Scatter (NoTrans): each entry and each row value once."""
    return K * nnz * s + K * nrows * s + 4 * nnz + 4 * (nrows + 1)


def time_calls(fn, warmup: int, steps: int) -> float:
    """Seconds per call: `warmup` untimed calls, then `steps` calls between
    HIP events recorded on torch's current stream (where the calls launch)."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(steps):
        fn()
    e1.record(st)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps * 1e-3


def line(kind: str, what: str, NgatherGlobal: int, N: int, p: int, t: float, B: int) -> str:
    gbs = B / t / 1e9
    return (f"{kind} {what + ':':14s} Ranks = 1, Global DOFS = {NgatherGlobal}, Max Local DOFS = {N}, "
            f"Degree = {p}, Time taken = {t:.6g} s, DOFS/s = {NgatherGlobal / t:.6g}, "
            f"DOFS/(s*rank) = {NgatherGlobal / t:.6g}, GB/s = {gbs:.1f}, %HBM = {100 * gbs / HBM_PEAK_GBS:.1f}")


def performance_test(nx: int, ny: int, nz: int, p: int, K: int = 1, warmup: int = 100, steps: int = 100,
                     dtype=None, seed: int = 0, log=print) -> dict:
    """Device half of PerformanceTest for one box; returns the timings and
    the host CSRs (for a CPU leg run beside it)."""
    import torch
    dtype = dtype or torch.float64
    s = torch.empty((), dtype=dtype).element_size()
    t0 = time.perf_counter()
    ids = make_ids(nx, ny, nz, p)
    N = len(ids)
    ogs = _ogs.ogs_t()
    ogs.Setup(N, ids, _ogs.Unsigned, _ogs.Auto, False, False)
    sids = ids.copy()
    sogs = _ogs.ogs_t()
    sogs.Setup(N, sids, _ogs.Signed, _ogs.Auto, True, False)
    del ids, sids
    setup_s = time.perf_counter() - t0
    gen = torch.Generator(device="cuda").manual_seed(seed)
    q = torch.empty(N * K, dtype=torch.float64, device="cuda").uniform_(-1.0, 1.0, generator=gen).to(dtype)
    gq = torch.zeros(sogs.Ngather * K, dtype=dtype, device="cuda")
    info = ogs.plan.info()
    res = {"nx": nx, "ny": ny, "nz": nz, "p": p, "K": K, "N": N, "NgatherGlobal": int(ogs.NgatherGlobal),
           "nnz": ogs.nnzT, "setup_s": setup_s, "tileD": int(info["tileD"]),
           "bundle": {k: int(info["bundle" + k]) for k in ("Ry", "Rz", "Ty", "Tz", "TileD")}}
    if info["bundleTy"]:
        order = (f"bundled blocks of {info['bundleTy']} lines x {info['bundleTz']} layers (row strides "
                 f"{info['bundleRy']}, {info['bundleRz']}), tiled for a layer distance of {info['bundleTileD']} blocks")
    else:
        order = f"block order tiled for a layer distance of {res['tileD']} blocks"
    log(f"setup {setup_s:.2f} s (both ogs); GatherScatter {order} (0: block id order)")
    t = time_calls(lambda: ogs.GatherScatter(q, K, _ogs.Add, _ogs.Sym), warmup, steps)
    B = gs_bytes(ogs.nnzT, ogs.NrowsT, K, s)
    res["gs"] = {"s": t, "bytes": B, "GBps": B / t / 1e9}
    log(line("Device", "GatherScatter", ogs.NgatherGlobal, N, p, t, B))
    t = time_calls(lambda: sogs.Gather(gq, q, K, _ogs.Add, _ogs.Trans), warmup, steps)
    B = gather_bytes(sogs.nnzT, sogs.NrowsT, K, s)
    res["gather"] = {"s": t, "bytes": B, "GBps": B / t / 1e9}
    log(line("Device", "Gather", ogs.NgatherGlobal, N, p, t, B))
    t = time_calls(lambda: sogs.Scatter(q, gq, K, _ogs.NoTrans), warmup, steps)
    res["scatter"] = {"s": t, "bytes": B, "GBps": B / t / 1e9}
    log(line("Device", "Scatter", ogs.NgatherGlobal, N, p, t, B))
    res["_ogs"], res["_sogs"] = ogs, sogs
    return res
