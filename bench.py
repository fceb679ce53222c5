#!/usr/bin/env python3
"""Benchmark of the rank-local gather-scatter on one MI355X.

Default workload (the headline configuration): ogs_t::GatherScatter, fp64,
Add, Sym, K=1 on the 48^3-element hex box at p=7 (56,623,104 local DOFs,
38,272,753 global DOFs), ids from the structured numbering, values
uniform[-1,1] from a fixed seed, all resident in HBM before timing.

One step = one GatherScatter call (one kernel launch).  W untimed warm-ups,
then K timed steps between synchronisations, timed with HIP events on the
stream the kernel is launched on.  DOFS/s follows the reference's
definition: NgatherGlobal / time per call (dawgsMain.cpp:207).

Prints ONE JSON line on stdout; diagnostics go to stderr.  After the timed
region the default line also carries summaries of the metric's other half,
"batched/fp32 vs plain": `precision` (FP32-compensated and bf16 storage, K=3
on the same mesh, against the plain fp64 method, with errors against the fp64
oracle) and `batched` (48 meshes in one graph launch against one launch per
mesh and field); --no-extras skips them.

    python bench.py [--gpus 1] [--steps 100] [--warmup 100] [--nx 48 [--ny NY] [--nz NZ]] [--p 7] [--no-cpu]
                    [--copy-probe] [--idle-probe]
    python bench.py --sweep [DEGREES] [--k K] | --batched | --precision | --readme
    python bench.py --cc [--nx 10 --ny 10 --nz 10 --p 3]   (the reference's -cc correctness check)
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GatherScatter DOFS/s + % HBM peak, fp64 48^3 p=7; batched/fp32 vs plain"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "gs_f64_add_sym_48cubed_p7_traffic.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu() -> str:
    """lscpu's model, its sockets x cores per socket (the thread count the
    reference picks when OMP_NUM_THREADS is unset, platformDeviceConfig.cpp:
    105-146), the CPUs this process may run on, and OMP_NUM_THREADS."""
    fields = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                fields[k.strip()] = v.strip()
    except Exception:
        pass
    model = fields.get("Model name", "unknown")
    try:
        rule = int(fields["Socket(s)"]) * int(fields["Core(s) per socket"])
        topo = f"{fields['Socket(s)']} socket(s) x {fields['Core(s) per socket']} cores"
    except (KeyError, ValueError):
        rule, topo = None, "topology unknown"
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count()
    omp = os.environ.get("OMP_NUM_THREADS")
    return (f"{model} ({topo} = {rule} by the reference's rule; {avail} CPUs available to this process; "
            f"OMP_NUM_THREADS={omp if omp else 'unset'})")


def algorithmic_bytes(nnz: int, nrows: int, K: int, s: int) -> int:
    """GatherScatter, Sym (gather list == scatter list): each value read and
    written once, each colId and row start read once."""
    return 2 * K * nnz * s + 4 * nnz + 4 * (nrows + 1)


def cpu_baseline(csr, N, NgatherGlobal, budget_s=12.0, max_calls=20):
    """The reference's host GatherScatter loop (ogsOperator.cpp:311-360,
    restated in the oracle, OpenMP over rows) on the same CSR and workload."""
    from oracle import oracle as O
    ref = O.OracleCsr(csr["rowStartsT"], csr["colIdsT"], Ncols=N)
    v = np.random.default_rng(1).uniform(-1, 1, N)
    for _ in range(3):
        ref.gather_scatter(v, 1, O.ADD, O.SYM)
    t0 = time.perf_counter()
    calls = 0
    while calls < max_calls and (calls < 2 or time.perf_counter() - t0 < budget_s):
        ref.gather_scatter(v, 1, O.ADD, O.SYM)
        calls += 1
    t = (time.perf_counter() - t0) / calls
    return {
        "value": NgatherGlobal / t,
        "unit": "DOFS/s",
        "cores": O.num_threads(),
        "kind": "port",
        "sample": f"full workload, {calls} host GatherScatter calls after 3 warm-ups, "
                  f"{t * 1e3:.1f} ms/call, OpenMP threads={O.num_threads()}, host CPU: {host_cpu()}",
    }


def host_lines(res, log) -> dict:
    """Host rows of PerformanceTest: the reference's host loops (oracle,
    OpenMP) over the same CSRs as the device rows."""
    from dawgs_amd.perf import line
    from oracle import oracle as O
    ogs, sogs = res["_ogs"], res["_sogs"]
    N, Ng, p, K = res["N"], res["NgatherGlobal"], res["p"], res["K"]
    c = ogs.csr()
    u = O.OracleCsr(c["rowStartsT"], c["colIdsT"], Ncols=N)
    c = sogs.csr()
    sg = O.OracleCsr(c["rowStartsT"], c["colIdsT"], c["rowStartsN"], c["colIdsN"], sogs.NrowsN, N)
    v = np.random.default_rng(2).uniform(-1, 1, N * K)
    out = {}

    def timed(fn, n=10):
        fn()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        return (time.perf_counter() - t0) / n

    t = timed(lambda: u.gather_scatter(v, K, O.ADD, O.SYM))
    log(line("Host  ", "GatherScatter", Ng, N, p, t, res["gs"]["bytes"]))
    out["gs"] = t
    gv = sg.gather(v, K, O.ADD, O.TRANS)
    t = timed(lambda: sg.gather(v, K, O.ADD, O.TRANS))
    log(line("Host  ", "Gather", Ng, N, p, t, res["gather"]["bytes"]))
    out["gather"] = t
    t = timed(lambda: sg.scatter(v, gv, K, O.NOTRANS))
    log(line("Host  ", "Scatter", Ng, N, p, t, res["scatter"]["bytes"]))
    out["scatter"] = t
    return out


def run_sweep(args) -> int:
    """Degree sweep p = 1..7 at ~56 M local DOFs per run."""
    import gc

    from dawgs_amd.perf import SWEEP_NX, performance_test
    from oracle import oracle as O
    rows = []
    for p in [int(x) for x in args.sweep.split(",")]:
        nx = SWEEP_NX[p]
        log(f" ---- degree {p}: {nx}^3 elements ----")
        res = performance_test(nx, nx, nx, p, K=args.k, warmup=args.warmup, steps=args.steps, log=log)
        row = {k: v for k, v in res.items() if not k.startswith("_")}
        if not args.no_cpu:
            row["host_s"] = host_lines(res, log)
        rows.append(row)
        del res
        gc.collect()
    out = {"metric": "ogs degree sweep (Device/Host GatherScatter, Gather, Scatter; fp64 Add; ~56M local DOFs)",
           "fields": args.k,
           "unit": "DOFS/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "dtype": "f64",
           "data": "synthetic", "host_threads": O.num_threads() if not args.no_cpu else None,
           "host_cpu": host_cpu() if not args.no_cpu else None,
           "rows": [{"p": r["p"], "nx": r["nx"], "local_dofs": r["N"], "global_dofs": r["NgatherGlobal"],
                     "gs_tileD": r["tileD"], "gs_bundle": r["bundle"],
                     "gs_DOFSps": r["NgatherGlobal"] / r["gs"]["s"], "gs_GBps": r["gs"]["GBps"],
                     "gs_frac": r["gs"]["GBps"] / HBM_PEAK_GBS,
                     "gather_DOFSps": r["NgatherGlobal"] / r["gather"]["s"], "gather_GBps": r["gather"]["GBps"],
                     "scatter_DOFSps": r["NgatherGlobal"] / r["scatter"]["s"], "scatter_GBps": r["scatter"]["GBps"],
                     "host_gs_DOFSps": (r["NgatherGlobal"] / r["host_s"]["gs"]) if "host_s" in r else None}
                    for r in rows]}
    print(json.dumps(out), flush=True)
    return 0


# The reference README's sample run (-sw: one rank, 34^3 elements, p = 1):
# DOFS/s = NgatherGlobal / time per call, on hardware it does not name
# (README.md:78-83).
README_SAMPLE = {"nx": 34, "p": 1, "global_dofs": 42875, "local_dofs": 314432,
                 "DOFSps": {"device_gs": 2.21005e9, "device_gather": 2.9569e9, "device_scatter": 2.82072e9,
                            "host_gs": 2.38924e7, "host_gather": 3.78253e7, "host_scatter": 3.28872e7}}


def run_readme(args) -> int:
    """The README's sample configuration timed as PerformanceTest times it
    (10 warm-ups, then the mean of 20 calls; dawgsMain.cpp:146-311), Device
    and (unless --no-cpu) Host lines in its format, and one JSON line with
    the README's published figures beside this build's."""
    from dawgs_amd.perf import performance_test
    from oracle import oracle as O
    r = README_SAMPLE
    res = performance_test(r["nx"], r["nx"], r["nx"], r["p"], warmup=10, steps=20, log=log)
    assert res["NgatherGlobal"] == r["global_dofs"] and res["N"] == r["local_dofs"]
    ng = res["NgatherGlobal"]
    got = {"device_gs": ng / res["gs"]["s"], "device_gather": ng / res["gather"]["s"],
           "device_scatter": ng / res["scatter"]["s"]}
    if not args.no_cpu:
        h = host_lines(res, log)
        got.update(host_gs=ng / h["gs"], host_gather=ng / h["gather"], host_scatter=ng / h["scatter"])
    out = {"metric": "README sample run (34^3 elements, p=1, one rank): DOFS/s per operation", "unit": "DOFS/s",
           "n_gpus": 1, "steps": 20, "warmup": 10, "dtype": "f64", "data": "synthetic",
           "config": {"nx": r["nx"], "p": r["p"], "local_dofs": r["local_dofs"], "global_dofs": r["global_dofs"]},
           "this_build": got, "readme_published": r["DOFSps"],
           "ratio": {k: got[k] / r["DOFSps"][k] for k in got},
           "host_threads": O.num_threads() if not args.no_cpu else None,
           "host_cpu": host_cpu() if not args.no_cpu else None}
    print(json.dumps(out), flush=True)
    return 0


def batch_meshes(seed: int = 48, n: int = 48):
    """The batched configuration: n meshes, nx, ny, nz ~ U{4..32},
    p ~ U{1..7}, K ~ {1, 3, 8, 16}, from one seed."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nx, ny, nz = (int(x) for x in rng.integers(4, 33, 3))
        p = int(rng.integers(1, 8))
        K = int(rng.choice([1, 3, 8, 16]))
        out.append((nx, ny, nz, p, K))
    return out


def batched_result(args) -> dict:
    import torch
    import dawgs_amd as D
    from dawgs_amd.mesh import make_ids
    from dawgs_amd.perf import time_calls
    specs = batch_meshes()
    t0 = time.perf_counter()
    ogss, bufs, fields = [], [], []
    gen = torch.Generator(device="cuda").manual_seed(0)
    fdofs = 0  # sum over meshes of NgatherGlobal * K
    for (nx, ny, nz, p, K) in specs:
        ids = make_ids(nx, ny, nz, p)
        g = D.ogs_t()
        g.Setup(len(ids), ids, D.Unsigned, D.Auto, False, False)
        v = torch.empty(len(ids) * K, dtype=torch.float64, device="cuda").uniform_(-1.0, 1.0, generator=gen)
        ogss.append(g)
        bufs.append(v)
        # the plain method's layout: one contiguous array per field
        fields.append([v.view(len(ids), K)[:, k].contiguous() for k in range(K)])
        fdofs += g.NgatherGlobal * K
    nnz = sum(g.nnzT for g in ogss)
    log(f"batch: {len(specs)} meshes, {sum(K for *_, K in specs)} fields, {nnz} local DOFs, "
        f"{fdofs} global field-DOFs; setup {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    batch = D.Batch([(g.plan, K, v) for g, (*_, K), v in zip(ogss, specs, bufs)], D.Add, tune=True)
    info = batch.info()
    log(f"batch create + autotune {time.perf_counter() - t0:.2f} s: {info}")

    def plain():
        for g, fs in zip(ogss, fields):
            for f in fs:
                g.GatherScatter(f, 1, D.Add, D.Sym)

    warm = max(args.warmup // 5, 10)
    steps = max(args.steps // 2, 20)
    t_plain = time_calls(plain, warm, steps)
    t_batch = time_calls(batch.run, warm, steps)
    nlaunch = sum(len(fs) for fs in fields)
    B = sum(2 * K * g.nnzT * 8 + 4 * g.nnzT + 4 * (g.NrowsT + 1) for g, (*_, K) in zip(ogss, specs))
    log(f"plain: {nlaunch} launches, {t_plain * 1e3:.3f} ms/step; batched: 1 graph launch, "
        f"{t_batch * 1e3:.3f} ms/step; speed-up {t_plain / t_batch:.2f}x")
    out = {"metric": "batched GatherScatter field-DOFS/s (48 meshes, K in {1,3,8,16}, fp64 Add), one graph launch",
           "value": fdofs / t_batch, "unit": "DOFS/s", "n_gpus": 1, "steps": steps, "warmup": warm,
           "ms_per_step": t_batch * 1e3, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "48 hex meshes, nx,ny,nz ~ U{4..32}, p ~ U{1..7}, K ~ {1,3,8,16}, seed 48",
                      "meshes": [list(x) for x in specs], "global_field_dofs": int(fdofs), "local_dofs": int(nnz)},
           "plain": {"launches": nlaunch, "ms_per_step": t_plain * 1e3, "value": fdofs / t_plain},
           "speedup_vs_plain": t_plain / t_batch,
           "roofline": {"bound": "hbm", "algorithmic_bytes": B, "achieved": round(B / t_batch / 1e9, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(B / t_batch / 1e9 / HBM_PEAK_GBS, 4)},
           "autotune": info}
    return out


def run_batched(args) -> int:
    print(json.dumps(batched_result(args)), flush=True)
    return 0


def error_stats(res: np.ndarray, y: np.ndarray) -> dict:
    """Error of `res` against the fp64 result `y`: max absolute, normwise
    relative ||res - y||_2 / ||y||_2, and max elementwise relative over
    |y| > 1e-300."""
    d = res.astype(np.float64) - y
    big = np.abs(y) > 1e-300
    return {"max_abs": float(np.max(np.abs(d))),
            "normwise_rel": float(np.linalg.norm(d) / np.linalg.norm(y)),
            "max_elem_rel": float(np.max(np.abs(d[big]) / np.abs(y[big])))}


def precision_result(args) -> dict:
    """Precision form: FLOAT_C (float storage) and BF16 storage, Add reduced
    in float with two-sum compensation, K=3 interleaved fields on the 48^3
    p=7 mesh, values log-uniform in magnitude over [1e-3, 1e3] with random
    sign.  Timed against the plain fp64 method (one K=1 launch per field, the
    reference's layout) and fp64 K=3 in one launch; errors against the fp64
    oracle of the unrounded values."""
    import torch
    import dawgs_amd as D
    from dawgs_amd.mesh import make_ids
    from dawgs_amd.perf import time_calls
    from oracle import oracle as O  # the checker: fp64 results and bit-exact comparisons
    nx, p, K = args.nx, args.p, 3
    t0 = time.perf_counter()
    ids = make_ids(nx, nx, nx, p)
    N = len(ids)
    ogs = D.ogs_t()
    ogs.Setup(N, ids, D.Unsigned, D.Auto, False, False)
    del ids
    gen = torch.Generator(device="cuda").manual_seed(0)
    mag = torch.empty(N * K, dtype=torch.float64, device="cuda").uniform_(-3.0, 3.0, generator=gen)
    sign = torch.randint(0, 2, (N * K,), device="cuda", generator=gen).to(torch.float64) * 2 - 1
    x = torch.pow(10.0, mag) * sign
    del mag, sign
    x32 = x.to(torch.float32)
    xb = x32.to(torch.bfloat16)  # storage rounding: fp64 -> float -> bf16, each to nearest even
    log(f"setup {time.perf_counter() - t0:.1f} s: N={N}, K={K}, nnz={ogs.nnzT}")
    fdofs = ogs.NgatherGlobal * K

    # errors first, on fresh copies: one GatherScatter each
    c = ogs.csr()
    ref = O.OracleCsr(c["rowStartsT"], c["colIdsT"], Ncols=N)
    xh = x.cpu().numpy()
    y = ref.gather_scatter(xh.copy(), K, O.ADD, O.SYM)
    errors = {}
    for name, dev, ty in (("fp32_compensated", x32, D.FloatC), ("bf16", xb, D.BF16)):
        d = dev.clone()
        ogs.GatherScatter(d, K, D.Add, D.Sym, type=ty)
        if ty == D.BF16:
            got = d.view(torch.int16).cpu().numpy().view(np.uint16)
            want = ref.gather_scatter(xb.view(torch.int16).cpu().numpy().view(np.uint16), K, O.ADD, O.SYM,
                                      type=O.BF16)
            res = O.bf16_to_float(got)
        else:
            got = d.cpu().numpy()
            want = ref.gather_scatter(x32.cpu().numpy(), K, O.ADD, O.SYM, type=O.FLOAT_C)
            res = got
        st = error_stats(res, y)
        st["bit_exact_vs_oracle"] = got.tobytes() == want.tobytes()
        errors[name] = st
        log(f"{name}: {st}")
        del d, got, want, res
    # the plain fp32 sum (reference arithmetic, no compensation) beside it
    d = x32.clone()
    ogs.GatherScatter(d, K, D.Add, D.Sym)
    errors["fp32_plain"] = error_stats(d.cpu().numpy(), y)
    log(f"fp32_plain: {errors['fp32_plain']}")
    del d, y, xh, ref, c

    fields = [x.view(N, K)[:, k].contiguous() for k in range(K)]  # plain method: one array per field
    x64 = x.clone()
    f32 = x32.clone()
    bf = xb.clone()
    b32 = D.Batch([(ogs.plan, K, f32)], D.Add, tune=True, type=D.FloatC)
    bb = D.Batch([(ogs.plan, K, bf)], D.Add, tune=True, type=D.BF16)
    b64 = D.Batch([(ogs.plan, K, x64)], D.Add, tune=True)

    def plain64():
        for f in fields:
            ogs.GatherScatter(f, 1, D.Add, D.Sym)

    forms = {
        "fp64_plain_per_field": (plain64, 8, K),
        "fp64_k3": (lambda: ogs.GatherScatter(x64, K, D.Add, D.Sym), 8, 1),
        "fp64_k3_batched": (b64.run, 8, 1),
        "fp32_compensated_k3": (lambda: ogs.GatherScatter(f32, K, D.Add, D.Sym, type=D.FloatC), 4, 1),
        "fp32_compensated_k3_batched": (b32.run, 4, 1),
        "bf16_k3": (lambda: ogs.GatherScatter(bf, K, D.Add, D.Sym), 2, 1),
        "bf16_k3_batched": (bb.run, 2, 1),
    }
    res = {}
    for name, (fn, s, launches) in forms.items():
        t = time_calls(fn, args.warmup, args.steps)
        kk = 1 if name == "fp64_plain_per_field" else K
        B = launches * (2 * kk * ogs.nnzT * s + 4 * ogs.nnzT + 4 * (ogs.NrowsT + 1))
        res[name] = {"us": round(t * 1e6, 1), "DOFSps": fdofs / t, "launches": launches,
                     "algorithmic_bytes": B, "GBps": round(B / t / 1e9, 1),
                     "frac": round(B / t / 1e9 / HBM_PEAK_GBS, 4)}
    base = res["fp64_plain_per_field"]["DOFSps"]
    for name, r in res.items():
        r["speedup_vs_fp64_plain"] = round(r["DOFSps"] / base, 3)
        r["speedup_vs_fp64_k3"] = round(r["DOFSps"] / res["fp64_k3"]["DOFSps"], 3)
        log(f"{name:30s} {r['us']:8.1f} us  {r['DOFSps']:.4g} field-DOFS/s  {r['GBps']:7.1f} GB/s "
            f"({100 * r['frac']:.1f}%)  x{r['speedup_vs_fp64_plain']:.2f} vs plain fp64")
    head = res["fp32_compensated_k3_batched"]
    out = {"metric": "precision form: batched FP32-compensated GatherScatter field-DOFS/s vs plain fp64 "
                     "(48^3 p=7, K=3)",
           "value": head["DOFSps"], "unit": "DOFS/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": head["us"] * 1e-3, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"ogs GatherScatter Add Sym, K={K} interleaved fields on a {nx}^3 hex mesh, "
                                  f"p={p}; values 10^U[-3,3] with random sign (fp64, seed 0), stored as float "
                                  f"/ bf16", "nx": nx, "p": p, "K": K, "local_dofs": N,
                      "global_field_dofs": int(fdofs)},
           "speedup_vs_fp64_plain": head["speedup_vs_fp64_plain"],
           "speedup_vs_fp64_k3": head["speedup_vs_fp64_k3"],
           "forms": res, "errors_vs_fp64": errors,
           "roofline": {"bound": "hbm", "achieved": head["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": head["frac"], "algorithmic_bytes": head["algorithmic_bytes"]}}
    return out


def run_precision(args) -> int:
    out = precision_result(args)
    if args.copy_probe:
        out["copy_probe"] = copy_probe(1285553088)
    print(json.dumps(out), flush=True)
    return 0


def run_cc(args) -> int:
    """CorrectnessTest (dawgsMain.cpp:314-416) on an nx x ny x nz box: the
    reference's own input (mt19937(rank 0), U(-0.25, 0.25), dawgsMain.cpp:
    326-330), its golden result from gslib's gs_vec (the reference's legacy
    comparator, compiled from its sources into oracle/_ref, dawgsGS.cpp:34-71),
    and Error = sum |q - qcheck| (CheckCorrectness, dawgsMain.cpp:29-47) for
    the device path and for the reference's host loops (the oracle)."""
    import torch
    import dawgs_amd as D
    from dawgs_amd.mesh import make_ids
    from oracle import oracle as O  # the checker: gslib and the host loops
    nx, ny, nz, p = args.nx, args.ny, args.nz, args.p
    ids = make_ids(nx, ny, nz, p)
    N = len(ids)
    q = O.reference_q(0, N)
    if not O.ref_available():
        log("gslib comparator not built (make -C oracle where the reference tree is present)")
        return 1
    qcheck = O.GsRef(ids).gs_vec(q, 1)
    errors = {}

    def check(name, qtest):
        bad = np.flatnonzero(np.abs(qtest - qcheck) > 1.0e-6)
        for n in bad[:10]:
            log(f"Rank 0, Entry {n}, baseId {ids[n]} q = {qtest[n]:f}, qRef = {qcheck[n]:f}")
        err = float(np.sum(np.abs(qtest - qcheck)))
        errors[name] = err
        log(f"{name}:  Error = {err:g}")

    ogs = D.ogs_t()
    ogs.Setup(N, ids.copy(), D.Unsigned, D.Auto, False, True)
    sids = ids.copy()
    sogs = D.ogs_t()
    sogs.Setup(N, sids, D.Signed, D.Auto, True, True)
    hogs, hsogs = O.OracleOgs(ids, O.UNSIGNED), O.OracleOgs(ids, O.SIGNED, True)
    for tag, dev, host in (("ogs", ogs, hogs), ("sogs", sogs, hsogs)):
        d = torch.from_numpy(q.copy()).cuda()
        dev.GatherScatter(d, 1, D.Add, D.Sym)
        check(f"Device GatherScatter ({tag})", d.cpu().numpy())
        check(f"Host   GatherScatter ({tag})", host.gather_scatter(q.copy(), 1, O.ADD, O.SYM))
    d = torch.from_numpy(q.copy()).cuda()
    gq = torch.zeros(sogs.Ngather, dtype=torch.float64, device="cuda")
    sogs.Gather(gq, d, 1, D.Add, D.Trans)
    sogs.Scatter(d, gq, 1, D.NoTrans)
    check("Device Gather+Scatter", d.cpu().numpy())
    g = hsogs.gather(q, 1, O.ADD, O.TRANS)
    check("Host   Gather+Scatter", hsogs.scatter(np.zeros(N), g, 1, O.NOTRANS))
    out = {"metric": "correctness check against gslib (-cc): Error = sum |q - qcheck|", "errors": errors,
           "pass": all(e == 0.0 for e in errors.values()), "n_gpus": 1, "dtype": "f64",
           "data": "the reference's own input: mt19937(0), uniform(-0.25, 0.25)",
           "config": {"nx": nx, "ny": ny, "nz": nz, "p": p, "local_dofs": N,
                      "global_dofs": int(ogs.NgatherGlobal)}}
    print(json.dumps(out), flush=True)
    return 0 if out["pass"] else 3


def copy_probe(nbytes: int) -> dict:
    """A plain device copy moving `nbytes` (torch.neg: one elementwise kernel,
    reading nbytes/2 and writing nbytes/2, its name distinct from every other
    kernel of a run): the practical streaming ceiling, and the known-size
    kernel that scales rocprofv3's FETCH_SIZE / WRITE_SIZE."""
    import torch
    n = nbytes // 16
    gen = torch.Generator(device="cuda").manual_seed(1)
    src = torch.empty(n, dtype=torch.float64, device="cuda").uniform_(generator=gen)
    dst = torch.empty_like(src)
    st = torch.cuda.current_stream()
    for _ in range(3):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(10):
        torch.neg(src, out=dst)
    e1.record(st)
    torch.cuda.synchronize()
    cms = e0.elapsed_time(e1) / 10
    log(f"copy probe: {16 * n} B in {cms * 1e3:.1f} us = {16 * n / (cms * 1e-3) / 1e9:.0f} GB/s")
    # the same in place (torch.neg_: each element read and written back at
    # its own address, the access pattern of a GatherScatter's values)
    for _ in range(3):
        src.neg_()
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(10):
        src.neg_()
    e1.record(st)
    torch.cuda.synchronize()
    ims = e0.elapsed_time(e1) / 10
    log(f"in-place probe: {16 * n} B in {ims * 1e3:.1f} us = {16 * n / (ims * 1e-3) / 1e9:.0f} GB/s")
    return {"bytes": 16 * n, "us": round(cms * 1e3, 1), "GBps": round(16 * n / (cms * 1e-3) / 1e9, 1),
            "inplace_us": round(ims * 1e3, 1), "inplace_GBps": round(16 * n / (ims * 1e-3) / 1e9, 1)}


def idle_probe(call, stream, gaps_ms=(0, 5, 20, 100, 500), n=120) -> dict:
    """Per-call times (us) after each idle gap: `n` calls with a HIP event
    between consecutive launches, reported as means of runs of 10."""
    import torch
    out = {}
    for gap in gaps_ms:
        torch.cuda.synchronize()
        time.sleep(gap / 1e3)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(n + 1)]
        ev[0].record(stream)
        for i in range(n):
            call()
            ev[i + 1].record(stream)
        torch.cuda.synchronize()
        t = [ev[i].elapsed_time(ev[i + 1]) * 1e3 for i in range(n)]
        out[f"{gap}ms"] = [round(sum(t[i:i + 10]) / 10, 1) for i in range(0, n, 10)]
        log(f"idle {gap:4d} ms, then per 10 calls (us): " + " ".join(f"{x:.1f}" for x in out[f"{gap}ms"]))
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    # the kernel time settles after ~100 calls (~25 ms of work) on a fresh box
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--nx", type=int, default=None, help="elements in x (dawgsSettings.cpp -nx; default 48, "
                                                         "10 with --cc)")
    ap.add_argument("--ny", type=int, default=None, help="elements in y (default: nx)")
    ap.add_argument("--nz", type=int, default=None, help="elements in z (default: nx)")
    ap.add_argument("--p", type=int, default=None, help="degree (default 7, 3 with --cc)")
    ap.add_argument("--cc", action="store_true",
                    help="the reference's correctness check (-cc, dawgsMain.cpp:314-416): the device and host "
                         "GatherScatter (Unsigned and Signed-unique) and Gather+Scatter against gslib on the "
                         "reference's own input, its 'Error = ' lines, and one JSON line")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extras", action="store_true",
                    help="headline only: skip the batched and precision summaries of the default line")
    ap.add_argument("--sweep", nargs="?", const="1,2,3,4,5,6,7", default=None, metavar="DEGREES",
                    help="degree sweep at ~56 M local DOFs (48^3 at p=7): Device (and, unless --no-cpu, "
                         "Host) GatherScatter/Gather/Scatter lines in the reference's format; one JSON "
                         "summary line on stdout")
    ap.add_argument("--k", type=int, default=1, help="fields per node (interleaved) for --sweep")
    ap.add_argument("--batched", action="store_true",
                    help="batched form: 48 meshes (nx,ny,nz ~ U{4..32}, p ~ U{1..7}, K ~ {1,3,8,16}), fp64 "
                         "Add GatherScatter in one persistent graph-captured launch vs one plain launch per "
                         "mesh and field; one JSON line on stdout")
    ap.add_argument("--precision", action="store_true",
                    help="precision form: FP32-compensated and BF16 storage, K=3 fields on the 48^3 p=7 mesh, "
                         "values log-uniform over [1e-3,1e3] with random sign; errors against the fp64 oracle "
                         "and DOFS/s against the plain fp64 kernel; one JSON line on stdout")
    ap.add_argument("--readme", action="store_true",
                    help="the reference README's sample run (34^3 elements, p=1): Device/Host lines in its "
                         "format and one JSON line with its published DOFS/s beside this build's")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the timed region, time a plain device copy of the same byte count "
                         "(the practical streaming ceiling, and a known-size kernel for scaling "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--idle-probe", action="store_true",
                    help="after the timed region, time the GatherScatter call by call after idle gaps of "
                         "0, 5, 20, 100 and 500 ms (means of runs of 10 calls; how the board's power "
                         "management shapes the first ~100 calls after a pause)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        log("bench.py measures one GPU only: the rank-local gather-scatter runs on one MI355X and "
            "has no multi-GPU path (several GPUs would each run it on their own sub-box).")
        return 2

    import torch
    import dawgs_amd as D
    from dawgs_amd.mesh import make_ids

    if not torch.cuda.is_available():
        log("no HIP device visible")
        return 1
    torch.cuda.set_device(0)
    args.nx = args.nx or (10 if args.cc else 48)
    args.p = args.p or (3 if args.cc else 7)
    args.ny = args.ny or args.nx
    args.nz = args.nz or args.nx
    if args.cc:
        return run_cc(args)
    if args.sweep:
        return run_sweep(args)
    if args.batched:
        return run_batched(args)
    if args.precision:
        return run_precision(args)
    if args.readme:
        return run_readme(args)
    nx, ny, nz, p = args.nx, args.ny, args.nz, args.p
    box = f"{nx}^3" if nx == ny == nz else f"{nx}x{ny}x{nz}"

    t0 = time.perf_counter()
    ids = make_ids(nx, ny, nz, p)
    N = len(ids)
    # Inputs are made before the Setup, so that the warm-ups follow the
    # Setup's own device work without an idle gap: on MI355X a memory-bound
    # kernel that resumes after >= 5 ms of idle runs ~10 calls at full
    # speed, then 10-30 calls up to 20% slower while the board's power
    # management reacts, and settles after ~100 calls
    # (profiles/gs_f64_power_transient_48cubed_p7.md).
    gen = torch.Generator(device="cuda").manual_seed(0)
    q = torch.empty(N, dtype=torch.float64, device="cuda").uniform_(-1.0, 1.0, generator=gen)
    stream = torch.cuda.current_stream()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ogs = D.ogs_t()
    # the headline runs K = 1 only: the plan keeps what K = 1 reads
    ogs.Setup(N, ids, D.Unsigned, D.Auto, False, False, fields=D.PlanK1)
    del ids

    for _ in range(args.warmup):
        ogs.GatherScatter(q, 1, D.Add, D.Sym)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(args.steps):
        ogs.GatherScatter(q, 1, D.Add, D.Sym)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / args.steps
    info = ogs.plan.info()
    log(f"setup {time.perf_counter() - t0:.2f} s (incl. timing): N={N} NgatherGlobal={ogs.NgatherGlobal} "
        f"nnz={ogs.nnzT} blocks={info['nblocksK1']} W={info['widthK1']} maxRow={info['maxRowT']} "
        f"deltaE={info['deltaE']} deltaFallback={info['deltaFallback']} byteItems={info['byteItems']} "
        f"byteFallback={info['byteFallback']} planMB={info['bytes'] / 1e6:.0f}")

    B = algorithmic_bytes(ogs.nnzT, ogs.NrowsT, 1, 8)
    achieved = B / (ms * 1e-3) / 1e9
    # measured HBM bytes per launch (two rocprofv3 --pmc runs, scaled by a
    # copy kernel of known size; profiles/k_gs_sym_byte_f64_add_48cubed_p7.md)
    traffic = None
    if os.path.exists(TRAFFIC_FILE) and (nx, ny, nz, p) == (48, 48, 48, 7):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "algorithmic_bytes": B}
    if traffic:
        # the kernel reads no row starts (the pack carries row heads), so it
        # moves fewer bytes than the CSR formulation counted in `achieved`
        phys = traffic / (ms * 1e-3) / 1e9
        roofline["physical_GBps"] = round(phys, 1)
        roofline["physical_frac"] = round(phys / HBM_PEAK_GBS, 4)
    if args.copy_probe:
        roofline["copy_probe"] = copy_probe(B)
    if args.idle_probe:
        roofline["idle_probe"] = idle_probe(lambda: ogs.GatherScatter(q, 1, D.Add, D.Sym), stream)
    log(f"GPU: {ms * 1e3:.1f} us/call, {achieved:.0f} GB/s algorithmic ({100 * achieved / HBM_PEAK_GBS:.1f}% of 8 TB/s), "
        f"{ogs.NgatherGlobal / (ms * 1e-3):.4g} DOFS/s")

    cpu = None
    if not args.no_cpu:
        csr = ogs.csr()
        cpu = cpu_baseline(csr, N, ogs.NgatherGlobal)
        log(f"CPU: {cpu['sample']}: {cpu['value']:.4g} DOFS/s")

    out = {
        "metric": METRIC,
        "value": ogs.NgatherGlobal / (ms * 1e-3),
        "unit": "DOFS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"ogs GatherScatter fp64 Add Sym K=1 on a {box} hex mesh, p={p} "
                               f"({N} local DOFs, {ogs.NgatherGlobal} global), structured numbering, "
                               f"values uniform[-1,1]",
                   "nx": nx, "ny": ny, "nz": nz, "p": p, "K": 1, "op": "Add", "trans": "Sym",
                   "local_dofs": N, "global_dofs": int(ogs.NgatherGlobal), "nnz": ogs.nnzT,
                   "plan_bytes": info["bytes"], "delta_pack_E": info["deltaE"],
                   "byte_codes": {"items": info["byteItems"], "fallback": info["byteFallback"]}},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if not args.no_extras:
        # the metric's second half ("batched/fp32 vs plain"): the batched and
        # precision forms, summarised; their full lines come from --batched /
        # --precision.  A failure here leaves the headline line intact.
        del q
        ex = argparse.Namespace(**vars(args))
        ex.warmup, ex.steps = 50, 50
        try:
            pr = precision_result(ex)
            f = pr["forms"]
            out["precision"] = {
                "workload": pr["config"]["workload"],
                "fp32_compensated_batched_vs_fp64_plain": pr["speedup_vs_fp64_plain"],
                "fp32_compensated_batched_DOFSps": pr["value"],
                "fp64_plain_per_field_DOFSps": f["fp64_plain_per_field"]["DOFSps"],
                "bf16_vs_fp64_plain": f["bf16_k3"]["speedup_vs_fp64_plain"],
                "errors_vs_fp64": pr["errors_vs_fp64"]}
        except Exception as e:  # noqa: BLE001
            out["precision"] = {"error": repr(e)}
        try:
            bt = batched_result(ex)
            out["batched"] = {"workload": bt["config"]["workload"], "value": bt["value"], "unit": bt["unit"],
                              "speedup_vs_plain": bt["speedup_vs_plain"], "ms_per_step": bt["ms_per_step"],
                              "plain_ms_per_step": bt["plain"]["ms_per_step"], "roofline_frac": bt["roofline"]["frac"]}
        except Exception as e:  # noqa: BLE001
            out["batched"] = {"error": repr(e)}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
