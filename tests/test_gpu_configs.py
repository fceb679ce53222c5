"""GPU parity of the configurations at their real sizes, bit-exact against
the oracle on the same CSR and inputs:

  * the batched form exactly as the benchmark runs it: 48 meshes with
    nx, ny, nz ~ U{4..32}, p ~ U{1..7}, K ~ {1, 3, 8, 16} (seed 48), fp64
    Add, tuned graph, every mesh checked after one replay;
  * the precision form's one-descriptor batches at 48^3 p=7 K=3 (float
    storage with compensated Add, and bfloat16 storage), against the oracle
    of the same storage type;
  * the degree sweep p = 2..6 at ~56 M local DOFs (p = 1 and 7 run in the
    parity suite): GatherScatter Sym, Gather Trans and Scatter NoTrans.
"""
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available()
    return torch


@pytest.fixture(scope="module")
def D():
    import dawgs_amd
    return dawgs_amd


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.dtype == b.dtype and a.shape == b.shape and a.tobytes() == b.tobytes()


def test_batched_config_48_meshes(torch, D):
    sys.path.insert(0, ROOT)
    from bench import batch_meshes
    specs = batch_meshes()
    assert len(specs) == 48
    rng = np.random.default_rng(48)
    ogss, host, bufs = [], [], []
    for (nx, ny, nz, p, K) in specs:
        ids = O.make_ids(nx, ny, nz, p)
        g = D.ogs_t()
        g.Setup(len(ids), ids.copy(), D.Unsigned)
        v = rng.uniform(-1, 1, len(ids) * K)
        ogss.append(g)
        host.append(v)
        bufs.append(torch.from_numpy(v.copy()).cuda())
    batch = D.Batch([(g.plan, K, b) for g, (*_, K), b in zip(ogss, specs, bufs)], D.Add, tune=True)
    info = batch.info()
    assert info["candidates"] > 0 and info["items"] > 0
    batch.run()
    torch.cuda.synchronize()
    for i, (g, (nx, ny, nz, p, K), h, b) in enumerate(zip(ogss, specs, host, bufs)):
        c = g.csr()
        want = O.OracleCsr(c["rowStartsT"], c["colIdsT"], Ncols=g.N).gather_scatter(h, K, O.ADD, O.SYM)
        assert same(b.cpu().numpy(), want), (i, nx, ny, nz, p, K, info)


@pytest.fixture(scope="module")
def mesh48(D):
    ids = O.make_ids(48, 48, 48, 7)
    g = D.ogs_t()
    g.Setup(len(ids), ids, D.Unsigned)
    c = g.csr()
    return g, O.OracleCsr(c["rowStartsT"], c["colIdsT"], Ncols=g.N)


@pytest.mark.parametrize("ty", ["FloatC", "BF16"])
def test_precision_batch_48cubed_p7_k3(torch, D, mesh48, ty):
    g, ref = mesh48
    K, N = 3, g.N
    rng = np.random.default_rng(3)
    x = (10.0 ** rng.uniform(-3, 3, N * K)) * rng.choice([-1.0, 1.0], N * K)
    f = x.astype(np.float32)
    if ty == "BF16":
        h = O.float_to_bf16(f)
        buf = torch.from_numpy(h.view(np.int16).copy()).cuda().view(torch.bfloat16)
        code, oty = D.BF16, O.BF16
    else:
        h = f
        buf = torch.from_numpy(h.copy()).cuda()
        code, oty = D.FloatC, O.FLOAT_C
    batch = D.Batch([(g.plan, K, buf)], D.Add, tune=True, type=code)
    batch.run()
    torch.cuda.synchronize()
    got = buf.view(torch.int16).cpu().numpy().view(np.uint16) if ty == "BF16" else buf.cpu().numpy()
    assert same(got, ref.gather_scatter(h.copy(), K, O.ADD, O.SYM, type=oty))


@pytest.mark.parametrize("p", [2, 3, 4, 5, 6])
def test_degree_sweep_full_size(torch, D, p):
    from dawgs_amd.perf import SWEEP_NX
    nx = SWEEP_NX[p]
    ids = O.make_ids(nx, nx, nx, p)
    N = len(ids)
    g = D.ogs_t()
    g.Setup(N, ids.copy(), D.Unsigned)
    s = D.ogs_t()
    sid = ids.copy()
    s.Setup(N, sid, D.Signed, D.Auto, True, False)  # the sweep's Signed-unique sogs
    del ids
    c = g.csr()
    ref = O.OracleCsr(c["rowStartsT"], c["colIdsT"], Ncols=N)
    cs = s.csr()
    sref = O.OracleCsr(cs["rowStartsT"], cs["colIdsT"], cs["rowStartsN"], cs["colIdsN"], s.NrowsN, N)
    rng = np.random.default_rng(100 + p)
    v = rng.uniform(-1, 1, N)
    d = torch.from_numpy(v.copy()).cuda()
    g.GatherScatter(d, 1, D.Add, D.Sym)
    torch.cuda.synchronize()
    assert same(d.cpu().numpy(), ref.gather_scatter(v.copy(), 1, O.ADD, O.SYM)), "GatherScatter"
    wg = sref.gather(v, 1, O.ADD, O.TRANS)
    gq = torch.zeros(len(wg), dtype=torch.float64, device="cuda")
    s.Gather(gq, torch.from_numpy(v).cuda(), 1, D.Add, D.Trans)
    torch.cuda.synchronize()
    assert same(gq.cpu().numpy(), wg), "Gather"
    base = rng.uniform(-1, 1, N)
    d = torch.from_numpy(base.copy()).cuda()
    s.Scatter(d, gq, 1, D.NoTrans)
    torch.cuda.synchronize()
    assert same(d.cpu().numpy(), sref.scatter(base.copy(), wg, 1, O.NOTRANS)), "Scatter"
