"""GPU parity of the batched GatherScatter (ogs_hip_batch_*): one persistent,
graph-captured launch over many meshes with different K, checked per mesh
against the oracle's host loop on the same CSR, bit-exact."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

SPECS = [  # nx, ny, nz, p, K
    (3, 4, 2, 1, 1), (2, 2, 5, 7, 3), (4, 3, 3, 3, 8), (1, 1, 1, 2, 16),
    (5, 2, 2, 5, 1), (2, 3, 4, 4, 3), (6, 5, 4, 2, 100), (3, 3, 3, 6, 1),
]


@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available()
    return torch


@pytest.fixture(scope="module")
def meshes():
    import dawgs_amd as D
    rng = np.random.default_rng(31)
    out = []
    for i, (nx, ny, nz, p, K) in enumerate(SPECS):
        ids = O.make_ids(nx, ny, nz, p).copy()
        kind = O.UNSIGNED
        if i == 2:  # one ragged, flagged Signed mesh: Sym uses all members
            ids[rng.random(len(ids)) < 0.1] = 0
            neg = rng.random(len(ids)) < 0.3
            ids[neg] = -ids[neg]
            kind = O.SIGNED
        g = D.ogs_t()
        g.Setup(len(ids), ids.copy(), kind)
        c = g.csr()
        out.append((g, O.OracleCsr(c["rowStartsT"], c["colIdsT"], Ncols=len(ids)), K, len(ids)))
    return out


def values(rng, dt, n):
    if np.issubdtype(dt, np.floating):
        return rng.uniform(-1, 1, n).astype(dt)
    return rng.integers(-2 ** 20, 2 ** 20 + 1, n).astype(dt)


@pytest.mark.parametrize("tune", [True, False])
@pytest.mark.parametrize("dt,op", [(np.float64, O.ADD), (np.float32, O.ADD), (np.float64, O.MUL),
                                   (np.int64, O.MIN), (np.int32, O.MAX)])
def test_batch_matches_oracle(torch, meshes, dt, op, tune):
    import dawgs_amd as D
    rng = np.random.default_rng(7 + op)
    host = [values(rng, dt, N * K) for (_, _, K, N) in meshes]
    bufs = [torch.from_numpy(h.copy()).cuda() for h in host]
    batch = D.Batch([(g.plan, K, b) for (g, _, K, _), b in zip(meshes, bufs)], op, tune=tune)
    torch.cuda.synchronize()
    for h, b in zip(host, bufs):  # creation (and its tuning) leaves the buffers alone
        assert b.cpu().numpy().tobytes() == h.tobytes()
    info = batch.info()
    assert info["items"] > 0 and (info["candidates"] > 0) == tune
    want = [ref.gather_scatter(h.copy(), K, op, O.SYM) for (_, ref, K, _), h in zip(meshes, host)]
    batch.run()
    torch.cuda.synchronize()
    for w, b in zip(want, bufs):
        assert b.cpu().numpy().tobytes() == w.tobytes()
    # a second replay of the graph applies the operator again
    want2 = [ref.gather_scatter(w.copy(), K, op, O.SYM) for (_, ref, K, _), w in zip(meshes, want)]
    batch.run()
    torch.cuda.synchronize()
    for w, b in zip(want2, bufs):
        assert b.cpu().numpy().tobytes() == w.tobytes()


def test_batch_equals_per_field_launches(torch, meshes):
    """One batched launch over interleaved K fields == K separate K=1 launches
    on each field's own array (the plain method), bit for bit."""
    import dawgs_amd as D
    rng = np.random.default_rng(99)
    host = [values(rng, np.float64, N * K) for (_, _, K, N) in meshes]
    bufs = [torch.from_numpy(h.copy()).cuda() for h in host]
    D.Batch([(g.plan, K, b) for (g, _, K, _), b in zip(meshes, bufs)], D.Add, tune=False).run()
    for (g, _, K, N), h, b in zip(meshes, host, bufs):
        fields = [torch.from_numpy(np.ascontiguousarray(h.reshape(N, K)[:, k])).cuda() for k in range(K)]
        for f in fields:
            g.GatherScatter(f, 1, D.Add, D.Sym)
        torch.cuda.synchronize()
        plain = np.stack([f.cpu().numpy() for f in fields], axis=1).reshape(-1)
        assert b.cpu().numpy().tobytes() == plain.tobytes()


def test_batch_mixed_byte_coded_and_small(torch):
    """Two meshes large enough for the byte-coded kernels (K = 1 and K = 3)
    beside small ones: whichever form the tuner picks (batched kernels or
    one launch per descriptor through its plan), each mesh is bit-exact
    after two runs of the graph."""
    import dawgs_amd as D
    specs = [(16, 16, 16, 7, 1), (16, 16, 16, 7, 3), (3, 4, 2, 1, 1), (2, 2, 5, 7, 3), (4, 3, 3, 3, 8)]
    rng = np.random.default_rng(57)
    descs, host, refs = [], [], []
    keep = []
    for nx, ny, nz, p, K in specs:
        ids = O.make_ids(nx, ny, nz, p)
        g = D.ogs_t()
        g.Setup(len(ids), ids.copy(), O.UNSIGNED)
        c = g.csr()
        refs.append((O.OracleCsr(c["rowStartsT"], c["colIdsT"], Ncols=len(ids)), K))
        h = rng.uniform(-1, 1, K * len(ids))
        host.append(h)
        keep.append(g)
        descs.append((g.plan, K, torch.from_numpy(h.copy()).cuda()))
    batch = D.Batch(descs, D.Add, tune=True)
    info = batch.info()
    print("mixed batch:", info)
    assert info["items"] > 0 and info["plans"] in (0, 1), info
    for _ in range(2):
        batch.run()
    torch.cuda.synchronize()
    for (ref, K), h, (_, _, b) in zip(refs, host, descs):
        want = ref.gather_scatter(ref.gather_scatter(h.copy(), K, O.ADD, O.SYM), K, O.ADD, O.SYM)
        assert b.cpu().numpy().tobytes() == want.tobytes()
