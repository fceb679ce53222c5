"""GPU checks of the handle contracts in include/ogs_hip.h:

  * one ogs_t handle shared by two host threads, each launching on its own
    stream: every call runs on the stream it was given (a launch on the other
    thread's stream would let the thread's next stream-ordered step overtake
    it), results bit-exact against the oracle;
  * plan lifetime: a batch and a borrowed plan keep working after the ogs_t
    that made the plan is freed, freed in either order.
"""
import threading

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


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


def test_shared_handle_two_threads_two_streams(torch, D):
    ids = O.make_ids(20, 20, 20, 3)
    N = len(ids)
    g = D.ogs_t()
    sid = ids.copy()
    g.Setup(N, sid, D.Signed, D.Auto, True, False)
    c = g.csr()
    ref = O.OracleCsr(c["rowStartsT"], c["colIdsT"], c["rowStartsN"], c["colIdsN"], g.NrowsN, N)
    rounds = 40
    rng = np.random.default_rng(5)
    starts = [rng.uniform(-1, 1, N) for _ in range(2)]
    # each round: GatherScatter on the thread's stream, then a stream-ordered
    # halving of the buffer and a Gather Trans into the thread's row buffer
    want = []
    for v in starts:
        x = v.copy()
        for _ in range(rounds):
            x = ref.gather_scatter(x, 1, O.ADD, O.SYM) * 0.5
        want.append((x, ref.gather(x, 1, O.ADD, O.TRANS)))
    bufs = [torch.from_numpy(v.copy()).cuda() for v in starts]
    rows = [torch.zeros(g.NrowsT, dtype=torch.float64, device="cuda") for _ in range(2)]
    streams = [torch.cuda.Stream() for _ in range(2)]
    torch.cuda.synchronize()
    errors = []
    barrier = threading.Barrier(2)

    def work(i):
        try:
            s = streams[i]
            with torch.cuda.stream(s):
                barrier.wait()
                for _ in range(rounds):
                    g.GatherScatter(bufs[i], 1, D.Add, D.Sym, stream=s.cuda_stream)
                    bufs[i].mul_(0.5)
                g.Gather(rows[i], bufs[i], 1, D.Add, D.Trans, stream=s.cuda_stream)
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    threads = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    torch.cuda.synchronize()
    assert not errors, errors
    for i in range(2):
        assert same(bufs[i].cpu().numpy(), want[i][0]), i
        assert same(rows[i].cpu().numpy(), want[i][1]), i


def _mesh(D, box=(12, 10, 8, 3)):
    ids = O.make_ids(*box)
    g = D.ogs_t()
    g.Setup(len(ids), ids.copy(), D.Unsigned)
    c = g.csr()
    return g, O.OracleCsr(c["rowStartsT"], c["colIdsT"], Ncols=len(ids)), len(ids)


@pytest.mark.parametrize("order", ["owner_first", "batch_first"])
def test_batch_outlives_or_precedes_its_plans_owner(torch, D, order):
    g1, r1, n1 = _mesh(D, (12, 10, 8, 3))
    g2, r2, n2 = _mesh(D, (6, 7, 5, 5))
    rng = np.random.default_rng(12)
    h1, h2 = rng.uniform(-1, 1, n1 * 3), rng.uniform(-1, 1, n2)
    b1, b2 = torch.from_numpy(h1.copy()).cuda(), torch.from_numpy(h2.copy()).cuda()
    batch = D.Batch([(g1.plan, 3, b1), (g2.plan, 1, b2)], D.Add, tune=False)
    if order == "owner_first":
        g1.Free()
        g2.Free()
        del g1, g2
        batch.run()
        torch.cuda.synchronize()
        assert same(b1.cpu().numpy(), r1.gather_scatter(h1.copy(), 3, O.ADD, O.SYM))
        assert same(b2.cpu().numpy(), r2.gather_scatter(h2.copy(), 1, O.ADD, O.SYM))
        del batch
    else:
        batch.run()
        torch.cuda.synchronize()
        del batch
        # the owners still run their own operator after the batch is gone
        g1.GatherScatter(b1, 3, D.Add, D.Sym)
        torch.cuda.synchronize()
        w = r1.gather_scatter(r1.gather_scatter(h1.copy(), 3, O.ADD, O.SYM), 3, O.ADD, O.SYM)
        assert same(b1.cpu().numpy(), w)
        g1.Free()
        g2.Free()


def test_borrowed_plan_outlives_owner(torch, D):
    g, ref, N = _mesh(D)
    p = g.plan
    g.Free()
    del g
    v = np.random.default_rng(3).uniform(-1, 1, N)
    d = torch.from_numpy(v.copy()).cuda()
    p.GatherScatter(d, 1, D.Add, D.Sym)
    torch.cuda.synchronize()
    assert same(d.cpu().numpy(), ref.gather_scatter(v.copy(), 1, O.ADD, O.SYM))


def test_owner_map_not_built_over_repeated_t_columns(torch, D):
    """A hand-built Signed operator with at most one N member per row but a
    column repeated across T rows (T = {a, b}, {a, c}; N = {a}, {c}): the
    column -> owner map would let column b read a value that column a's
    owner-row write has replaced, so the plan does not build it and the
    row-based kernel (both rows' gathers before any write, in one block)
    runs: b and c get their rows' owners' original values."""
    import dawgs_amd as D2
    rsT = np.array([0, 2, 4], np.int32)
    ciT = np.array([0, 1, 0, 2], np.int32)
    rsN = np.array([0, 1, 2], np.int32)
    ciN = np.array([0, 2], np.int32)
    dev = lambda a: torch.from_numpy(a).cuda()
    plan = D2.Plan(3, 2, 2, dev(rsN), dev(ciN), dev(rsT), dev(ciT))
    info = plan.info()
    assert info["ownerMap"] == 0 and info["nMask"] == 0, info
    # the same rows with distinct T columns (T = {a, b}, {c, d}) get both
    ok = D2.Plan(4, 2, 2, dev(rsN), dev(np.array([0, 2], np.int32)), dev(rsT), dev(np.array([0, 1, 2, 3], np.int32)))
    assert ok.info()["ownerMap"] == 1 and ok.info()["nMask"] == 1
    v0 = np.array([10.0, 20.0, 30.0])
    d = dev(v0.copy())
    plan.GatherScatter(d, 1, D.Add, D.NoTrans)
    torch.cuda.synchronize()
    r = d.cpu().numpy()
    assert r[1] == 10.0 and r[2] == 30.0 and r[0] in (10.0, 30.0), r
