"""GPU parity of the precision form: OGS_BF16 (bfloat16 storage) and
OGS_FLOAT_C (float storage), Add reduced in float with two-sum compensation,
Max/Min exact, Mul refused.

  * bit-exact against the pinned oracle of the same type, for every
    {op x trans x K} on Unsigned / Signed / Signed-unique meshes, through the
    handle layer (ogs_*) and the operator layer (ogs_hip_*), and batched;
  * hand-built rows whose exact sums plain float summation misses;
  * the 48^3 p=7, K=3 workload (values log-uniform in magnitude over
    [1e-3, 1e3], random sign): bit-exact against the oracle, and the error
    against the fp64 oracle of the unrounded fp64 values within a bound
    derived from the storage and accumulation roundoffs:
        |res_i - y_i| <= u_s a_i + (u_f + u_o)(|y_i| + u_s a_i) + 2 g^2 a_i
    with y = GS_fp64(x), a = GS_fp64(|x|), u_s the storage roundoff of the
    inputs (2^-24 float, 2^-8 bf16), u_f = 2^-24 the float accumulation,
    u_o the output rounding (0 float, 2^-8 bf16) and g = gamma_7 (rows hold
    at most 8 members), plus a 1e-3 relative slack on the bound itself.
    FLOAT_C must also meet the normwise target ||res - y||_2 / ||y||_2 <= 1e-6.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

PTYPES = [O.BF16, O.FLOAT_C]
POPS = [O.ADD, O.MAX, O.MIN]
TRANS = [O.SYM, O.NOTRANS, O.TRANS]


@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available()
    return torch


@pytest.fixture(scope="module")
def D():
    import dawgs_amd
    return dawgs_amd


def logu(rng, n):
    return 10.0 ** rng.uniform(-3, 3, n) * rng.choice([-1.0, 1.0], n)


def stored(ty, f64):
    """Host storage of fp64 values in the given precision-form type."""
    f = np.asarray(f64, np.float64).astype(np.float32)
    return O.float_to_bf16(f) if ty == O.BF16 else f


def widen(ty, h):
    return O.bf16_to_float(h) if ty == O.BF16 else h


def to_dev(torch, ty, h):
    if ty == O.BF16:
        return torch.from_numpy(np.ascontiguousarray(h).view(np.int16)).cuda().view(torch.bfloat16)
    return torch.from_numpy(np.ascontiguousarray(h)).cuda()


def to_host(torch, ty, t):
    torch.cuda.synchronize()
    if ty == O.BF16:
        return t.view(torch.int16).cpu().numpy().view(np.uint16)
    return t.cpu().numpy()


def empty_dev(torch, ty, n, fill=0):
    dt = torch.bfloat16 if ty == O.BF16 else torch.float32
    return torch.full((max(n, 1),), fill, dtype=dt, device="cuda")


MESHES = {}


def mesh(kind):
    if kind in MESHES:
        return MESHES[kind]
    import dawgs_amd as D
    rng = np.random.default_rng(17)
    ids = O.make_ids(5, 6, 4, 3).copy()
    ids[rng.random(len(ids)) < 0.08] = 0
    k, u = O.UNSIGNED, False
    if kind != "unsigned":
        neg = rng.random(len(ids)) < 0.35
        ids[neg] = -ids[neg]
        k, u = O.SIGNED, kind == "unique"
    o = O.OracleOgs(ids, k, u)
    g = D.ogs_t()
    dids = ids.copy()
    g.Setup(len(dids), dids, k, D.Auto, u, False)
    MESHES[kind] = (dids, o, g)
    return MESHES[kind]


@pytest.mark.parametrize("K", [1, 3, 8])
@pytest.mark.parametrize("trans", TRANS)
@pytest.mark.parametrize("op", POPS)
@pytest.mark.parametrize("ty", PTYPES)
@pytest.mark.parametrize("kind", ["unsigned", "signed", "unique"])
def test_precision_operators(torch, D, kind, ty, op, trans, K):
    ids, o, g = mesh(kind)
    rng = np.random.default_rng(100 * ty + 10 * op + trans + K)
    v = stored(ty, logu(rng, len(ids) * K))
    # GatherScatter through the handle layer
    want = o.op_gather_scatter(v.copy(), K, op, trans, type=ty)
    d = to_dev(torch, ty, v)
    g.GatherScatter(d, K, op, trans, type=D.FloatC if ty == O.FLOAT_C else None)
    assert to_host(torch, ty, d).tobytes() == want.tobytes()
    # Gather and Scatter through the operator layer
    wg = o.op_gather(v, K, op, trans, type=ty)
    gv = empty_dev(torch, ty, o.NlocalT * K, 7)
    g.plan.Gather(gv, to_dev(torch, ty, v), K, op, trans, type=D.FloatC if ty == O.FLOAT_C else None)
    got = to_host(torch, ty, gv)
    assert got[: len(wg)].tobytes() == wg.tobytes()
    if op == O.ADD:
        base = stored(ty, logu(rng, len(ids) * K))
        gvals = stored(ty, logu(rng, o.NlocalT * K))
        ws = o.op_scatter(base.copy(), gvals, K, trans, type=ty)
        s = to_dev(torch, ty, base)
        g.plan.Scatter(s, to_dev(torch, ty, gvals), K, trans, type=D.FloatC if ty == O.FLOAT_C else None)
        assert to_host(torch, ty, s).tobytes() == ws.tobytes()


@pytest.mark.parametrize("K", [16, 100])
@pytest.mark.parametrize("ty", PTYPES)
def test_precision_many_fields(torch, D, ty, K):
    ids, o, g = mesh("signed")
    rng = np.random.default_rng(K + ty)
    v = stored(ty, logu(rng, len(ids) * K))
    for trans in TRANS:
        want = o.op_gather_scatter(v.copy(), K, O.ADD, trans, type=ty)
        d = to_dev(torch, ty, v)
        g.plan.GatherScatter(d, K, O.ADD, trans, type=ty)
        assert to_host(torch, ty, d).tobytes() == want.tobytes(), trans


KNOWN = [  # float32 rows and their exact sums; plain float summation misses each
    ([1.0, 2.0 ** -30, -1.0], 2.0 ** -30),
    ([2.0 ** 24, 1.0, 1.0], 2.0 ** 24 + 2),
    ([1.0, 2.0 ** -25, 2.0 ** -25, 2.0 ** -25, 2.0 ** -25], 1.0 + 2.0 ** -23),
    ([3.0e7, 1.0, -3.0e7, 0.5], 1.5),
]


def test_compensation_known_answers(torch, D):
    rows = [np.array(r, np.float32) for r, _ in KNOWN]
    v = np.concatenate(rows)
    rs = np.cumsum([0] + [len(r) for r in rows]).astype(np.int32)
    ci = np.arange(len(v), dtype=np.int32)
    drs, dci = torch.from_numpy(rs).cuda(), torch.from_numpy(ci).cuda()
    plan = D.Plan(len(v), len(rows), len(rows), drs, dci, drs, dci)
    d = to_dev(torch, O.FLOAT_C, v)
    plan.GatherScatter(d, 1, D.Add, D.Sym, type=D.FloatC)
    got = to_host(torch, O.FLOAT_C, d)
    for (r, exact), a, b in zip(KNOWN, rs[:-1], rs[1:]):
        assert np.all(got[a:b] == np.float32(exact)), (r, got[a:b])
    # the plain float type keeps the reference's uncompensated sum
    d = to_dev(torch, O.FLOAT_C, v)
    plan.GatherScatter(d, 1, D.Add, D.Sym)
    assert to_host(torch, O.FLOAT_C, d)[0] == np.float32(0.0)
    # bf16: 1 + 2^-12 - 1 = 2^-12 (all three representable)
    b = O.float_to_bf16(np.array([1.0, 2.0 ** -12, -1.0], np.float32))
    r1 = torch.tensor([0, 3], dtype=torch.int32, device="cuda")
    c1 = torch.arange(3, dtype=torch.int32, device="cuda")
    one = D.Plan(3, 1, 1, r1, c1, r1, c1)
    db = to_dev(torch, O.BF16, b)
    one.GatherScatter(db, 1, D.Add, D.Sym)
    assert np.all(widen(O.BF16, to_host(torch, O.BF16, db)) == np.float32(2.0 ** -12))


def test_precision_special_values(torch, D):
    ids, o, g = mesh("unsigned")
    rng = np.random.default_rng(5)
    f = logu(rng, len(ids)).astype(np.float32)
    pos = rng.choice(len(ids), 60, replace=False)
    f[pos[:10]] = np.inf
    f[pos[10:20]] = -np.inf
    f[pos[20:30]] = np.nan
    f[pos[30:40]] = -0.0
    f[pos[40:50]] = 3.0e38
    f[pos[50:60]] = -3.0e38
    for ty in PTYPES:
        v = stored(ty, f)
        for op in POPS:
            want = o.op_gather_scatter(v.copy(), 1, op, O.SYM, type=ty)
            d = to_dev(torch, ty, v)
            g.plan.GatherScatter(d, 1, op, O.SYM, type=ty)
            assert to_host(torch, ty, d).tobytes() == want.tobytes(), (ty, op)


def test_precision_refuses_mul(torch, D):
    from dawgs_amd._lib import OgsError
    ids, o, g = mesh("unsigned")
    for ty in PTYPES:
        d = empty_dev(torch, ty, len(ids))
        with pytest.raises(OgsError) as e:
            g.GatherScatter(d, 1, D.Mul, D.Sym, type=D.FloatC if ty == O.FLOAT_C else None)
        assert e.value.code == -6
        with pytest.raises(OgsError) as e:
            g.plan.Gather(empty_dev(torch, ty, o.NlocalT), d, 1, D.Mul, D.Trans, type=ty)
        assert e.value.code == -6
        with pytest.raises(OgsError) as e:
            D.Batch([(g.plan, 1, d)], D.Mul, tune=False, type=ty)
        assert e.value.code == -6


BATCH_SPECS = [(3, 4, 2, 1, 1), (2, 2, 5, 7, 3), (4, 3, 3, 3, 8), (1, 1, 1, 2, 16), (2, 3, 4, 4, 3), (3, 2, 2, 6, 5)]


@pytest.mark.parametrize("op", POPS)
@pytest.mark.parametrize("ty", PTYPES)
def test_precision_batch(torch, D, ty, op):
    rng = np.random.default_rng(60 + ty + op)
    meshes = []
    for nx, ny, nz, p, K in BATCH_SPECS:
        ids = O.make_ids(nx, ny, nz, p)
        g = D.ogs_t()
        g.Setup(len(ids), ids.copy(), D.Unsigned)
        c = g.csr()
        meshes.append((g, O.OracleCsr(c["rowStartsT"], c["colIdsT"], Ncols=len(ids)), K, len(ids)))
    host = [stored(ty, logu(rng, N * K)) for (_, _, K, N) in meshes]
    bufs = [to_dev(torch, ty, h) for h in host]
    batch = D.Batch([(g.plan, K, b) for (g, _, K, _), b in zip(meshes, bufs)], op, tune=True, type=ty)
    torch.cuda.synchronize()
    for h, b in zip(host, bufs):
        assert to_host(torch, ty, b).tobytes() == h.tobytes()
    batch.run()
    for (_, ref, K, _), h, b in zip(meshes, host, bufs):
        want = ref.gather_scatter(h.copy(), K, op, O.SYM, type=ty)
        assert to_host(torch, ty, b).tobytes() == want.tobytes()


def error_stats(res, y, a):
    d = res.astype(np.float64) - y
    big = np.abs(y) > 1e-300
    return {"max_abs": float(np.max(np.abs(d))),
            "normwise_rel": float(np.linalg.norm(d) / np.linalg.norm(y)),
            "max_elem_rel": float(np.max(np.abs(d[big]) / np.abs(y[big])))}


def test_precision_48cubed_p7_k3(torch, D):
    """The precision workload at full size: bit-exact and within the bound."""
    K = 3
    ids = O.make_ids(48, 48, 48, 7)
    N = len(ids)
    g = D.ogs_t()
    g.Setup(N, ids.copy(), D.Unsigned)
    c = g.csr()
    ref = O.OracleCsr(c["rowStartsT"], c["colIdsT"], Ncols=N)
    rng = np.random.default_rng(48)
    x = logu(rng, N * K)
    y = ref.gather_scatter(x.copy(), K, O.ADD, O.SYM)
    a = ref.gather_scatter(np.abs(x), K, O.ADD, O.SYM)
    uf = 2.0 ** -24
    g7 = 7 * uf / (1 - 7 * uf)
    for ty, us, uo in ((O.FLOAT_C, 2.0 ** -24, 0.0), (O.BF16, 2.0 ** -8, 2.0 ** -8)):
        v = stored(ty, x)
        want = ref.gather_scatter(v.copy(), K, O.ADD, O.SYM, type=ty)
        d = to_dev(torch, ty, v)
        g.plan.GatherScatter(d, K, D.Add, D.Sym, type=ty)
        got = to_host(torch, ty, d)
        del d
        assert got.tobytes() == want.tobytes(), ty
        res = widen(ty, got).astype(np.float64)
        bound = (1 + 1e-3) * (us * a + (uf + uo) * (np.abs(y) + us * a)) + 2 * g7 ** 2 * a
        err = np.abs(res - y)
        assert np.all(err <= bound), (ty, float(np.max(err / bound)))
        st = error_stats(res, y, a)
        print(f"type {ty}: {st}")
        if ty == O.FLOAT_C:
            assert st["normwise_rel"] <= 1e-6
