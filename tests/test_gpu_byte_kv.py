"""GPU parity of the one-list GatherScatter Sym and the Gather with three
interleaved fields over the K = 1 kernel's one-byte codes (k_gs_sym_byte_kv,
k_gather_byte_kv; the Gather on partition 0, never on bundled blocks): bit-exact
against the oracle for the reference's four types and every op, and for
float-compensated storage, on meshes whose partition 0 fills the device, for
blocks that fall back to the 16-bit and 32-bit words, on bundled low-degree
blocks and on Signed operators.  bfloat16 storage, and lists with rows of
more than 8 members, keep the entry-major kernel: checked on the same meshes.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TYPES = {0: np.float32, 1: np.float64, 2: np.int32, 3: np.int64}
K = 3


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


def values(rng, dt, n):
    """Uniform values with 1% signed zeros and NaN among them (no infinities,
    whose NaN results carry a sign that differs between host and GPU)."""
    if np.issubdtype(dt, np.floating):
        v = rng.uniform(-1, 1, n).astype(dt)
        at = rng.random(n) < 0.01
        v[at] = np.array([0.0, -0.0, np.nan], dtype=dt)[rng.integers(0, 3, int(at.sum()))]
        return v
    return rng.integers(-2 ** 20, 2 ** 20 + 1, n).astype(dt)


_CACHE = {}


def mesh(D, box, fields=0, kind=O.UNSIGNED, ids=None, key=None):
    key = key or (box, fields, kind)
    if key not in _CACHE:
        ids = O.make_ids(*box) if ids is None else ids
        g = D.ogs_t()
        g.Setup(len(ids), ids.copy(), kind, D.Auto, False, False, fields=fields)
        c = g.csr()
        _CACHE[key] = (g, O.OracleCsr(c["rowStartsT"], c["colIdsT"], Ncols=len(ids)), len(ids))
    return _CACHE[key]


def run(torch, D, g, ref, N, dt, op, seed):
    v = values(np.random.default_rng(seed), dt, K * N)
    want = ref.gather_scatter(v.copy(), K, op, O.SYM)
    d = torch.from_numpy(v.copy()).cuda()
    g.GatherScatter(d, K, op, D.Sym)
    torch.cuda.synchronize()
    return d.cpu().numpy(), want


BOXES = [(16, 16, 16, 7), (22, 22, 22, 5), (36, 32, 32, 3)]


@pytest.mark.parametrize("box", BOXES, ids=lambda b: "x".join(map(str, b)))
@pytest.mark.parametrize("ty", list(TYPES))
@pytest.mark.parametrize("op", [0, 1, 2, 3])
def test_byte_k3(torch, D, box, ty, op):
    g, ref, N = mesh(D, box)
    info = g.plan.info()
    assert info["byteK3"] == 1 and info["maxRowT"] <= 8, info
    got, want = run(torch, D, g, ref, N, TYPES[ty], op, hash((box, ty, op, "k3")) % 2 ** 32)
    assert same(got, want), (box, np.dtype(TYPES[ty]).name, op)


def test_byte_k3_k1_plan(torch, D):
    """A plan prepared for K = 1 only runs K = 3 over the same codes."""
    g, ref, N = mesh(D, (16, 16, 16, 7), fields=D.PlanK1)
    assert g.plan.info()["byteK3"] == 1
    got, want = run(torch, D, g, ref, N, np.float64, O.ADD, 17)
    assert same(got, want)


@pytest.mark.parametrize("ty,op", [(O.FLOAT_C, 0), (O.FLOAT_C, 2), (O.BF16, 0), (O.BF16, 3)])
def test_byte_k3_precision_types(torch, D, ty, op):
    """Float-compensated storage runs the byte kernel (each row folded once,
    by the thread holding its head); bfloat16 the entry-major kernel."""
    g, ref, N = mesh(D, (16, 16, 16, 7))
    rng = np.random.default_rng(50 + ty + op)
    f = (rng.uniform(-1, 1, K * N) * 10.0 ** rng.uniform(-3, 3, K * N)).astype(np.float32)
    if ty == O.BF16:
        v = O.float_to_bf16(f)
        d = torch.from_numpy(v.view(np.int16).copy()).cuda().view(torch.bfloat16)
    else:
        v = f
        d = torch.from_numpy(v.copy()).cuda()
    want = ref.gather_scatter(v.copy(), K, op, O.SYM, type=ty)
    g.GatherScatter(d, K, op, D.Sym, type=(D.FloatC if ty == O.FLOAT_C else None))
    torch.cuda.synchronize()
    got = d.view(torch.int16).cpu().numpy().view(np.uint16) if ty == O.BF16 else d.cpu().numpy()
    assert same(got, want)


def shuffled(box, frac, seed):
    ids = O.make_ids(*box).copy()
    n = len(ids)
    a = int(n * (1 - frac))
    rng = np.random.default_rng(seed)
    ids[a:] = ids[a:][rng.permutation(n - a)]
    return ids


@pytest.mark.parametrize("fields", [0, 1], ids=["all_k", "k1_plan"])
def test_byte_k3_fallback_blocks(torch, D, fields):
    """Shuffled tail: blocks without codes run from their 16-bit words, and
    blocks without those from their 32-bit words (in a K = 1 plan, the
    compacted copy found through the exception word's slot)."""
    ids = shuffled((36, 32, 32, 3), 0.08, 3)
    g, ref, N = mesh(D, None, fields=fields, ids=ids, key=("shuffled", fields))
    info = g.plan.info()
    if info["plainK1"] > 0:
        pytest.skip(f"blocks left in CSR order: the entry-major kernel serves ({info})")
    assert info["byteK3"] == 1 and info["byteFallback"] > 0 and info["deltaFallback"] > 0, info
    for dt, op in ((np.float64, 0), (np.int32, 1), (np.float32, 3), (np.int64, 2)):
        got, want = run(torch, D, g, ref, N, dt, op, 90 + op)
        assert same(got, want), (np.dtype(dt).name, op)


def test_byte_k3_bundled_blocks(torch, D):
    """128^3 p=2: the plan's bundled blocks (rows of 4 lines side by side)
    carry the codes the three-field kernel reads."""
    g, ref, N = mesh(D, (128, 128, 128, 2), fields=D.PlanK1)
    info = g.plan.info()
    assert info["bundleTy"] > 0 and info["byteK3"] == 1, info
    got, want = run(torch, D, g, ref, N, np.float64, O.ADD, 23)
    assert same(got, want)


@pytest.mark.parametrize("unique,flag", [(True, 0.0), (False, 0.3)])
def test_byte_k3_signed(torch, D, unique, flag):
    """Signed operators (5% of the ids 0; flagged: 30% negative): Sym runs
    over the T list, bit-exact whichever kernel the plan's lists admit."""
    rng = np.random.default_rng(31)
    ids = O.make_ids(36, 32, 32, 3).copy()
    ids[rng.random(len(ids)) < 0.05] = 0
    if flag:
        neg = rng.random(len(ids)) < flag
        ids[neg] = -ids[neg]
    o = O.OracleOgs(ids, O.SIGNED, unique)
    g = D.ogs_t()
    g.Setup(len(ids), ids.copy(), D.Signed, D.Auto, unique, False)
    info = g.plan.info()
    print("signed", unique, flag, {k: info[k] for k in ("nblocksK1", "plainK1", "deltaE", "deltaFallback",
                                                         "byteItems", "byteFallback", "byteK3", "maxRowT")})
    for op in (0, 3):
        v = values(rng, np.float64, K * len(ids))
        want = o.op_gather_scatter(v.copy(), K, op, O.SYM)
        d = torch.from_numpy(v.copy()).cuda()
        g.GatherScatter(d, K, op, D.Sym)
        torch.cuda.synchronize()
        assert same(d.cpu().numpy(), want), (op, info)


def test_rows_over_8_keep_entry_major(torch, D):
    """Rows of up to 12 members: no three-field byte launch; still bit-exact."""
    ids = O.make_ids(16, 16, 16, 7)
    g0 = ids - 1
    merged = np.where((g0 // 2) % 3 == 0, (g0 // 2) * 2, g0) + 1
    g, ref, N = mesh(D, None, ids=merged, key="merged")
    info = g.plan.info()
    assert info["maxRowT"] > 8 and info["byteK3"] == 0, info
    got, want = run(torch, D, g, ref, N, np.float64, O.ADD, 41)
    assert same(got, want)


def run_gather(torch, D, g, ref, N, dt, op, seed, ty=None):
    rng = np.random.default_rng(seed)
    v = values(rng, dt, K * N) if ty is None else \
        (rng.uniform(-1, 1, K * N) * 10.0 ** rng.uniform(-3, 3, K * N)).astype(np.float32)
    dv = torch.from_numpy(v.copy()).cuda()
    for trans in (O.TRANS, O.NOTRANS):
        want = ref.gather(v, K, op, trans, type=ty)
        gq = torch.zeros(len(want), dtype=dv.dtype, device="cuda")
        g.plan.Gather(gq, dv, K, op, trans, type=(D.FloatC if ty == O.FLOAT_C else None))
        torch.cuda.synchronize()
        assert same(gq.cpu().numpy(), want), (np.dtype(dt).name, op, trans)


@pytest.mark.parametrize("box", BOXES, ids=lambda b: "x".join(map(str, b)))
@pytest.mark.parametrize("ty", list(TYPES))
@pytest.mark.parametrize("op", [0, 1, 2, 3])
def test_byte_gather_k3(torch, D, box, ty, op):
    """Gather Trans and NoTrans (the Unsigned operator's lists alias) with
    three fields, each row's results at its row."""
    g, ref, N = mesh(D, box)
    run_gather(torch, D, g, ref, N, TYPES[ty], op, hash((box, ty, op, "g3")) % 2 ** 32)


@pytest.mark.parametrize("op", [0, 2])
def test_byte_gather_k3_float_c(torch, D, op):
    g, ref, N = mesh(D, (16, 16, 16, 7))
    run_gather(torch, D, g, ref, N, np.float32, op, 60 + op, ty=O.FLOAT_C)


@pytest.mark.parametrize("fields", [0, 1], ids=["all_k", "k1_plan"])
def test_byte_gather_k3_fallback_blocks(torch, D, fields):
    ids = shuffled((36, 32, 32, 3), 0.08, 3)
    g, ref, N = mesh(D, None, fields=fields, ids=ids, key=("shuffled", fields))
    info = g.plan.info()
    if info["plainK1"] > 0:
        pytest.skip(f"blocks left in CSR order: the entry-major kernel serves ({info})")
    for dt, op in ((np.float64, 0), (np.int64, 3)):
        run_gather(torch, D, g, ref, N, dt, op, 70 + op)


def test_byte_gather_k3_bundled_plan(torch, D):
    """A plan with bundled GatherScatter blocks gathers over partition 0."""
    g, ref, N = mesh(D, (128, 128, 128, 2), fields=D.PlanK1)
    assert g.plan.info()["bundleTy"] > 0
    run_gather(torch, D, g, ref, N, np.float64, O.ADD, 80)


@pytest.mark.parametrize("k", [2, 4])
@pytest.mark.parametrize("ty", list(TYPES))
@pytest.mark.parametrize("op", [0, 3])
def test_byte_k2_k4(torch, D, k, ty, op):
    """Two and four fields over the same codes (four fields of 8-byte values
    on the entry-major kernels): GatherScatter Sym and Gather Trans,
    bit-exact."""
    g, ref, N = mesh(D, (16, 16, 16, 7))
    dt = TYPES[ty]
    rng = np.random.default_rng(hash((k, ty, op, "k24")) % 2 ** 32)
    v = values(rng, dt, k * N)
    want = ref.gather_scatter(v.copy(), k, op, O.SYM)
    d = torch.from_numpy(v.copy()).cuda()
    g.GatherScatter(d, k, op, D.Sym)
    torch.cuda.synchronize()
    assert same(d.cpu().numpy(), want), ("gs", k, np.dtype(dt).name, op)
    wg = ref.gather(v, k, op, O.TRANS)
    gq = torch.zeros(len(wg), dtype=d.dtype, device="cuda")
    g.plan.Gather(gq, torch.from_numpy(v.copy()).cuda(), k, op, D.Trans)
    torch.cuda.synchronize()
    assert same(gq.cpu().numpy(), wg), ("gather", k, np.dtype(dt).name, op)


@pytest.mark.parametrize("k", [2, 4])
def test_byte_k2_k4_float_c(torch, D, k):
    g, ref, N = mesh(D, (16, 16, 16, 7))
    rng = np.random.default_rng(200 + k)
    v = (rng.uniform(-1, 1, k * N) * 10.0 ** rng.uniform(-3, 3, k * N)).astype(np.float32)
    want = ref.gather_scatter(v.copy(), k, O.ADD, O.SYM, type=O.FLOAT_C)
    d = torch.from_numpy(v.copy()).cuda()
    g.GatherScatter(d, k, O.ADD, D.Sym, type=D.FloatC)
    torch.cuda.synchronize()
    assert same(d.cpu().numpy(), want)


# --------------------------- Scatter, three fields by columns at low degree
@pytest.mark.parametrize("box", [(104, 100, 96, 1), (80, 64, 64, 2)], ids=["104x100x96_p1", "80x64x64_p2"])
@pytest.mark.parametrize("dt", [np.float64, np.int64, np.float32])
def test_scatter_k3_low_degree(torch, D, box, dt):
    """Scatter with three fields on lists of at least 3 entries per row (by
    columns for fields of up to 24 bytes): NoTrans, and Trans of the
    Unsigned operator, bit-exact."""
    g, ref, N = mesh(D, box)
    rng = np.random.default_rng(hash((box, np.dtype(dt).name, "s3")) % 2 ** 32)
    for trans in (O.NOTRANS, O.TRANS):
        gv = values(rng, dt, K * ref.op.NrowsT)
        base = values(rng, dt, K * N)
        want = ref.scatter(base.copy(), gv, K, trans)
        d = torch.from_numpy(base.copy()).cuda()
        g.plan.Scatter(d, torch.from_numpy(gv.copy()).cuda(), K, trans)
        torch.cuda.synchronize()
        assert same(d.cpu().numpy(), want), (box, np.dtype(dt).name, trans)
