"""GPU parity of the K = 1 one-list GatherScatter over the plan's 16-bit
delta pack (k_gs_sym_delta) and its one-byte codes (k_gs_sym_byte):
bit-exact against the oracle on meshes large enough for the pack
(partition 0 fills the device), for every type and op, Unsigned and Signed
(Sym), with fallback blocks (columns without short runs) mixed in, with the
pack dropped when most blocks would fall back, and with rows of more than 8
members (which the 8-position group fold does not take).
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TYPES = {0: np.float32, 1: np.float64, 2: np.int32, 3: np.int64}


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
    """Uniform values with 1% signed zeros and NaN among them (no
    infinities: a NaN made by inf - inf or inf * 0 has a sign that differs
    between the host and the GPU, while an input NaN passes through both)."""
    if np.issubdtype(dt, np.floating):
        v = rng.uniform(-1, 1, n).astype(dt)
        at = rng.random(n) < 0.01
        v[at] = np.array([0.0, -0.0, np.nan], dtype=dt)[rng.integers(0, 3, int(at.sum()))]
        return v
    return rng.integers(-2 ** 20, 2 ** 20 + 1, n).astype(dt)


def setup(D, ids, kind=O.UNSIGNED, unique=False):
    g = D.ogs_t()
    d = ids.copy()
    g.Setup(len(d), d, kind, D.Auto, unique, False)
    return g, d


BOXES = [(16, 16, 16, 7), (36, 32, 32, 3), (22, 22, 22, 5), (104, 100, 96, 1)]
_CACHE = {}


def unsigned_mesh(D, box):
    if box not in _CACHE:
        ids = O.make_ids(*box)
        g, _ = setup(D, ids)
        c = g.csr()
        _CACHE[box] = (g, O.OracleCsr(c["rowStartsT"], c["colIdsT"], Ncols=len(ids)), len(ids))
    return _CACHE[box]


@pytest.mark.parametrize("box", BOXES, ids=lambda b: "x".join(map(str, b)))
def test_delta_pack_built(D, box):
    g, _, N = unsigned_mesh(D, box)
    info = g.plan.info()
    assert info["deltaE"] > 0 and info["deltaFallback"] <= info["nblocksK1"] // 8, info
    assert info["nblocksK1"] >= 4 * 256  # partition 0 fills the device: the K = 1 launch uses it
    assert 0 < info["bytes"]


@pytest.mark.parametrize("box", BOXES, ids=lambda b: "x".join(map(str, b)))
@pytest.mark.parametrize("ty", list(TYPES))
@pytest.mark.parametrize("op", [0, 1, 2, 3])
def test_delta_gather_scatter(torch, D, box, ty, op):
    g, ref, N = unsigned_mesh(D, box)
    dt = TYPES[ty]
    rng = np.random.default_rng(hash((box, ty, op)) % 2 ** 32)
    v = values(rng, dt, N)
    want = ref.gather_scatter(v.copy(), 1, op, O.SYM)
    d = torch.from_numpy(v.copy()).cuda()
    g.GatherScatter(d, 1, op, D.Sym)
    torch.cuda.synchronize()
    assert same(d.cpu().numpy(), want), (box, np.dtype(dt).name, op)


def test_byte_codes_built(D):
    """48^3-like p=7 numbering (16^3 elements here): the one-byte codes exist
    over the delta pack's blocks, most blocks encoded (48^3: 116 of 27,743
    fall back)."""
    g, _, _ = unsigned_mesh(D, (16, 16, 16, 7))
    info = g.plan.info()
    assert info["maxRowT"] == 8 and info["byteItems"] > 0, info
    assert info["byteFallback"] <= info["nblocksK1"] // 16, info


@pytest.mark.parametrize("ty", list(TYPES))
@pytest.mark.parametrize("op", [0, 2])
def test_delta_rows_over_8(torch, D, ty, op):
    """Ids of a 16^3 p=7 mesh with every third pair of global ids merged
    (rows of up to 12 members): the kernels fold at the heads of the
    entries instead of in groups of 8; bit-exact."""
    ids = O.make_ids(16, 16, 16, 7)
    g0 = ids - 1
    merged = np.where((g0 // 2) % 3 == 0, (g0 // 2) * 2, g0)
    g, _ = setup(D, merged + 1)
    info = g.plan.info()
    assert info["maxRowT"] > 8 and info["deltaE"] > 0, info
    c = g.csr()
    ref = O.OracleCsr(c["rowStartsT"], c["colIdsT"], Ncols=len(ids))
    dt = TYPES[ty]
    rng = np.random.default_rng(1000 + 10 * ty + op)
    v = values(rng, dt, len(ids))
    want = ref.gather_scatter(v.copy(), 1, op, O.SYM)
    d = torch.from_numpy(v.copy()).cuda()
    g.GatherScatter(d, 1, op, D.Sym)
    torch.cuda.synchronize()
    assert same(d.cpu().numpy(), want), (np.dtype(dt).name, op)


@pytest.mark.parametrize("ty,op", [(O.FLOAT_C, 0), (O.BF16, 0), (O.BF16, 2), (O.FLOAT_C, 3)])
def test_delta_precision_types(torch, D, ty, op):
    box = BOXES[0]
    g, ref, N = unsigned_mesh(D, box)
    rng = np.random.default_rng(5 + ty + op)
    f = (rng.uniform(-1, 1, N) * 10.0 ** rng.uniform(-3, 3, N)).astype(np.float32)
    if ty == O.BF16:
        v = O.float_to_bf16(f)
        d = torch.from_numpy(v.view(np.int16).copy()).cuda().view(torch.bfloat16)
    else:
        v = f
        d = torch.from_numpy(v.copy()).cuda()
    want = ref.gather_scatter(v.copy(), 1, op, O.SYM, type=ty)
    g.GatherScatter(d, 1, op, D.Sym, type=(D.FloatC if ty == O.FLOAT_C else None))
    torch.cuda.synchronize()
    got = d.view(torch.int16).cpu().numpy().view(np.uint16) if ty == O.BF16 else d.cpu().numpy()
    assert same(got, want)


@pytest.mark.parametrize("unique,flag", [(True, 0.0), (False, 0.3)])
def test_delta_signed_sym(torch, D, unique, flag):
    box = BOXES[1]
    rng = np.random.default_rng(11)
    ids = O.make_ids(*box).copy()
    ids[rng.random(len(ids)) < 0.05] = 0
    if flag:
        neg = rng.random(len(ids)) < flag
        ids[neg] = -ids[neg]
    o = O.OracleOgs(ids, O.SIGNED, unique)
    g, _ = setup(D, ids, O.SIGNED, unique)
    assert g.plan.info()["deltaE"] > 0
    for op in (0, 2):
        v = values(rng, np.float64, len(ids))
        want = o.op_gather_scatter(v.copy(), 1, op, O.SYM)
        d = torch.from_numpy(v.copy()).cuda()
        g.GatherScatter(d, 1, op, D.Sym)
        torch.cuda.synchronize()
        assert same(d.cpu().numpy(), want), op


def shuffled(box, frac, seed):
    """Structured ids whose last `frac` of local positions are shuffled among
    themselves: the blocks of those rows hold columns without short runs."""
    ids = O.make_ids(*box).copy()
    n = len(ids)
    a = int(n * (1 - frac))
    rng = np.random.default_rng(seed)
    ids[a:] = ids[a:][rng.permutation(n - a)]
    return ids


def test_delta_fallback_blocks(torch, D):
    ids = shuffled((36, 32, 32, 3), 0.08, 3)
    g, _ = setup(D, ids)
    info = g.plan.info()
    assert info["deltaE"] > 0 and 0 < info["deltaFallback"] < info["nblocksK1"] // 4, info
    c = g.csr()
    ref = O.OracleCsr(c["rowStartsT"], c["colIdsT"], Ncols=len(ids))
    rng = np.random.default_rng(4)
    for dt, op in ((np.float64, 0), (np.int32, 1), (np.float32, 3)):
        v = values(rng, dt, len(ids))
        want = ref.gather_scatter(v.copy(), 1, op, O.SYM)
        d = torch.from_numpy(v.copy()).cuda()
        g.GatherScatter(d, 1, op, D.Sym)
        torch.cuda.synchronize()
        assert same(d.cpu().numpy(), want), (np.dtype(dt).name, op)


def test_delta_dropped_when_mostly_fallback(torch, D):
    ids = shuffled((36, 32, 32, 3), 1.0, 9)
    g, _ = setup(D, ids)
    assert g.plan.info()["deltaE"] == 0
    c = g.csr()
    ref = O.OracleCsr(c["rowStartsT"], c["colIdsT"], Ncols=len(ids))
    v = np.random.default_rng(8).uniform(-1, 1, len(ids))
    want = ref.gather_scatter(v.copy(), 1, O.ADD, O.SYM)
    d = torch.from_numpy(v.copy()).cuda()
    g.GatherScatter(d, 1, D.Add, D.Sym)
    torch.cuda.synchronize()
    assert same(d.cpu().numpy(), want)


@pytest.mark.parametrize("box", BOXES, ids=lambda b: "x".join(map(str, b)))
@pytest.mark.parametrize("ty", list(TYPES))
@pytest.mark.parametrize("op", [0, 1, 2, 3])
def test_delta_gather(torch, D, box, ty, op):
    """Gather K = 1 (Trans, and NoTrans of the Unsigned operator) over the
    delta pack, each row's result at its row."""
    g, ref, N = unsigned_mesh(D, box)
    dt = TYPES[ty]
    rng = np.random.default_rng(hash((box, ty, op, "g")) % 2 ** 32)
    v = values(rng, dt, N)
    dv = torch.from_numpy(v.copy()).cuda()
    for trans in (O.TRANS, O.NOTRANS):
        want = ref.gather(v, 1, op, trans)
        gq = torch.zeros(len(want), dtype=dv.dtype, device="cuda")
        g.plan.Gather(gq, dv, 1, op, trans)
        torch.cuda.synchronize()
        assert same(gq.cpu().numpy(), want), (box, np.dtype(dt).name, op, trans)


def test_delta_gather_fallback_blocks(torch, D):
    ids = shuffled((36, 32, 32, 3), 0.08, 3)
    g, _ = setup(D, ids)
    assert g.plan.info()["deltaFallback"] > 0
    c = g.csr()
    ref = O.OracleCsr(c["rowStartsT"], c["colIdsT"], Ncols=len(ids))
    v = values(np.random.default_rng(6), np.float64, len(ids))
    want = ref.gather(v, 1, O.ADD, O.TRANS)
    gq = torch.zeros(len(want), dtype=torch.float64, device="cuda")
    g.plan.Gather(gq, torch.from_numpy(v).cuda(), 1, D.Add, D.Trans)
    torch.cuda.synchronize()
    assert same(gq.cpu().numpy(), want)


def test_k1_plan_48cubed_p7(torch, D):
    """A plan prepared for K = 1 only (OGS_PLAN_K1) at the headline size: at
    most 0.5 GB of plan memory beside the CSR (the full plan holds ~2 GB),
    GatherScatter / Gather / Scatter at K = 1 bit-exact, and K = 3 still
    runs (from the CSR) bit-exact."""
    ids = O.make_ids(48, 48, 48, 7)
    N = len(ids)
    g = D.ogs_t()
    g.Setup(N, ids.copy(), D.Unsigned, D.Auto, False, False, fields=D.PlanK1)
    info = g.plan.info()
    assert info["bytes"] <= 500_000_000, info
    assert info["deltaE"] > 0 and info["deltaFallback"] == 0 and info["packedK1"] == 0, info
    c = g.csr()
    ref = O.OracleCsr(c["rowStartsT"], c["colIdsT"], Ncols=N)
    rng = np.random.default_rng(71)
    v = rng.uniform(-1, 1, N)
    d = torch.from_numpy(v.copy()).cuda()
    g.GatherScatter(d, 1, D.Add, D.Sym)
    torch.cuda.synchronize()
    assert same(d.cpu().numpy(), ref.gather_scatter(v.copy(), 1, O.ADD, O.SYM))
    wg = ref.gather(v, 1, O.MAX, O.TRANS)
    gq = torch.zeros(len(wg), dtype=torch.float64, device="cuda")
    g.plan.Gather(gq, torch.from_numpy(v).cuda(), 1, D.Max, D.Trans)
    torch.cuda.synchronize()
    assert same(gq.cpu().numpy(), wg)
    base = rng.uniform(-1, 1, N)
    d = torch.from_numpy(base.copy()).cuda()
    g.plan.Scatter(d, gq, 1, D.NoTrans)
    torch.cuda.synchronize()
    assert same(d.cpu().numpy(), ref.scatter(base.copy(), wg, 1, O.NOTRANS))
    del gq, d
    x3 = rng.uniform(-1, 1, 3 * N)
    d3 = torch.from_numpy(x3.copy()).cuda()
    g.GatherScatter(d3, 3, D.Add, D.Sym)
    torch.cuda.synchronize()
    assert same(d3.cpu().numpy(), ref.gather_scatter(x3.copy(), 3, O.ADD, O.SYM))


@pytest.mark.parametrize("unique,flag", [(True, 0.0), (False, 0.3)])
def test_k1_plan_signed(torch, D, unique, flag):
    """OGS_PLAN_K1 on Signed operators keeps the packs the two-list kernels
    read; every K = 1 call stays bit-exact."""
    rng = np.random.default_rng(21)
    ids = O.make_ids(36, 32, 32, 3).copy()
    if flag:
        neg = rng.random(len(ids)) < flag
        ids[neg] = -ids[neg]
    o = O.OracleOgs(ids, O.SIGNED, unique)
    g = D.ogs_t()
    g.Setup(len(ids), ids.copy(), D.Signed, D.Auto, unique, False, fields=D.PlanK1)
    v = rng.uniform(-1, 1, len(ids))
    for trans in (O.SYM, O.NOTRANS, O.TRANS):
        d = torch.from_numpy(v.copy()).cuda()
        g.GatherScatter(d, 1, D.Add, trans)
        torch.cuda.synchronize()
        assert same(d.cpu().numpy(), o.op_gather_scatter(v.copy(), 1, O.ADD, trans)), trans


def test_delta_fallback_blocks_k1_plan(torch, D):
    """A K = 1 plan keeps only its fallback blocks' 32-bit words (each
    fallback block finds its words by the slot in its exception words):
    GatherScatter and Gather bit-exact with fallback blocks mixed in."""
    ids = shuffled((36, 32, 32, 3), 0.08, 3)
    g = D.ogs_t()
    g.Setup(len(ids), ids.copy(), D.Unsigned, D.Auto, False, False, fields=D.PlanK1)
    info = g.plan.info()
    assert info["deltaE"] > 0 and 0 < info["deltaFallback"] < info["nblocksK1"] // 4, info
    assert info["packedK1"] == 0, info  # partition 0's full 32-bit pack is gone
    c = g.csr()
    ref = O.OracleCsr(c["rowStartsT"], c["colIdsT"], Ncols=len(ids))
    rng = np.random.default_rng(12)
    for dt, op in ((np.float64, 0), (np.int64, 2), (np.float32, 1)):
        v = values(rng, dt, len(ids))
        want = ref.gather_scatter(v.copy(), 1, op, O.SYM)
        d = torch.from_numpy(v.copy()).cuda()
        g.GatherScatter(d, 1, op, D.Sym)
        torch.cuda.synchronize()
        assert same(d.cpu().numpy(), want), (np.dtype(dt).name, op)
        wg = ref.gather(v, 1, op, O.TRANS)
        gq = torch.zeros(len(wg), dtype=d.dtype, device="cuda")
        g.plan.Gather(gq, torch.from_numpy(v.copy()).cuda(), 1, op, D.Trans)
        torch.cuda.synchronize()
        assert same(gq.cpu().numpy(), wg), ("gather", np.dtype(dt).name, op)


@pytest.mark.parametrize("box,fields", [((192, 192, 192, 1), 1), ((200, 150, 120, 1), 0), ((128, 128, 128, 2), 1)],
                         ids=["192cubed_p1_k1plan", "200x150x120_p1", "128cubed_p2_k1plan"])
def test_bundled_blocks(torch, D, box, fields):
    """Low degree: the plan finds the mesh's line and layer row strides in
    the T list (first-encounter order), builds GatherScatter blocks of 4
    lines x 2 or 1 layers and keeps them (they race faster than partition 0
    at these sizes).  Bit-exact against the oracle for every type and op:
    each row is still reduced in its own CSR order, whichever block holds
    it."""
    ids = O.make_ids(*box)
    N = len(ids)
    g = D.ogs_t()
    g.Setup(N, ids.copy(), D.Unsigned, D.Auto, False, False, fields=fields)
    info = g.plan.info()
    nx, ny, p = box[0], box[1], box[3]
    # rows each element introduces in first-encounter order: local indices
    # >= 1 along every axis where a neighbour came before, so an element
    # line holds (p+1) p^2 + (nx-1) p^3 rows and a layer adds the first
    # line's (p+1)^2 p + (nx-1) (p+1) p^2
    Ry = (p + 1) * p ** 2 + (nx - 1) * p ** 3
    Rz = (p + 1) ** 2 * p + (nx - 1) * (p + 1) * p ** 2 + (ny - 1) * Ry
    assert (info["bundleRy"], info["bundleRz"]) == (Ry, Rz), info
    assert info["bundleTy"] == 4 and info["bundleTz"] in (1, 2), info
    c = g.csr()
    ref = O.OracleCsr(c["rowStartsT"], c["colIdsT"], Ncols=N)
    del ids
    rng = np.random.default_rng(sum(box))
    for dt, op in ((np.float64, 0), (np.float32, 3), (np.int32, 1), (np.int64, 2)):
        v = values(rng, dt, N)
        want = ref.gather_scatter(v.copy(), 1, op, O.SYM)
        d = torch.from_numpy(v.copy()).cuda()
        g.GatherScatter(d, 1, op, D.Sym)
        torch.cuda.synchronize()
        assert same(d.cpu().numpy(), want), (box, np.dtype(dt).name, op)
        # the Gather on the same plan (partition 0, byte codes, rows of at
        # most 8 members folded per group of 8 positions)
        wg = ref.gather(v, 1, op, O.TRANS)
        gq = torch.zeros(len(wg), dtype=torch.from_numpy(v[:1]).dtype, device="cuda")
        g.plan.Gather(gq, torch.from_numpy(v).cuda(), 1, op, D.Trans)
        torch.cuda.synchronize()
        assert same(gq.cpu().numpy(), wg), (box, np.dtype(dt).name, op)
        del d, want, gq, wg


def test_bundled_blocks_signed_operator(torch, D):
    """A Signed-unique operator at 192^3 p=1 (the sweep's sogs): its T list
    gets the same bundled blocks; GatherScatter Sym runs over them bit-exact,
    NoTrans and Trans (two lists, other kernels) stay bit-exact."""
    ids = O.make_ids(192, 192, 192, 1)
    N = len(ids)
    g = D.ogs_t()
    sid = ids.copy()
    g.Setup(N, sid, D.Signed, D.Auto, True, False)
    del ids
    info = g.plan.info()
    assert (info["bundleRy"], info["bundleRz"], info["bundleTy"]) == (193, 193 * 193, 4), info
    c = g.csr()
    ref = O.OracleCsr(c["rowStartsT"], c["colIdsT"], c["rowStartsN"], c["colIdsN"], g.NrowsN, N)
    rng = np.random.default_rng(193)
    v = values(rng, np.float64, N)
    for trans in (O.SYM, O.NOTRANS, O.TRANS):
        want = ref.gather_scatter(v.copy(), 1, O.ADD, trans)
        d = torch.from_numpy(v.copy()).cuda()
        g.GatherScatter(d, 1, D.Add, trans)
        torch.cuda.synchronize()
        assert same(d.cpu().numpy(), want), trans


@pytest.mark.parametrize("frac", [0.0, 0.08, 1.0], ids=["structured", "partly_shuffled", "shuffled"])
def test_row_map_scatter(torch, D, frac):
    """Scatter K = 1 by columns over the column -> row map in 4-bit row-delta
    codes (lists of fewer than two entries per row, here p = 7): built on a
    structured list, with fallback chunks (reading the int32 map) where part
    of the numbering is shuffled, dropped where all of it is; bit-exact
    against the oracle in every case, for 8- and 4-byte values, NoTrans and
    Trans of an Unsigned operator."""
    ids = shuffled((36, 36, 30, 7), frac, 21) if frac else O.make_ids(36, 36, 30, 7)
    g, _ = setup(D, ids)
    info = g.plan.info()
    chunks = -(-len(ids) // 2048)
    if frac == 0.0:  # (element layers where many rows jump back hold the chunks that fall back: ~11%)
        assert info["rowMapE"] > 0 and info["rowMapFallback"] < chunks // 6, info
    elif frac < 1.0:
        assert info["rowMapE"] > 0 and chunks // 16 < info["rowMapFallback"] <= chunks // 4, info
    else:
        assert info["rowMapE"] == 0, info
    c = g.csr()
    ref = O.OracleCsr(c["rowStartsT"], c["colIdsT"], Ncols=len(ids))
    rng = np.random.default_rng(22)
    for dt in (np.float64, np.int32):
        gv = values(rng, dt, g.NrowsT)
        base = values(rng, dt, len(ids))
        for trans in (O.NOTRANS, O.TRANS):
            want = ref.scatter(base.copy(), gv, 1, trans)
            d = torch.from_numpy(base.copy()).cuda()
            g.plan.Scatter(d, torch.from_numpy(gv.copy()).cuda(), 1, trans)
            torch.cuda.synchronize()
            assert same(d.cpu().numpy(), want), (np.dtype(dt).name, trans)


def test_row_map_only_for_short_rows(D):
    """Lists of two or more entries per row (p = 3) keep the int32 map."""
    g, _ = setup(D, O.make_ids(36, 32, 32, 3))
    assert g.plan.info()["rowMapE"] == 0
