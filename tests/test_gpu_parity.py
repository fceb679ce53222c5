"""GPU parity: the gfx950 path (through the C-ABI of libogs_hip.so) against
the pinned CPU oracle, bit-exact, on the same seeded inputs.

Covers every {type x op x trans x K} of GatherScatter / Gather / Scatter on
Unsigned, Signed (ragged, flagged) and Signed-unique setups, the reference's
own "-cc" check (gslib, Error = 0), the 24^3 p=3 configuration, the full
48^3 p=7 headline size, extract, hand-built operators with long rows,
Scatter by columns (dense lists) and by rows (sparse lists), and the error
contract.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TYPES = {0: np.float32, 1: np.float64, 2: np.int32, 3: np.int64}
OPS = [0, 1, 2, 3]       # Add Mul Max Min
TRANS = [0, 1, 2]        # Sym NoTrans Trans


@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available()
    return torch


@pytest.fixture(scope="module")
def D():
    import dawgs_amd
    return dawgs_amd


def to_dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def to_host(torch, t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def values(rng, dt, n):
    if np.issubdtype(dt, np.floating):
        return rng.uniform(-1, 1, n).astype(dt)
    return rng.integers(-2 ** 20, 2 ** 20 + 1, n).astype(dt)


def same(a, b):
    """Bit-exact equality (NaN payloads and signed zeros included)."""
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype != b.dtype or a.shape != b.shape:
        return False
    return a.tobytes() == b.tobytes()


MESHES = {}


def mesh(kind):
    """Small ragged meshes, built once: (ids after setup, oracle, ogs_t)."""
    if kind in MESHES:
        return MESHES[kind]
    import dawgs_amd as D
    rng = np.random.default_rng(7)
    ids = O.make_ids(6, 5, 4, 3).copy()
    ids[rng.random(len(ids)) < 0.08] = 0
    if kind == "unsigned":
        k, u = O.UNSIGNED, False
    elif kind == "signed":
        neg = rng.random(len(ids)) < 0.35
        ids[neg] = -ids[neg]
        k, u = O.SIGNED, False
    else:
        k, u = O.SIGNED, True
    o = O.OracleOgs(ids, k, u)
    g = D.ogs_t()
    dids = ids.copy()
    g.Setup(len(dids), dids, k, D.Auto, u, False)
    MESHES[kind] = (dids, o, g)
    return MESHES[kind]


# ------------------------------------------------------------------ setup
@pytest.mark.parametrize("kind", ["unsigned", "signed", "unique"])
def test_setup_matches_oracle(kind):
    ids, o, g = mesh(kind)
    assert np.array_equal(ids, o.ids)  # unique: same signs written back
    assert (g.N, g.Ngather, g.NgatherGlobal, g.NlocalP, g.NlocalT, g.NhaloP, g.NhaloT) == \
           (o.N, o.Ngather, o.NgatherGlobal, o.NlocalP, o.NlocalT, o.NhaloP, o.NhaloT)
    assert g.gather_defined == o.gather_defined
    c = g.csr()
    for name in ("rowStartsT", "colIdsT", "rowStartsN", "colIdsN", "blockRowStartsT", "blockRowStartsN"):
        assert np.array_equal(c[name], getattr(o, name)), name


@pytest.mark.parametrize("box,kind,unique", [((10, 10, 10, 3), O.UNSIGNED, False),
                                             ((10, 10, 10, 3), O.SIGNED, True),
                                             ((24, 24, 24, 3), O.UNSIGNED, False)])
def test_setup_reference_configs(D, box, kind, unique):
    ids = O.make_ids(*box)
    o = O.OracleOgs(ids, kind, unique)
    dids = ids.copy()
    g = D.ogs_t()
    g.Setup(len(dids), dids, kind, D.Auto, unique, False)
    c = g.csr()
    assert g.NgatherGlobal == o.NgatherGlobal
    assert np.array_equal(c["rowStartsT"], o.rowStartsT)
    assert np.array_equal(c["colIdsT"], o.colIdsT)
    assert np.array_equal(c["colIdsN"], o.colIdsN)
    assert g.NrowBlocksT == o.NrowBlocksT


# ------------------------------------------------------------------ operators
@pytest.mark.parametrize("K", [1, 3])
@pytest.mark.parametrize("trans", TRANS)
@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("ty", list(TYPES))
@pytest.mark.parametrize("kind", ["unsigned", "signed", "unique"])
def test_gather_scatter(torch, kind, ty, op, trans, K):
    ids, o, g = mesh(kind)
    rng = np.random.default_rng(1000 * ty + 10 * op + trans)
    v = values(rng, TYPES[ty], len(ids) * K)
    want = o.op_gather_scatter(v.copy(), K, op, trans)
    d = to_dev(torch, v)
    g.GatherScatter(d, K, op, trans)
    assert same(to_host(torch, d), want)


@pytest.mark.parametrize("K", [1, 3])
@pytest.mark.parametrize("trans", TRANS)
@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("ty", list(TYPES))
@pytest.mark.parametrize("kind", ["unsigned", "signed", "unique"])
def test_gather(torch, kind, ty, op, trans, K):
    ids, o, g = mesh(kind)
    rng = np.random.default_rng(2000 * ty + 10 * op + trans)
    v = values(rng, TYPES[ty], len(ids) * K)
    want = o.op_gather(v, K, op, trans)
    gv = torch.full((max(o.NlocalT, 1) * K,), 77, dtype=to_dev(torch, v[:1]).dtype, device="cuda")
    g.plan.Gather(gv, to_dev(torch, v), K, op, trans)
    got = to_host(torch, gv)
    assert same(got[: len(want)], want)
    assert np.all(got[len(want):] == 77)  # rows beyond the op's row count untouched


@pytest.mark.parametrize("K", [1, 3])
@pytest.mark.parametrize("trans", TRANS)
@pytest.mark.parametrize("ty", list(TYPES))
@pytest.mark.parametrize("kind", ["unsigned", "signed", "unique"])
def test_scatter(torch, kind, ty, trans, K):
    ids, o, g = mesh(kind)
    rng = np.random.default_rng(3000 * ty + trans)
    gvals = values(rng, TYPES[ty], o.NlocalT * K)
    base = values(rng, TYPES[ty], len(ids) * K)
    want = o.op_scatter(base.copy(), gvals, K, trans)
    d = to_dev(torch, base)
    g.plan.Scatter(d, to_dev(torch, gvals), K, trans)
    assert same(to_host(torch, d), want)


@pytest.mark.parametrize("ty", list(TYPES))
def test_ogs_level_gather_scatter_unique(torch, D, ty):
    """sogs Gather(Trans) then Scatter(NoTrans), as dawgs times them."""
    ids, o, g = mesh("unique")
    rng = np.random.default_rng(40 + ty)
    v = values(rng, TYPES[ty], len(ids))
    gv = torch.zeros(g.Ngather, dtype=to_dev(torch, v[:1]).dtype, device="cuda")
    g.Gather(gv, to_dev(torch, v), 1, D.Add, D.Trans)
    assert same(to_host(torch, gv), o.gather(v, 1, O.ADD, O.TRANS))
    out = to_dev(torch, np.zeros_like(v))
    g.Scatter(out, gv, 1, D.NoTrans)
    assert same(to_host(torch, out), o.scatter(np.zeros_like(v), o.gather(v, 1, O.ADD, O.TRANS), 1, O.NOTRANS))


@pytest.mark.parametrize("K", [8, 16, 100])
@pytest.mark.parametrize("kind", ["signed", "unique"])
def test_many_fields(torch, K, kind):
    """K large enough to need field chunks (K=100 runs in passes); on the
    unique setup NoTrans/Trans run through the column -> owner map."""
    ids, o, g = mesh(kind)
    rng = np.random.default_rng(K)
    v = rng.uniform(-1, 1, len(ids) * K)
    for trans in TRANS:
        want = o.op_gather_scatter(v.copy(), K, O.ADD, trans)
        d = to_dev(torch, v)
        g.plan.GatherScatter(d, K, O.ADD, trans)
        assert same(to_host(torch, d), want), trans
        gv = torch.zeros(o.NlocalT * K, dtype=torch.float64, device="cuda")
        g.plan.Gather(gv, to_dev(torch, v), K, O.MAX, trans)
        wg = o.op_gather(v, K, O.MAX, trans)
        assert same(to_host(torch, gv)[: len(wg)], wg)
        s = to_dev(torch, v)
        g.plan.Scatter(s, to_dev(torch, v[: o.NlocalT * K]), K, trans)
        assert same(to_host(torch, s), o.op_scatter(v.copy(), v[: o.NlocalT * K], K, trans))


@pytest.mark.parametrize("K", [2, 4, 5, 12])
@pytest.mark.parametrize("kind", ["signed", "unique"])
def test_signed_gather_scatter_other_field_counts(torch, kind, K):
    """GatherScatter NoTrans/Trans of the Signed setups at field counts the
    other tests skip: compiled shapes (2, 4) and runtime loops (5, 12) of the
    by-column owner kernel (unique) and of the column-mask Trans kernel and
    the two-list kernels (both setups)."""
    ids, o, g = mesh(kind)
    rng = np.random.default_rng(70 + K)
    for dt, op in ((np.float64, O.ADD), (np.int32, O.MAX), (np.float32, O.MIN)):
        v = values(rng, dt, len(ids) * K)
        for trans in (O.NOTRANS, O.TRANS):
            want = o.op_gather_scatter(v.copy(), K, op, trans)
            d = to_dev(torch, v)
            g.plan.GatherScatter(d, K, op, trans)
            assert same(to_host(torch, d), want), (np.dtype(dt).name, op, trans)


# ------------------------------------------------------------------ reference checks
@pytest.mark.skipif(not O.ref_available(), reason="gslib comparator not built")
def test_reference_cc_error_zero(torch, D):
    """CorrectnessTest (dawgsMain.cpp:314-416), device rows: Error = 0."""
    ids = O.make_ids(10, 10, 10, 3)
    q = O.reference_q(0, len(ids))
    qcheck = O.GsRef(ids).gs_vec(q, 1)
    ogs = D.ogs_t()
    ogs.Setup(len(ids), ids.copy(), D.Unsigned, D.Pairwise, False, False)
    sids = ids.copy()
    sogs = D.ogs_t()
    sogs.Setup(len(sids), sids, D.Signed, D.Pairwise, True, False)
    d = to_dev(torch, q)
    ogs.GatherScatter(d, 1, D.Add, D.Sym)
    assert same(to_host(torch, d), qcheck)
    d = to_dev(torch, q)
    sogs.GatherScatter(d, 1, D.Add, D.Sym)
    assert same(to_host(torch, d), qcheck)
    d = to_dev(torch, q)
    gq = torch.zeros(sogs.Ngather, dtype=torch.float64, device="cuda")
    sogs.Gather(gq, d, 1, D.Add, D.Trans)
    sogs.Scatter(d, gq, 1, D.NoTrans)
    assert same(to_host(torch, d), qcheck)


def test_config_24cubed_p3(torch, D):
    """One MI355X, plain method: 24^3 p=3, fp64/fp32 Add, int32/int64
    Min/Max/Mul in [-2^20, 2^20], Sym/Trans/NoTrans, k=1, bit-exact."""
    ids = O.make_ids(24, 24, 24, 3)
    N = len(ids)
    o = O.OracleOgs(ids, O.UNSIGNED)
    so = O.OracleOgs(ids, O.SIGNED, True)
    ogs = D.ogs_t()
    ogs.Setup(N, ids.copy(), D.Unsigned, D.Auto, False, False)
    sids = ids.copy()
    sogs = D.ogs_t()
    sogs.Setup(N, sids, D.Signed, D.Auto, True, False)
    rng = np.random.default_rng(24)
    cases = [(np.float64, O.ADD), (np.float32, O.ADD)] + \
            [(dt, op) for dt in (np.int32, np.int64) for op in (O.MIN, O.MAX, O.MUL)]
    for dt, op in cases:
        v = values(rng, dt, N)
        for trans in TRANS:
            for ref, dev in ((o, ogs), (so, sogs)):
                want = ref.op_gather_scatter(v.copy(), 1, op, trans)
                d = to_dev(torch, v)
                dev.GatherScatter(d, 1, op, trans)
                assert same(to_host(torch, d), want), (np.dtype(dt).name, op, trans)
        # Gather Trans + Scatter NoTrans on the unique setup
        gq = torch.zeros(sogs.Ngather, dtype=to_dev(torch, v[:1]).dtype, device="cuda")
        sogs.Gather(gq, to_dev(torch, v), 1, op, D.Trans)
        assert same(to_host(torch, gq), so.gather(v, 1, op, O.TRANS))


def test_three_fields_on_whole_shape_blocks_32cubed_p3(torch, D):
    """K = 3 one-list GatherScatter at 32^3 p=3 (2.1M entries): large enough
    that every type runs the three-field partitions (blocks of up to 1,365
    entries for 4-byte values, 682 for 8-byte ones) rather than the
    power-of-two widths below them; bit-exact against the oracle for each
    type and op: GatherScatter Sym on the Unsigned and Signed-unique setups,
    NoTrans on the Unsigned one (its N list is its T list), Trans and NoTrans
    on the Signed-unique one (rows of the T list stored at their owners; each
    column from its owner), and the
    operator's Gather Trans and Scatter NoTrans (T list) on both."""
    ids = O.make_ids(32, 32, 32, 3)
    N = len(ids)
    o = O.OracleOgs(ids, O.UNSIGNED)
    so = O.OracleOgs(ids, O.SIGNED, True)
    ogs = D.ogs_t()
    ogs.Setup(N, ids.copy(), D.Unsigned, D.Auto, False, False)
    sogs = D.ogs_t()
    sogs.Setup(N, ids.copy(), D.Signed, D.Auto, True, False)
    rng = np.random.default_rng(323)
    cases = [(np.float64, O.ADD), (np.float32, O.ADD), (np.float64, O.MAX), (np.float32, O.MIN)] + \
            [(dt, op) for dt in (np.int32, np.int64) for op in (O.ADD, O.MIN, O.MAX, O.MUL)]
    for dt, op in cases:
        v = values(rng, dt, 3 * N)
        for ref, dev, trans in ((o, ogs, O.SYM), (so, sogs, O.SYM), (o, ogs, O.NOTRANS), (so, sogs, O.TRANS),
                                (so, sogs, O.NOTRANS)):
            want = ref.op_gather_scatter(v.copy(), 3, op, trans)
            d = to_dev(torch, v)
            dev.GatherScatter(d, 3, op, trans)
            assert same(to_host(torch, d), want), (np.dtype(dt).name, op, trans)
        for ref, dev in ((o, ogs), (so, sogs)):
            want = ref.op_gather(v, 3, op, O.TRANS)
            gq = torch.zeros(len(want), dtype=to_dev(torch, v[:1]).dtype, device="cuda")
            dev.plan.Gather(gq, to_dev(torch, v), 3, op, O.TRANS)
            assert same(to_host(torch, gq), want), (np.dtype(dt).name, op, "gather")
            base = values(rng, dt, 3 * N)
            d = to_dev(torch, base)
            dev.plan.Scatter(d, gq, 3, O.NOTRANS)
            assert same(to_host(torch, d), ref.op_scatter(base.copy(), want, 3, O.NOTRANS)), \
                (np.dtype(dt).name, op, "scatter")


def test_headline_48cubed_p7(torch, D):
    """fp64 Add Sym GatherScatter at 48^3 p=7 (56.6M DOFs): bit-exact against
    the oracle's host loop run on the same CSR, and the size-independent
    property GS(1) = multiplicity."""
    from dawgs_amd.mesh import make_ids, mesh_counts
    ids = make_ids(48, 48, 48, 7)
    N = len(ids)
    g = D.ogs_t()
    g.Setup(N, ids, D.Unsigned, D.Auto, False, False)
    assert g.NgatherGlobal == mesh_counts(48, 48, 48, 7)["NgatherGlobal"] == 38272753
    c = g.csr()
    del ids
    hist = np.bincount(np.diff(c["rowStartsT"]))
    assert {k: int(v) for k, v in enumerate(hist) if v} == {1: 24389000, 2: 11858100, 4: 1921830, 8: 103823}
    ref = O.OracleCsr(c["rowStartsT"], c["colIdsT"], Ncols=N)
    rng = np.random.default_rng(48)
    v = rng.uniform(-1, 1, N)
    want = ref.gather_scatter(v.copy(), 1, O.ADD, O.SYM)
    d = to_dev(torch, v)
    g.GatherScatter(d, 1, D.Add, D.Sym)
    assert same(to_host(torch, d), want)
    # plan-level Gather Trans and Scatter NoTrans at full size (the widest partitions)
    wg = ref.gather(v, 1, O.ADD, O.TRANS)
    gq = torch.zeros(len(wg), dtype=torch.float64, device="cuda")
    g.plan.Gather(gq, to_dev(torch, v), 1, O.ADD, O.TRANS)
    assert same(to_host(torch, gq), wg)
    base = rng.uniform(-1, 1, N)
    d = to_dev(torch, base)
    g.plan.Scatter(d, gq, 1, O.NOTRANS)
    assert same(to_host(torch, d), ref.scatter(base.copy(), wg, 1, O.NOTRANS))
    del gq, base
    ones = torch.ones(N, dtype=torch.float64, device="cuda")
    g.GatherScatter(ones, 1, D.Add, D.Sym)
    m = to_host(torch, ones)
    lens = np.diff(c["rowStartsT"]).astype(np.float64)
    expect = np.empty(N)
    expect[c["colIdsT"]] = np.repeat(lens, np.diff(c["rowStartsT"]))
    assert same(m, expect)


def test_tiled_block_order_192cubed_p1(torch, D):
    """Degree 1 at 192^3 (56.6M DOFs): rows sharing a cache line sit a mesh
    layer of blocks apart, the plan finds that distance from the list and
    picks the tiled block order (measured faster on scratch values).  The
    result is bit-exact against the oracle's host loop on the same CSR, and
    GS(1) = multiplicity."""
    from dawgs_amd.mesh import make_ids
    ids = make_ids(192, 192, 192, 1)
    N = len(ids)
    g = D.ogs_t()
    g.Setup(N, ids, D.Unsigned, D.Auto, False, False)
    del ids
    info = g.plan.info()
    assert 140 <= info["tileD"] <= 148, info  # 193^2 vertices per layer / ~259 rows per block
    c = g.csr()
    ref = O.OracleCsr(c["rowStartsT"], c["colIdsT"], Ncols=N)
    rng = np.random.default_rng(192)
    v = rng.uniform(-1, 1, N)
    want = ref.gather_scatter(v.copy(), 1, O.ADD, O.SYM)
    d = to_dev(torch, v)
    g.GatherScatter(d, 1, D.Add, D.Sym)
    assert same(to_host(torch, d), want)
    ones = torch.ones(N, dtype=torch.float64, device="cuda")
    g.GatherScatter(ones, 1, D.Add, D.Sym)
    expect = np.empty(N)
    expect[c["colIdsT"]] = np.repeat(np.diff(c["rowStartsT"]).astype(np.float64), np.diff(c["rowStartsT"]))
    assert same(to_host(torch, ones), expect)
    del ones, expect
    # plan-level Gather Trans and Scatter NoTrans (each with its own block-order choice)
    wg = ref.gather(v, 1, O.ADD, O.TRANS)
    gq = torch.zeros(len(wg), dtype=torch.float64, device="cuda")
    g.plan.Gather(gq, to_dev(torch, v), 1, O.ADD, O.TRANS)
    assert same(to_host(torch, gq), wg)
    base = rng.uniform(-1, 1, N)
    d = to_dev(torch, base)
    g.plan.Scatter(d, gq, 1, O.NOTRANS)
    assert same(to_host(torch, d), ref.scatter(base.copy(), wg, 1, O.NOTRANS))


# ------------------------------------------------------------------ edge cases
def test_empty_and_degenerate(torch, D):
    g = D.ogs_t()
    g.Setup(0, np.zeros(0, np.int64), D.Unsigned)
    assert g.Ngather == 0 and g.NgatherGlobal == 0
    d = torch.zeros(1, dtype=torch.float64, device="cuda")
    g.GatherScatter(d, 1, D.Add, D.Sym)
    z = np.zeros(100, np.int64)  # every node ignored
    g.Setup(100, z, D.Unsigned)
    d = to_dev(torch, np.arange(100, dtype=np.float64))
    g.GatherScatter(d, 1, D.Add, D.Sym)
    assert same(to_host(torch, d), np.arange(100, dtype=np.float64))
    one = np.array([5], np.int64)
    g.Setup(1, one, D.Unsigned)
    d = to_dev(torch, np.array([-0.0]))
    g.GatherScatter(d, 1, D.Add, D.Sym)
    r = to_host(torch, d)
    assert r[0] == 0.0 and not np.signbit(r[0])  # +0.0 start


def test_special_values(torch, D):
    ids = np.array([1, 2, 2, 3, 3, 3, 4, 4], np.int64)
    o = O.OracleOgs(ids, O.UNSIGNED)
    g = D.ogs_t()
    g.Setup(len(ids), ids.copy(), D.Unsigned)
    inf, nan = np.inf, np.nan
    v = np.array([-0.0, nan, 1.0, inf, -inf, 2.0, -0.0, 0.0])
    for dt in (np.float64, np.float32):
        for op in OPS:
            x = v.astype(dt)
            want = o.gather_scatter(x.copy(), 1, op, O.SYM)
            d = to_dev(torch, x)
            g.GatherScatter(d, 1, op, D.Sym)
            assert same(to_host(torch, d), want), (np.dtype(dt).name, op)
    for dt in (np.int32, np.int64):
        big = np.iinfo(dt).max
        x = np.array([big, big, 3, big, -big, 7, -1, big], dtype=dt)
        for op in OPS:
            want = o.gather_scatter(x.copy(), 1, op, O.SYM)
            d = to_dev(torch, x)
            g.GatherScatter(d, 1, op, D.Sym)
            assert same(to_host(torch, d), want), (np.dtype(dt).name, op)


def test_single_owner_notrans_special_values(torch, D):
    """GatherScatter NoTrans K = 1 on Signed operators whose N list holds at
    most one member per row (the by-column path: each column takes f of its
    row's owner value, f = the op from its identity): -0.0 becomes +0.0 under
    Add, NaN and infinities follow the host's compare-select, and a group
    with no positive id gets the identity -- on every column, whichever
    order the columns are written in."""
    ids = np.array([1, -1, -1, -2, -2, 3, -3, 4, -4, -4, -4, 5, 0, -6, 6, -7], np.int64)
    inf, nan = np.inf, np.nan
    fv = np.array([-0.0, 1.0, 2.0, 5.0, 6.0, nan, 3.0, inf, 0.0, 1.0, 2.0, -inf, 9.0, 4.0, -0.0, 8.0])
    for unique in (False, True):
        o = O.OracleOgs(ids.copy(), O.SIGNED, unique)
        g = D.ogs_t()
        dids = ids.copy()
        g.Setup(len(dids), dids, D.Signed, D.Auto, unique, False)
        info = g.plan.info()
        assert info["maxRowN"] <= 1 and info["colsT"] == 1, info
        for dt in (np.float64, np.float32, np.int32, np.int64):
            x = fv.astype(dt) if np.issubdtype(dt, np.floating) else \
                np.array([7, -3, 2, np.iinfo(dt).max, 5, -np.iinfo(dt).max, 1, 0, 4, 4, 9, -8, 6, 2, 3, 1], dt)
            for op in OPS:
                want = o.op_gather_scatter(x.copy(), 1, op, O.NOTRANS)
                d = to_dev(torch, x)
                g.plan.GatherScatter(d, 1, op, O.NOTRANS)
                assert same(to_host(torch, d), want), (unique, np.dtype(dt).name, op)


def test_row_limits_and_errors(torch, D):
    from dawgs_amd._lib import OgsError
    g = D.ogs_t()
    ids = np.full(512, 9, np.int64)
    g.Setup(512, ids, D.Unsigned)
    d = to_dev(torch, np.ones(512))
    g.GatherScatter(d, 1, D.Add, D.Sym)
    assert np.all(to_host(torch, d) == 512.0)
    with pytest.raises(OgsError) as e:
        g.Setup(513, np.full(513, 9, np.int64), D.Unsigned)
    assert e.value.code == -2
    with pytest.raises(OgsError) as e:
        g.Setup(8, np.arange(1, 9, dtype=np.int64), D.Unsigned, D.Auto, True)
    assert e.value.code == -4
    g.Setup(8, np.array([1, 1, 2, 2, 3, 3, 4, 4], np.int64), D.Unsigned)
    assert not g.gather_defined
    with pytest.raises(OgsError) as e:
        g.Gather(torch.zeros(4, dtype=torch.float64, device="cuda"), to_dev(torch, np.ones(8)), 1, D.Add, D.Trans)
    assert e.value.code == -3


def test_plan_rejects_entries_in_n_rows_beyond_nrowsn(torch, D):
    """Rows [NrowsN, NrowsT) of the N list must be empty (ogs_hip_plan_create's
    contract): a list with entries there is refused, for an aliased operator
    and for distinct lists."""
    from dawgs_amd._lib import OgsError
    E_CSR = -7
    rs = np.array([0, 2, 3, 5], np.int32)  # 3 rows, the last one non-empty
    ci = np.array([0, 4, 1, 2, 3], np.int32)
    d_rs, d_ci = to_dev(torch, rs), to_dev(torch, ci)
    with pytest.raises(OgsError) as e:
        D.Plan(5, 2, 3, d_rs, d_ci, d_rs, d_ci)
    assert e.value.code == E_CSR
    rsN = np.array([0, 1, 2, 3], np.int32)
    ciN = np.array([0, 1, 3], np.int32)
    with pytest.raises(OgsError) as e:
        D.Plan(5, 2, 3, to_dev(torch, rsN), to_dev(torch, ciN), d_rs, d_ci)
    assert e.value.code == E_CSR
    ok = D.Plan(5, 3, 3, d_rs, d_ci, d_rs, d_ci)  # all rows in N: accepted
    assert ok.info()["NrowsN"] == 3


def test_hand_built_operator_long_rows(torch, D):
    """Plan layer on a hand-built Signed operator (the halo pack/unpack shape):
    rows of 1..1500 members, distinct N and T lists, K = 1 and 3."""
    rng = np.random.default_rng(5)
    Ncols = 6000
    lens = np.array([1500, 1, 7, 300, 2, 1, 900, 40] + list(rng.integers(1, 9, 200)), dtype=np.int64)
    lens = lens[: np.searchsorted(np.cumsum(lens), Ncols)]
    nnzT = int(lens.sum())
    perm = rng.permutation(Ncols)[:nnzT].astype(np.int32)
    rsT = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ciT = np.concatenate([np.sort(perm[rsT[r]:rsT[r + 1]]) for r in range(len(lens))]).astype(np.int32)
    # N list: a subset of each row (first row positive ones only); last rows empty in N
    NrowsT = len(lens)
    NrowsN = NrowsT - 5
    keep = rng.random(nnzT) < 0.6
    rows = np.repeat(np.arange(NrowsT), lens)
    keep &= rows < NrowsN
    ciN = ciT[keep]
    rsN = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=NrowsT))]).astype(np.int32)
    ref = O.OracleCsr(rsT, ciT, rsN, ciN, NrowsN, Ncols)
    plan = D.Plan(Ncols, NrowsN, NrowsT, to_dev(torch, rsN), to_dev(torch, ciN), to_dev(torch, rsT), to_dev(torch, ciT))
    info = plan.info()
    assert info["maxRowT"] == 1500 and info["nnzT"] == nnzT
    for K in (1, 3):
        v = rng.uniform(-1, 1, Ncols * K)
        for trans in TRANS:
            want = ref.gather_scatter(v.copy(), K, O.ADD, trans)
            d = to_dev(torch, v)
            plan.GatherScatter(d, K, O.ADD, trans)
            assert same(to_host(torch, d), want), (K, trans)
            wg = ref.gather(v, K, O.MIN, trans)
            gv = torch.zeros(NrowsT * K, dtype=torch.float64, device="cuda")
            plan.Gather(gv, to_dev(torch, v), K, O.MIN, trans)
            assert same(to_host(torch, gv)[: len(wg)], wg), (K, trans)
            gvals = rng.uniform(-1, 1, NrowsT * K)
            ws = ref.scatter(v.copy(), gvals, K, trans)
            d = to_dev(torch, v)
            plan.Scatter(d, to_dev(torch, gvals), K, trans)
            assert same(to_host(torch, d), ws), (K, trans)


def rand_csr(rng, Ncols, nnz, lens_hi=9):
    """Rows of 1..lens_hi-1 distinct columns drawn from [0, Ncols) without
    repeats, nnz entries in all."""
    lens = rng.integers(1, lens_hi, nnz)
    cs = np.cumsum(lens)
    n = int(np.searchsorted(cs, nnz)) + 1
    lens = lens[:n]
    lens[-1] -= cs[n - 1] - nnz  # rows sum to nnz, the last one non-empty
    cols = rng.choice(Ncols, nnz, replace=False).astype(np.int32)
    rs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate([np.sort(cols[rs[r]:rs[r + 1]]) for r in range(len(lens))]).astype(np.int32)
    return rs, ci


def test_gather_scatter_trans_by_column_mask_and_without(torch, D):
    """GatherScatter of hand-built Signed operators whose N rows are ordered
    subsets of their T rows (the plan keeps the N list as a column mask and
    runs Trans over the T list alone) and whose N rows are not (an N entry
    replaced by a column in no T row: the plan keeps no mask and runs the
    two-list kernels), every trans, K = 1 and 3, against the oracle's host
    loops."""
    rng = np.random.default_rng(61)
    Ncols = 30_000
    rsT, ciT = rand_csr(rng, Ncols, 24_000)
    NrowsT = len(rsT) - 1
    rows = np.repeat(np.arange(NrowsT), np.diff(rsT))
    keep = rng.random(len(ciT)) < 0.6
    NrowsN = NrowsT - 7
    keep &= rows < NrowsN
    ciN = ciT[keep].copy()
    rsN = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=NrowsT))]).astype(np.int32)
    moved = ciN.copy()
    # one N entry of a multi-member row replaced by a column no T row holds
    r = int(np.flatnonzero(np.diff(rsN) >= 2)[0])
    free = np.setdiff1d(np.arange(Ncols), ciT)
    moved[rsN[r]] = free[0]
    moved[rsN[r]:rsN[r + 1]] = np.sort(moved[rsN[r]:rsN[r + 1]])
    for cN in (ciN, moved):
        ref = O.OracleCsr(rsT, ciT, rsN, cN, NrowsN, Ncols)
        plan = D.Plan(Ncols, NrowsN, NrowsT, to_dev(torch, rsN), to_dev(torch, cN), to_dev(torch, rsT),
                      to_dev(torch, ciT))
        for K in (1, 3):
            for trans in TRANS:
                v = rng.uniform(-1, 1, Ncols * K)
                want = ref.gather_scatter(v.copy(), K, O.ADD, trans)
                d = to_dev(torch, v)
                plan.GatherScatter(d, K, O.ADD, trans)
                assert same(to_host(torch, d), want), (cN is moved, K, trans)


@pytest.mark.parametrize("ty", list(TYPES))
def test_scatter_by_columns_and_by_rows(torch, D, ty):
    """Scatter over a dense T list (>= Ncols/2 entries: by columns, each
    column's row held by the plan) and a sparse N list (by rows), and over a
    plan whose T list is sparse too, for compile-time and runtime K."""
    rng = np.random.default_rng(40 + ty)
    Ncols = 20_000
    rsT, ciT = rand_csr(rng, Ncols, 16_000)
    NrowsT = len(rsT) - 1
    rows = np.repeat(np.arange(NrowsT), np.diff(rsT))
    keep = rng.random(len(ciT)) < 0.4
    ciN = ciT[keep]
    rsN = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=NrowsT))]).astype(np.int32)
    dense = O.OracleCsr(rsT, ciT, rsN, ciN, NrowsT, Ncols)
    p1 = D.Plan(Ncols, NrowsT, NrowsT, to_dev(torch, rsN), to_dev(torch, ciN), to_dev(torch, rsT), to_dev(torch, ciT))
    rsS, ciS = rand_csr(rng, Ncols, 6_000)
    sparse = O.OracleCsr(rsS, ciS, Ncols=Ncols)
    dS, cS = to_dev(torch, rsS), to_dev(torch, ciS)
    p2 = D.Plan(Ncols, len(rsS) - 1, len(rsS) - 1, dS, cS, dS, cS)
    i1, i2 = p1.info(), p2.info()
    assert (i1["colsT"], i1["colsN"], i2["colsT"], i2["colsN"]) == (1, 0, 0, 0), (i1, i2)
    dt = TYPES[ty]
    for plan, ref, nrows in ((p1, dense, NrowsT), (p2, sparse, len(rsS) - 1)):
        for K in (1, 2, 3, 5, 16):
            for trans in TRANS:
                v = values(rng, dt, Ncols * K)
                gv = values(rng, dt, nrows * K)
                want = ref.scatter(v.copy(), gv, K, trans)
                d = to_dev(torch, v)
                plan.Scatter(d, to_dev(torch, gv), K, trans)
                assert same(to_host(torch, d), want), (plan is p1, K, trans)


def test_scatter_by_columns_largest_row_wins(torch, D):
    """A hand-built list whose columns sit in several rows (Setup never
    builds one).  The reference's host loop runs over rows in parallel, so
    any of those rows' values may land last; scattering by columns (K values
    of at most 16 bytes, or K outside the row-wise kernels' shapes) always
    leaves the largest row's value -- a row-order sequential loop's result,
    brute force here."""
    rng = np.random.default_rng(44)
    Ncols = 5000
    NrowsT = 3000
    lens = rng.integers(1, 5, NrowsT)
    rs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(Ncols, l, replace=False)) for l in lens]).astype(np.int32)
    assert len(np.unique(ci)) < len(ci)  # repeated columns across rows
    d_rs, d_ci = to_dev(torch, rs), to_dev(torch, ci)
    plan = D.Plan(Ncols, NrowsT, NrowsT, d_rs, d_ci, d_rs, d_ci)
    assert plan.info()["colsT"] == 1
    for K, dt in ((1, np.float64), (2, np.float64), (5, np.float64), (7, np.int64), (2, np.float32),
                  (4, np.float32), (1, np.int32)):
        v = values(rng, dt, Ncols * K)
        gv = values(rng, dt, NrowsT * K)
        want = v.copy().reshape(Ncols, K)
        for r in range(NrowsT):
            want[ci[rs[r]:rs[r + 1]]] = gv[r * K:(r + 1) * K]
        d = to_dev(torch, v)
        plan.Scatter(d, to_dev(torch, gv), K, O.NOTRANS)
        assert same(to_host(torch, d), want.ravel()), K


def test_partition_cut_follows_empty_rows(torch, D):
    """T lists without empty rows (every Setup-built operator) are cut into
    blocks of W entries; a T list with empty rows -- here scattered ones and a
    run of 5000 -- is cut on entries + rows (merge path), which bounds a
    block's rows too.  Both must give the oracle's results."""
    import math
    g = D.ogs_t()
    ids = O.make_ids(6, 5, 4, 3)
    g.Setup(len(ids), ids.copy(), D.Unsigned)
    i = g.plan.info()
    assert i["nblocksK1"] == math.ceil(i["nnzT"] / i["widthK1"]), i
    rng = np.random.default_rng(17)
    Ncols = 30_000
    lens = rng.integers(0, 9, 9000)
    lens[3000:8000] = 0
    lens = lens[: np.searchsorted(np.cumsum(lens), Ncols)]
    NrowsT = len(lens)
    nnzT = int(lens.sum())
    cols = rng.permutation(Ncols)[:nnzT].astype(np.int32)
    rsT = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ciT = np.concatenate([np.sort(cols[rsT[r]:rsT[r + 1]]) for r in range(NrowsT)]).astype(np.int32)
    keep = rng.random(nnzT) < 0.7
    rows = np.repeat(np.arange(NrowsT), lens)
    ciN = ciT[keep]
    rsN = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=NrowsT))]).astype(np.int32)
    ref = O.OracleCsr(rsT, ciT, rsN, ciN, NrowsT, Ncols)
    plan = D.Plan(Ncols, NrowsT, NrowsT, to_dev(torch, rsN), to_dev(torch, ciN), to_dev(torch, rsT), to_dev(torch, ciT))
    i = plan.info()
    assert i["nblocksK1"] == math.ceil((nnzT + NrowsT) / i["widthK1"]), i
    for K in (1, 2, 3, 8):
        v = rng.uniform(-1, 1, Ncols * K)
        for trans in TRANS:
            for op in (O.ADD, O.MAX):
                want = ref.gather_scatter(v.copy(), K, op, trans)
                d = to_dev(torch, v)
                plan.GatherScatter(d, K, op, trans)
                assert same(to_host(torch, d), want), (K, trans, op)
            wg = ref.gather(v, K, O.ADD, trans)
            gv = torch.zeros(NrowsT * K, dtype=torch.float64, device="cuda")
            plan.Gather(gv, to_dev(torch, v), K, O.ADD, trans)
            assert same(to_host(torch, gv)[: len(wg)], wg), (K, trans)
            gvals = rng.uniform(-1, 1, NrowsT * K)
            ws = ref.scatter(v.copy(), gvals, K, trans)
            d = to_dev(torch, v)
            plan.Scatter(d, to_dev(torch, gvals), K, trans)
            assert same(to_host(torch, d), ws), (K, trans)


def test_unsortable_blocks_fall_back_to_csr_order(torch, D):
    """Columns scattered over 4M entries: a block's sorted columns need more
    than 4 runs of 2^19, so the plan leaves such blocks in CSR order.  Both
    kinds of block must give the oracle's result."""
    rng = np.random.default_rng(9)
    Ncols = 4_000_000
    lens = rng.integers(1, 9, 40_000)
    lens[:5000] = 1  # short rows over a narrow column range: two whole packable blocks
    nnz = int(lens.sum())
    narrow = np.sort(rng.choice(np.arange(Ncols - 50_000, Ncols), 5000, replace=False))
    wide = rng.choice(Ncols - 50_000, nnz - 5000, replace=False)
    cols = np.concatenate([narrow, wide]).astype(np.int32)
    rsT = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ciT = np.concatenate([np.sort(cols[rsT[r]:rsT[r + 1]]) for r in range(len(lens))]).astype(np.int32)
    assert len(np.unique(ciT)) == nnz
    ref = O.OracleCsr(rsT, ciT, Ncols=Ncols)
    d_rs, d_ci = to_dev(torch, rsT), to_dev(torch, ciT)
    plan = D.Plan(Ncols, len(lens), len(lens), d_rs, d_ci, d_rs, d_ci)
    info = plan.info()
    assert info["plainK1"] > 0 and info["packedK1"] > 0, info
    v = rng.uniform(-1, 1, Ncols)
    for op in OPS:
        want = ref.gather_scatter(v.copy(), 1, op, O.SYM)
        d = to_dev(torch, v)
        plan.GatherScatter(d, 1, op, O.SYM)
        assert same(to_host(torch, d), want), op
    wg = ref.gather(v, 1, O.ADD, O.TRANS)
    gv = torch.zeros(len(lens), dtype=torch.float64, device="cuda")
    plan.Gather(gv, to_dev(torch, v), 1, O.ADD, O.TRANS)
    assert same(to_host(torch, gv), wg)
    ws = ref.scatter(v.copy(), wg, 1, O.NOTRANS)
    d = to_dev(torch, v)
    plan.Scatter(d, gv, 1, O.NOTRANS)
    assert same(to_host(torch, d), ws)


@pytest.mark.parametrize("ty", list(TYPES))
@pytest.mark.parametrize("K", [1, 3])
def test_extract(torch, D, ty, K):
    rng = np.random.default_rng(11)
    q = values(rng, TYPES[ty], 5000 * K)
    ids = rng.integers(0, 5000, 1234).astype(np.int32)
    want = O.extract(ids, q, K)
    gq = to_dev(torch, np.zeros(len(ids) * K, TYPES[ty]))
    D.extract(gq, to_dev(torch, q), to_dev(torch, ids), K)
    assert same(to_host(torch, gq), want)
