"""Pins the CPU oracle (oracle/ogs_oracle.c) before anything is checked
against it.  Each pin is independent of the oracle's own code path:

  * closed forms of the structured numbering (dawgsMain.cpp:126-134);
  * a dictionary-based brute-force setup written from the ogs semantics
    (ogs.hpp:42-120 and the row-order rule ogsSetup.cpp:409-433);
  * a numpy model of the greedy row blocks (ogsOperator.cpp:447-512) and
    the block counts of the reference configurations;
  * the reference's own legacy comparator, gslib gs_vec, compiled from its
    sources (dawgsGS.cpp:34-71, the "-cc ... Error = 0" check);
  * known answers (v = 1 gives multiplicities, integer-valued sums are
    exact, Scatter(Gather^T v) = GS_Sym v).
"""
import numpy as np
import pytest

from oracle import oracle as O

CONFIG_COUNTS = {  # (nx, p): (N, Ngather, {len: rows}, NrowBlocksT)
    (10, 3): (64000, 29791, {1: 10648, 2: 13068, 4: 5346, 8: 729}, 126),
    (24, 3): (884736, 389017, {1: 125000, 2: 172500, 4: 79350, 8: 12167}, 1729),
}


def mult_1d(n_el, p):
    """Multiplicity of each 1-D grid point: 2 on interior element boundaries."""
    m = np.ones(n_el * p + 1, dtype=np.int64)
    m[p:n_el * p:p] = 2
    return m


def expected_multiplicity(nx, ny, nz, p):
    """Per local node, the number of elements sharing its id."""
    mx, my, mz = mult_1d(nx, p), mult_1d(ny, p), mult_1d(nz, p)
    ids = O.make_ids(nx, ny, nz, p) - 1
    sx, sy = p * nx + 1, (p * nx + 1) * (p * ny + 1)
    gz, rem = np.divmod(ids, sy)
    gy, gx = np.divmod(rem, sx)
    return mx[gx] * my[gy] * mz[gz]


# ---------------------------------------------------------------- ids
@pytest.mark.parametrize("nx,ny,nz,p", [(1, 1, 1, 1), (2, 3, 4, 2), (3, 2, 2, 5), (10, 10, 10, 3)])
def test_make_ids_closed_form(nx, ny, nz, p):
    ids = O.make_ids(nx, ny, nz, p)
    Nq = p + 1
    assert len(ids) == nx * ny * nz * Nq ** 3
    assert ids.min() == 1
    assert ids.max() == (p * nx + 1) * (p * ny + 1) * (p * nz + 1)
    assert len(np.unique(ids)) == ids.max()
    # element (I,J,K) node (i,j,k) -> grid point (I*p+i, J*p+j, K*p+k)
    e = ids.reshape(nz, ny, nx, Nq, Nq, Nq)
    assert e[0, 0, 0, 0, 0, 0] == 1
    if nx > 1:
        assert e[0, 0, 1, 0, 0, 0] == 1 + p
        assert e[0, 0, 0, 0, 0, p] == e[0, 0, 1, 0, 0, 0]  # shared face node
    counts = np.bincount(ids)[ids]
    assert np.array_equal(counts, expected_multiplicity(nx, ny, nz, p))


# ---------------------------------------------------------------- setup
def brute_setup(ids, kind, unique):
    """Dictionary-based restatement of the single-rank setup semantics."""
    ids = np.array(ids, dtype=np.int64)
    groups = {}  # |id| -> members in local order
    order = []
    for n, g in enumerate(ids):
        if g == 0:
            continue
        a = abs(int(g))
        if a not in groups:
            groups[a] = []
            order.append(a)
        groups[a].append(n)
    if unique:
        for a in order:
            m = groups[a]
            for j, n in enumerate(m):
                ids[n] = a if j == 0 else -a
    positive = lambda n: kind == O.UNSIGNED or ids[n] > 0
    pos = [a for a in order if any(positive(n) for n in groups[a])]
    neg = [a for a in order if not any(positive(n) for n in groups[a])]
    rowsT, rowsN = [], []
    for a in pos + neg:
        rowsT.append(groups[a])
        rowsN.append([n for n in groups[a] if positive(n)])
    gather_defined = all(sum(positive(n) for n in groups[a]) == 1 for a in order)
    return ids, len(pos), len(order), rowsT, rowsN, gather_defined


def csr_rows(rs, ci):
    return [list(ci[rs[r]:rs[r + 1]]) for r in range(len(rs) - 1)]


@pytest.mark.parametrize("box", [(1, 1, 1, 1), (2, 1, 1, 2), (2, 2, 2, 1), (3, 2, 2, 2), (3, 3, 3, 3), (2, 3, 1, 3)])
@pytest.mark.parametrize("kind,unique,flag", [(O.UNSIGNED, False, 0.0), (O.SIGNED, False, 0.3),
                                              (O.SIGNED, False, 0.0), (O.SIGNED, True, 0.0),
                                              (O.SIGNED, False, 0.9)])
def test_setup_brute_force(box, kind, unique, flag, rng):
    ids = O.make_ids(*box)
    ids = ids.copy()
    # ragged: some ignored nodes, some flagged (negative) nodes
    ids[rng.random(len(ids)) < 0.1] = 0
    if flag:
        neg = rng.random(len(ids)) < flag
        ids[neg] = -ids[neg]
    want_ids, NlocalP, NlocalT, rowsT, rowsN, gdef = brute_setup(ids, kind, unique)
    o = O.OracleOgs(ids, kind, unique)
    assert (o.NlocalP, o.NlocalT, o.Ngather, o.NgatherGlobal) == (NlocalP, NlocalT, NlocalP, NlocalP)
    assert o.NhaloP == 0 and o.NhaloT == 0
    assert o.gather_defined == (True if unique else gdef)
    assert (o.NrowsN, o.NrowsT) == (NlocalP, NlocalT)
    assert csr_rows(o.rowStartsT, o.colIdsT) == rowsT
    assert csr_rows(o.rowStartsN, o.colIdsN) == rowsN
    if unique:
        assert np.array_equal(o.ids, want_ids)


@pytest.mark.parametrize("nx,p", list(CONFIG_COUNTS))
def test_setup_counts_match_reference_configs(nx, p):
    N, Ng, hist, nblk = CONFIG_COUNTS[(nx, p)]
    ids = O.make_ids(nx, nx, nx, p)
    o = O.OracleOgs(ids, O.UNSIGNED)
    assert o.N == N and o.NgatherGlobal == Ng == (p * nx + 1) ** 3
    h = np.bincount(np.diff(o.rowStartsT))
    assert {int(k): int(v) for k, v in enumerate(h) if v} == hist
    assert o.NrowBlocksT == nblk
    # CSR invariants: colIds a permutation of the local indices, ascending in
    # each row, first members strictly increasing (first-encounter order)
    assert np.array_equal(np.sort(o.colIdsT), np.arange(N))
    first = o.colIdsT[o.rowStartsT[:-1]]
    assert np.all(np.diff(first) > 0)
    rows = np.repeat(np.arange(o.NrowsT), np.diff(o.rowStartsT))
    same_row = rows[1:] == rows[:-1]
    assert np.all(np.diff(o.colIdsT)[same_row] > 0)


def greedy_blocks(rs, budget=512):
    """Independent model of the greedy row packing."""
    sizes = np.diff(rs)
    starts = [0] if len(sizes) else []
    acc = 0
    for i, s in enumerate(sizes):
        if acc + s > budget:
            starts.append(i)
            acc = s
        else:
            acc += s
    return np.array(starts + [len(sizes)], dtype=np.int32)


@pytest.mark.parametrize("box", [(10, 10, 10, 3), (7, 5, 3, 4), (4, 4, 4, 7)])
def test_row_blocks_model(box):
    o = O.OracleOgs(O.make_ids(*box), O.SIGNED, True)
    assert np.array_equal(o.blockRowStartsT, greedy_blocks(o.rowStartsT))
    assert np.array_equal(o.blockRowStartsN, greedy_blocks(o.rowStartsN))
    nnz = np.diff(o.rowStartsT)[:, None]
    blk = o.blockRowStartsT
    for b in range(len(blk) - 1):
        assert o.rowStartsT[blk[b + 1]] - o.rowStartsT[blk[b]] <= 512


def test_setup_errors():
    ids = O.make_ids(2, 2, 2, 2)
    with pytest.raises(O.OracleError) as e:
        O.OracleOgs(ids, O.UNSIGNED, True)
    assert e.value.code == -1
    with pytest.raises(O.OracleError) as e:
        O.OracleOgs(ids, O.HALO, True)
    assert e.value.code == -1
    # a row of 513 members exceeds the reference's 512 limit
    with pytest.raises(O.OracleError) as e:
        O.OracleOgs(np.full(513, 7, dtype=np.int64), O.UNSIGNED)
    assert e.value.code == -3
    O.OracleOgs(np.full(512, 7, dtype=np.int64), O.UNSIGNED)
    # Gather on a non-unique setup is undefined
    o = O.OracleOgs(ids, O.UNSIGNED)
    assert not o.gather_defined
    with pytest.raises(O.OracleError):
        o.gather(np.ones(len(ids)), 1, O.ADD, O.TRANS)


# ---------------------------------------------------------------- arithmetic
def test_known_answers(rng):
    box = (5, 4, 3, 3)
    ids = O.make_ids(*box)
    N = len(ids)
    m = expected_multiplicity(*box)
    o = O.OracleOgs(ids, O.UNSIGNED)
    # v = 1 -> multiplicity, exactly
    v = o.gather_scatter(np.ones(N), 1, O.ADD, O.SYM)
    assert np.array_equal(v, m.astype(np.float64))
    # Mul of 2s -> 2^multiplicity
    v = o.gather_scatter(np.full(N, 2, np.int64), 1, O.MUL, O.SYM)
    assert np.array_equal(v, 2 ** m)
    # Max/Min of the local index -> largest / smallest member index of the group
    idx = np.arange(N, dtype=np.int32)
    order = np.argsort(ids, kind="stable")
    mx = np.empty(N, np.int64)
    mn = np.empty(N, np.int64)
    bnd = np.flatnonzero(np.diff(ids[order])) + 1
    for grp in np.split(order, bnd):
        mx[grp] = grp.max()
        mn[grp] = grp.min()
    assert np.array_equal(o.gather_scatter(idx.copy(), 1, O.MAX, O.SYM), mx)
    assert np.array_equal(o.gather_scatter(idx.copy(), 1, O.MIN, O.SYM), mn)
    # integer-valued doubles: the sum is exact in any order -> compare to bincount
    vi = rng.integers(-1000, 1000, N).astype(np.float64)
    sums = np.bincount(ids, weights=vi)
    assert np.array_equal(o.gather_scatter(vi.copy(), 1, O.ADD, O.SYM), sums[ids])
    # Gather(1) over T = row lengths; Scatter(Gather^T v) = GS_Sym v
    s = O.OracleOgs(ids, O.SIGNED, True)
    gl = s.gather(np.ones(N), 1, O.ADD, O.TRANS)
    assert np.array_equal(gl, np.diff(s.rowStartsT).astype(np.float64))
    x = rng.uniform(-1, 1, N)
    g = s.gather(x, 1, O.ADD, O.TRANS)
    y = s.scatter(np.zeros(N), g, 1, O.NOTRANS)
    assert np.array_equal(y, s.gather_scatter(x.copy(), 1, O.ADD, O.SYM))
    # K fields are independent: field k of a K=3 run equals the K=1 run on field k
    X = rng.uniform(-1, 1, (N, 3))
    Y = o.gather_scatter(X.copy().reshape(-1), 3, O.ADD, O.SYM).reshape(N, 3)
    for k in range(3):
        assert np.array_equal(Y[:, k], o.gather_scatter(X[:, k].copy(), 1, O.ADD, O.SYM))


def test_identities_and_special_values():
    # rows with no positive member produce the op's init in Gather NoTrans:
    # 0, 1, -max, +max (ogsOperator.cpp:36-55), not +-inf or INT_MIN
    ids = np.array([1, -2, -2, 3], dtype=np.int64)
    o = O.OracleOgs(ids, O.SIGNED, False)
    assert (o.NrowsN, o.NrowsT) == (2, 3)
    v = np.array([5.0, 6.0, 7.0, 8.0])
    for op, init in [(O.ADD, 0.0), (O.MUL, 1.0), (O.MAX, -np.finfo(np.float64).max),
                     (O.MIN, np.finfo(np.float64).max)]:
        # NoTrans GS: gather N (row of -2 is empty) -> scatter T
        y = o.op_gather_scatter(v.copy(), 1, op, O.NOTRANS)
        assert y[1] == init and y[2] == init
    vi = np.array([5, 6, 7, 8], dtype=np.int32)
    y = o.op_gather_scatter(vi.copy(), 1, O.MAX, O.NOTRANS)
    assert y[1] == -np.iinfo(np.int32).max
    # +0.0 start: a lone -0.0 comes out +0.0; NaN is skipped by Max/Min
    z = O.OracleOgs(np.array([1, 2, 2], dtype=np.int64), O.UNSIGNED)
    r = z.gather_scatter(np.array([-0.0, np.nan, 1.0]), 1, O.ADD, O.SYM)
    assert r[0] == 0.0 and not np.signbit(r[0])
    r = z.gather_scatter(np.array([-0.0, np.nan, 1.0]), 1, O.MAX, O.SYM)
    assert r[1] == 1.0 and r[2] == 1.0
    r = z.gather_scatter(np.array([-0.0, 1.0, np.nan]), 1, O.MIN, O.SYM)
    assert r[1] == 1.0
    # integer Mul wraps modulo 2^32
    w = z.gather_scatter(np.array([3, 2 ** 20, 2 ** 20], dtype=np.int32), 1, O.MUL, O.SYM)
    assert w[1] == np.int32(np.uint32((2 ** 40) % 2 ** 32).astype(np.int64) - 0)


# ---------------------------------------------------------------- gslib
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="gslib comparator not built")


@needs_ref
def test_reference_cc_error_zero():
    """dawgsMain.cpp:314-416 with the reference's own input: Error = 0."""
    ids = O.make_ids(10, 10, 10, 3)
    q = O.reference_q(0, len(ids))
    assert np.all(np.abs(q) <= 0.25)
    qcheck = O.GsRef(ids).gs_vec(q, 1)
    o = O.OracleOgs(ids, O.UNSIGNED)
    assert np.array_equal(o.gather_scatter(q.copy(), 1, O.ADD, O.SYM), qcheck)
    s = O.OracleOgs(ids, O.SIGNED, True)
    assert np.array_equal(s.gather_scatter(q.copy(), 1, O.ADD, O.SYM), qcheck)
    g = s.gather(q, 1, O.ADD, O.TRANS)
    assert np.array_equal(s.scatter(np.zeros_like(q), g, 1, O.NOTRANS), qcheck)


@needs_ref
@pytest.mark.parametrize("dt", [np.float64, np.float32, np.int32, np.int64])
@pytest.mark.parametrize("op", [O.ADD, O.MUL, O.MAX, O.MIN])
@pytest.mark.parametrize("K", [1, 3])
def test_oracle_matches_gslib(dt, op, K, rng):
    box = (6, 5, 4, 3)
    ids = O.make_ids(*box)
    N = len(ids)
    if np.issubdtype(dt, np.floating):
        v = rng.uniform(-1, 1, N * K).astype(dt)
    else:
        v = rng.integers(-2 ** 20, 2 ** 20, N * K).astype(dt)
    want = O.GsRef(ids).gs_vec(v, K, op)
    got = O.OracleOgs(ids, O.UNSIGNED).gather_scatter(v.copy(), K, op, O.SYM)
    assert np.array_equal(got, want)


@needs_ref
def test_planted_order_mistake_is_caught(rng):
    """Summing a row in reverse CSR order (a plausible kernel mistake) must
    differ from gslib's in-order sum: the gslib pin detects order errors."""
    ids = O.make_ids(6, 6, 6, 3)
    o = O.OracleOgs(ids, O.UNSIGNED)
    v = rng.uniform(-1, 1, len(ids))
    rev = o.colIdsT.copy()
    for r in range(o.NrowsT):
        a, b = o.rowStartsT[r], o.rowStartsT[r + 1]
        rev[a:b] = rev[a:b][::-1]
    wrong = O.csr_gather_scatter(o.rowStartsT, rev, v.copy(), 1, O.ADD)
    assert not np.array_equal(wrong, O.GsRef(ids).gs_vec(v, 1))
    right = O.csr_gather_scatter(o.rowStartsT, o.colIdsT, v.copy(), 1, O.ADD)
    assert np.array_equal(right, O.GsRef(ids).gs_vec(v, 1))
