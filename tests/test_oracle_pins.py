"""Pins every operator function of the CPU oracle against a brute force that
works from the ids alone, never through the oracle's CSR or its list
selection:

  * GatherScatter Sym / NoTrans / Trans, Gather NoTrans / Trans and Scatter
    NoTrans / Trans, for float, double, int and long long, each op, K = 1
    and 3, on Unsigned, flagged Signed (ragged, with id 0), heavily
    flagged Signed and unique Signed boxes.  The semantics are restated per
    group of equal |id| from ogsOperator.cpp (host loops):
      - Gather (60-97): NoTrans reduces a group's positive members, Trans
        all its members, from the op's init, in ascending local index;
      - Scatter (206-238): Trans writes a row's value to the group's
        positive members, NoTrans to all members;
      - GatherScatter (311-360, list table 321-339): Sym reduces all
        members and writes all; Trans reduces all members and writes the
        positive ones (groups without one are untouched); NoTrans reduces
        the positive members (init when there are none) and writes all;
  * extract (540-559) for every storage type;
  * the Halo-kind setup at one rank (ogsSetup.cpp:55-64, 192-331, 779-847).

Planted mistakes (in the oracle's inputs, on the CPU) show the pins bite:
N and T lists swapped, Trans run as NoTrans, an off-by-one extract index.
"""
import numpy as np
import pytest

from oracle import oracle as O
from test_oracle import brute_setup

# inf - inf and NaN operands are part of the cases; numpy warns about them
pytestmark = pytest.mark.filterwarnings("ignore::RuntimeWarning")

TYPES = [np.float32, np.float64, np.int32, np.int64]
OPS = [O.ADD, O.MUL, O.MAX, O.MIN]


# ---------------------------------------------------------------- brute force
def _init(dt, op):
    """ogsOperator.cpp:36-55: 0, 1, -max, +max of the type."""
    if op == O.ADD:
        return dt(0)
    if op == O.MUL:
        return dt(1)
    big = np.finfo(dt).max if np.issubdtype(dt, np.floating) else np.iinfo(dt).max
    return dt(-big) if op == O.MAX else dt(big)


def _wrap(x: int, dt) -> int:
    bits = np.iinfo(dt).bits
    return ((x + (1 << (bits - 1))) % (1 << bits)) - (1 << (bits - 1))


def _apply(acc, x, dt, op):
    """One step of the host functor: a += b, a *= b, compare-selects."""
    if op == O.MAX:
        return x if x > acc else acc
    if op == O.MIN:
        return x if x < acc else acc
    if np.issubdtype(dt, np.integer):
        r = int(acc) + int(x) if op == O.ADD else int(acc) * int(x)
        return dt(_wrap(r, dt))
    return dt(acc + x) if op == O.ADD else dt(acc * x)


def _fold(vals, dt, op):
    acc = _init(dt, op)
    for x in vals:
        acc = _apply(acc, x, dt, op)
    return acc


class Brute:
    """The single-rank ogs semantics per group of equal |id|, from the ids."""

    def __init__(self, ids, kind, unique):
        self.ids, self.NrowsN, self.NrowsT, _, _, self.gdef = brute_setup(ids, kind, unique)
        self.kind = kind
        groups, order = {}, []
        for n, g in enumerate(self.ids):
            if g == 0:
                continue
            a = abs(int(g))
            if a not in groups:
                groups[a] = []
                order.append(a)
            groups[a].append(n)
        pos = lambda n: kind == O.UNSIGNED or self.ids[n] > 0
        self.members = [groups[a] for a in order]                    # ascending local index
        self.positive = [[n for n in groups[a] if pos(n)] for a in order]
        # row of a group: positive groups first, each class in first-encounter order
        has = [bool(p) for p in self.positive]
        rows = [i for i in range(len(order)) if has[i]] + [i for i in range(len(order)) if not has[i]]
        self.row = {g: r for r, g in enumerate(rows)}

    def gather_scatter(self, v, K, op, trans):
        v = v.copy()
        dt = v.dtype.type
        src = v.copy()
        for g, (mem, pos) in enumerate(zip(self.members, self.positive)):
            if trans == O.TRANS and not pos:
                continue                      # rows >= NrowsN: not touched
            gl = pos if trans == O.NOTRANS else mem
            sl = pos if trans == O.TRANS else mem
            for k in range(K):
                val = _fold([src[k + m * K] for m in gl], dt, op)
                for m in sl:
                    v[k + m * K] = val
        return v

    def gather(self, v, K, op, trans):
        dt = v.dtype.type
        nrows = self.NrowsN if trans == O.NOTRANS else self.NrowsT
        gv = np.zeros(nrows * K, dtype=v.dtype)
        for g, (mem, pos) in enumerate(zip(self.members, self.positive)):
            r = self.row[g]
            if r >= nrows:
                continue
            gl = pos if trans == O.NOTRANS else mem
            for k in range(K):
                gv[k + r * K] = _fold([v[k + m * K] for m in gl], dt, op)
        return gv

    def scatter(self, v, gv, K, trans):
        v = v.copy()
        nrows = self.NrowsN if trans == O.TRANS else self.NrowsT
        for g, (mem, pos) in enumerate(zip(self.members, self.positive)):
            r = self.row[g]
            if r >= nrows:
                continue
            for m in (pos if trans == O.TRANS else mem):
                for k in range(K):
                    v[k + m * K] = gv[k + r * K]
        return v


SPECIALS = [0.0, -0.0, np.nan, np.inf, -np.inf, 0.5, -0.5]


def _values(rng, dt, n, specials=False):
    """Uniform values; with `specials`, ~15% of them signed zeros, NaN,
    infinities and repeated values, so that ties (-0 against +0), NaN
    skipping and inf - inf reach the compare-selects and sums."""
    if np.issubdtype(dt, np.floating):
        v = rng.uniform(-1, 1, n).astype(dt)
        if specials:
            at = rng.random(n) < 0.15
            v[at] = np.array(SPECIALS, dtype=dt)[rng.integers(0, len(SPECIALS), int(at.sum()))]
        return v
    return rng.integers(-2 ** 20, 2 ** 20, n).astype(dt)


def _ids(box, kind, unique, flag, rng):
    ids = O.make_ids(*box).copy()
    ids[rng.random(len(ids)) < 0.1] = 0
    if flag:
        neg = rng.random(len(ids)) < flag
        ids[neg] = -ids[neg]
    return ids


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))


MESHES = [  # (box, kind, unique, flagged fraction)
    ((2, 2, 1, 2), O.UNSIGNED, False, 0.0),
    ((3, 2, 2, 1), O.SIGNED, False, 0.3),
    ((2, 2, 2, 2), O.SIGNED, False, 0.9),
    ((2, 1, 2, 3), O.SIGNED, True, 0.0),
]


# ---------------------------------------------------------------- operator pins
@pytest.mark.parametrize("mesh", range(len(MESHES)))
@pytest.mark.parametrize("dt", TYPES, ids=lambda t: t.__name__)
@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("K", [1, 3])
def test_operator_functions_vs_brute_force(mesh, dt, op, K):
    box, kind, unique, flag = MESHES[mesh]
    rng = np.random.default_rng(1000 * mesh + 10 * op + K)
    ids = _ids(box, kind, unique, flag, rng)
    b = Brute(ids, kind, unique)
    o = O.OracleOgs(ids, kind, unique)
    assert (o.NrowsN, o.NrowsT) == (b.NrowsN, b.NrowsT)
    N = len(ids)
    v = _values(rng, dt, N * K, specials=True)
    for trans in (O.SYM, O.NOTRANS, O.TRANS):
        want = b.gather_scatter(v, K, op, trans)
        assert _same(o.op_gather_scatter(v.copy(), K, op, trans), want), f"GatherScatter trans={trans}"
        # the ogs_t wrapper at one rank runs the local operator only
        assert _same(o.gather_scatter(v.copy(), K, op, trans), want)
    for trans in (O.NOTRANS, O.TRANS):
        want = b.gather(v, K, op, trans)
        assert _same(o.op_gather(v.copy(), K, op, trans), want), f"Gather trans={trans}"
    gv = _values(rng, dt, max(b.NrowsT, 1) * K)[: b.NrowsT * K]
    for trans in (O.NOTRANS, O.TRANS):
        want = b.scatter(v, gv, K, trans)
        assert _same(o.op_scatter(v.copy(), gv, K, trans), want), f"Scatter trans={trans}"
    if o.gather_defined:
        for trans in (O.NOTRANS, O.TRANS):
            assert _same(o.gather(v.copy(), K, op, trans)[: (b.NrowsN if trans == O.NOTRANS else b.NrowsT) * K],
                         b.gather(v, K, op, trans))
            assert _same(o.scatter(v.copy(), gv, K, trans), b.scatter(v, gv, K, trans))


@pytest.mark.parametrize("mesh", range(len(MESHES)))
@pytest.mark.parametrize("dt", [np.float32, np.float64], ids=lambda t: t.__name__)
@pytest.mark.parametrize("op", [O.MAX, O.MIN, O.ADD])
def test_signed_zero_and_nan_ties(mesh, dt, op):
    """Values drawn from {+0, -0, NaN, -1, 1} only: Max/Min keep the earlier
    of two equal values (-0 against +0) and skip NaN (the compare-select of
    ogsOperator.cpp:46-55); Add starts from +0, so a row of -0 gives +0."""
    box, kind, unique, flag = MESHES[mesh]
    rng = np.random.default_rng(77 + mesh)
    ids = _ids(box, kind, unique, flag, rng)
    b = Brute(ids, kind, unique)
    o = O.OracleOgs(ids, kind, unique)
    v = np.array([0.0, -0.0, np.nan, -1.0, 1.0], dtype=dt)[rng.integers(0, 5, len(ids))]
    for trans in (O.SYM, O.NOTRANS, O.TRANS):
        assert _same(o.op_gather_scatter(v.copy(), 1, op, trans), b.gather_scatter(v, 1, op, trans))
    for trans in (O.NOTRANS, O.TRANS):
        assert _same(o.op_gather(v.copy(), 1, op, trans), b.gather(v, 1, op, trans))


def test_brute_force_covers_the_distinct_cases():
    """The meshes above hold every case the list table distinguishes: groups
    with several positive members, with none (rows past NrowsN), with one,
    and ignored nodes (id 0)."""
    seen = set()
    for m, (box, kind, unique, flag) in enumerate(MESHES):
        rng = np.random.default_rng(1000 * m + 10 * O.ADD + 1)
        b = Brute(_ids(box, kind, unique, flag, rng), kind, unique)
        npos = [len(p) for p in b.positive]
        seen |= {"none" if c == 0 else "one" if c == 1 else "several" for c in npos}
        if b.NrowsN < b.NrowsT:
            seen.add("rows past NrowsN")
        if any(len(p) < len(mm) for p, mm in zip(b.positive, b.members) if p):
            seen.add("flagged members beside positive ones")
    assert seen == {"none", "one", "several", "rows past NrowsN", "flagged members beside positive ones"}


# ---------------------------------------------------------------- planted mistakes
def _flagged():
    box, kind, unique, flag = MESHES[1]
    rng = np.random.default_rng(7)
    ids = _ids(box, kind, unique, flag, rng)
    return ids, Brute(ids, kind, unique), O.OracleOgs(ids, kind, unique), rng


@pytest.mark.parametrize("op", OPS)
def test_planted_list_swap_is_caught(op):
    """The oracle's GatherScatter over a CSR whose N and T lists are swapped
    (the plausible mistake of a transposed list table) differs from the
    brute force, for Trans and for NoTrans."""
    ids, b, o, rng = _flagged()
    v = _values(rng, np.float64, len(ids))
    swapped = O.OracleCsr(o.rowStartsN, o.colIdsN, o.rowStartsT, o.colIdsT, NrowsN=o.NrowsN, Ncols=len(ids))
    right = O.OracleCsr(o.rowStartsT, o.colIdsT, o.rowStartsN, o.colIdsN, NrowsN=o.NrowsN, Ncols=len(ids))
    for trans in (O.TRANS, O.NOTRANS):
        want = b.gather_scatter(v, 1, op, trans)
        assert _same(right.gather_scatter(v.copy(), 1, op, trans), want)
        assert not _same(swapped.gather_scatter(v.copy(), 1, op, trans), want)
    # Trans and NoTrans differ from each other and from Sym on a flagged operator
    outs = [b.gather_scatter(v, 1, op, t) for t in (O.SYM, O.NOTRANS, O.TRANS)]
    assert not _same(outs[0], outs[1]) and not _same(outs[0], outs[2]) and not _same(outs[1], outs[2])


def test_planted_gather_and_scatter_list_swaps_are_caught():
    ids, b, o, rng = _flagged()
    v = _values(rng, np.float64, len(ids))
    gv = _values(rng, np.float64, b.NrowsT)
    # Gather NoTrans run over T lists, Scatter Trans run over T lists
    assert not _same(o.op_gather(v.copy(), 1, O.ADD, O.TRANS)[: b.NrowsN], b.gather(v, 1, O.ADD, O.NOTRANS))
    assert not _same(o.op_scatter(v.copy(), gv, 1, O.NOTRANS), b.scatter(v, gv, 1, O.TRANS))


# ---------------------------------------------------------------- extract
EXTRACT_TYPES = [(np.float32, None), (np.float64, None), (np.int32, None), (np.int64, None),
                 (np.uint16, O.BF16), (np.float32, O.FLOAT_C)]


@pytest.mark.parametrize("dt,code", EXTRACT_TYPES, ids=lambda x: getattr(x, "__name__", str(x)))
@pytest.mark.parametrize("K", [1, 3, 5])
def test_extract_vs_brute_force(dt, code, K):
    """extract (ogsOperator.cpp:540-559): gq[k + n*K] = q[k + ids[n]*K],
    element by element, with repeated and unordered ids."""
    rng = np.random.default_rng(K)
    Nq, Ns = 97, 211
    if dt == np.uint16:
        q = rng.integers(0, 1 << 16, Nq * K).astype(np.uint16)
    else:
        q = _values(rng, dt, Nq * K)
    ids = rng.integers(0, Nq, Ns)
    ids[:3] = [Nq - 1, 0, Nq - 1]
    want = np.empty(Ns * K, dtype=q.dtype)
    for n in range(Ns):
        for k in range(K):
            want[k + n * K] = q[k + ids[n] * K]
    assert _same(O.extract(ids, q, K, type=code), want)
    # planted: an off-by-one index is caught
    assert not _same(O.extract(np.minimum(ids + 1, Nq - 1), q, K, type=code), want)


def test_extract_empty():
    assert O.extract(np.zeros(0, np.int32), np.ones(4), 2).size == 0


# ---------------------------------------------------------------- Halo kind
def brute_halo(ids):
    """Halo-kind setup at one rank: groups keep their signs; each group
    needs exactly one positive id (ogsSetup.cpp:286-288); every group is
    local (no other rank), so NlocalP = NlocalT = groups and the halo
    operator has no rows; the local operator is not built (779-847)."""
    groups = {}
    for g in ids:
        if g != 0:
            groups.setdefault(abs(int(g)), []).append(int(g) > 0)
    if any(sum(p) != 1 for p in groups.values()):
        return None
    return len(groups)


@pytest.mark.parametrize("box", [(1, 1, 1, 1), (2, 2, 1, 2), (3, 2, 2, 3)])
def test_halo_setup_vs_brute_force(box):
    rng = np.random.default_rng(sum(box))
    ids = O.make_ids(*box).copy()
    ids[rng.random(len(ids)) < 0.15] = 0
    # one positive member per group: the unique labelling of the brute force
    ids, *_ = brute_setup(ids, O.SIGNED, True)
    ngroups = brute_halo(ids)
    assert ngroups is not None
    o = O.OracleOgs(ids, O.HALO)
    assert (o.NlocalP, o.NlocalT, o.NhaloP, o.NhaloT) == (ngroups, ngroups, 0, 0)
    assert o.Ngather == o.NgatherGlobal == ngroups
    assert (o.NrowsN, o.NrowsT) == (0, 0)     # no local operator
    assert o.NhaloT - o.NhaloP == 0            # halo_t::Nhalo (ogsSetup.cpp:63)
    # any group with no positive id, or with two, is refused
    bad = ids.copy()
    first = np.flatnonzero(bad > 0)[0]
    bad[first] = -bad[first]
    assert brute_halo(bad) is None
    with pytest.raises(O.OracleError) as e:
        O.OracleOgs(bad, O.HALO)
    assert e.value.code == -2
    two = ids.copy()
    neg = np.flatnonzero(two < 0)
    if len(neg):
        two[neg[0]] = -two[neg[0]]
        assert brute_halo(two) is None
        with pytest.raises(O.OracleError):
            O.OracleOgs(two, O.HALO)
