"""Pins the oracle's precision-form types (OR_BF16, OR_FLOAT_C) before the
GPU path is checked against them.  These types have no counterpart in the
reference; the oracle restates the library's own specification (float
accumulation in CSR order with a two-sum running error, bf16 rounding to
nearest even), so each pin here comes from outside that code:

  * torch's own float32 <-> bfloat16 conversions (round to nearest even);
  * exact rational row sums (fractions.Fraction) and the published error
    bound of the cascaded sum Sum2 (Ogita, Rump, Oishi 2005, Prop. 4.5):
    |res - S| <= u|S| + gamma_{n-1}^2 sum|x|, u = 2^-24;
  * known answers that plain float summation gets wrong, so dropping the
    compensation fails them (the oracle's own plain float Add is used as
    that planted mistake);
  * Min/Max against numpy's reductions of the widened stored values;
  * Scatter(Gather_T v) == GS_Sym v, and K interleaved fields == K separate
    single-field calls.
"""
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O

U = 2.0 ** -24


def gamma(k):
    return k * U / (1 - k * U)


def csr_rows(rows):
    """CSR over rows given as lists of column ids (Sym operator)."""
    rs = np.zeros(len(rows) + 1, np.int32)
    rs[1:] = np.cumsum([len(r) for r in rows])
    ci = np.array([c for r in rows for c in r], np.int32)
    return rs, ci


def row_sums(values, dtype_type, K=1):
    """Each value its own column; one row holding all of them; returns the
    Gather (Trans) of that single row."""
    n = len(values) // K
    ref = O.OracleCsr(np.array([0, n], np.int32), np.arange(n, dtype=np.int32), Ncols=n)
    return ref.gather(values, K, O.ADD, O.TRANS, type=dtype_type)


# --------------------------------------------------------------------------
# bfloat16 conversions against torch

def torch_bf16_round(f32):
    torch = pytest.importorskip("torch")
    t = torch.from_numpy(np.ascontiguousarray(f32, np.float32)).to(torch.bfloat16)
    return t.view(torch.int16).numpy().view(np.uint16)


def special_floats(rng, n=200000):
    u = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    half = (rng.integers(0, 2 ** 16, 4000, dtype=np.uint64).astype(np.uint32) << 16) | 0x8000  # ties
    near = (rng.integers(0, 2 ** 16, 4000, dtype=np.uint64).astype(np.uint32) << 16) | \
        rng.choice(np.array([0x7FFF, 0x8001, 0x0001, 0xFFFF], np.uint32), 4000)
    edge = np.array([0, 0x80000000, 0x7F800000, 0xFF800000, 0x7F7FFFFF, 0xFF7FFFFF, 0x7F7F8000, 0x7F7F7FFF,
                     0x00000001, 0x807FFFFF, 0x00008000, 0x00018000, 0x3F808000, 0x3F818000], np.uint32)
    return np.concatenate([u, half, near, edge]).view(np.float32)


def test_float_to_bf16_matches_torch(rng):
    f = special_floats(rng)
    want = torch_bf16_round(f)
    got_np = O.float_to_bf16(f)
    L = O.lib()
    idx = rng.choice(len(f), 3000, replace=False)
    got_c = np.array([L.oracle_float_to_bf16(float(f[i])) for i in idx], np.uint16)
    nan = np.isnan(f)
    assert np.array_equal(got_np[~nan], want[~nan])
    assert np.all(np.isnan(O.bf16_to_float(got_np[nan])))
    keep = ~nan[idx]
    assert np.array_equal(got_c[keep], want[idx][keep])


def test_bf16_widen_exact():
    torch = pytest.importorskip("torch")
    b = np.arange(2 ** 16, dtype=np.uint32).astype(np.uint16)
    want = torch.from_numpy(b.view(np.int16)).view(torch.bfloat16).to(torch.float32).numpy()
    got = O.bf16_to_float(b)
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), want[~nan].view(np.uint32))
    L = O.lib()
    for v in (0x0000, 0x8000, 0x3F80, 0x7F7F, 0xFF7F, 0x0001, 0x7F80):
        assert np.float32(L.oracle_bf16_to_float(v)).view(np.uint32) == got[v].view(np.uint32)


# --------------------------------------------------------------------------
# Compensated Add: known answers and the Sum2 bound

KNOWN = [  # (float32 row, exact sum)
    ([1.0, 2.0 ** -30, -1.0], 2.0 ** -30),
    ([2.0 ** 24, 1.0, 1.0], 2.0 ** 24 + 2),
    ([1.0, 2.0 ** -25, 2.0 ** -25, 2.0 ** -25, 2.0 ** -25], 1.0 + 2.0 ** -23),
    ([3.0e7, 1.0, -3.0e7, 0.5], 1.5),
    ([1.0e-3, 1.0e3, -1.0e3], float(np.float32(1.0e-3))),
]


@pytest.mark.parametrize("row,exact", KNOWN)
def test_compensated_known_answers(row, exact):
    v = np.array(row, np.float32)
    got = row_sums(v.copy(), O.FLOAT_C)[0]
    assert got == np.float32(exact)
    plain = row_sums(v.copy(), O.FLOAT)[0]  # the planted mistake: no compensation
    assert plain != np.float32(exact)


def test_compensated_special_values():
    cases = [([-0.0], 0.0, False), ([np.inf, 1.0], np.inf, False), ([np.inf, -np.inf], np.nan, True),
             ([3.0e38, 3.0e38], np.inf, False), ([np.nan, 1.0], np.nan, True), ([-np.inf, 2.0], -np.inf, False)]
    for row, want, isnan in cases:
        got = row_sums(np.array(row, np.float32), O.FLOAT_C)[0]
        if isnan:
            assert np.isnan(got)
        else:
            assert got == np.float32(want) and np.signbit(got) == np.signbit(np.float32(want))


def random_rows(rng, nrows):
    rows = []
    for i in range(nrows):
        n = int(rng.choice([1, 2, 3, 4, 8, 8, 17, 50]))
        mag = 10.0 ** rng.uniform(-3, 3, n)
        x = (mag * rng.choice([-1.0, 1.0], n)).astype(np.float32)
        if i % 3 == 0 and n >= 2:  # heavy cancellation: pair up near-opposites
            x[1] = np.float32(-x[0] * (1 + rng.uniform(-1e-6, 1e-6)))
        rows.append(x)
    return rows


def test_compensated_bound_vs_exact_sums(rng):
    rows = random_rows(rng, 3000)
    v = np.concatenate(rows)
    cols = np.cumsum([0] + [len(r) for r in rows])
    rs = cols.astype(np.int32)
    ci = np.arange(len(v), dtype=np.int32)
    ref = O.OracleCsr(rs, ci, Ncols=len(v))
    comp = ref.gather(v.copy(), 1, O.ADD, O.TRANS, type=O.FLOAT_C)
    plain = ref.gather(v.copy(), 1, O.ADD, O.TRANS, type=O.FLOAT)
    viol_plain = 0
    rounded_ok = 0
    for r, x in enumerate(rows):
        S = sum(Fraction(float(t)) for t in x)
        A = sum(abs(Fraction(float(t))) for t in x)
        bound = U * abs(S) + Fraction(gamma(len(x) - 1)) ** 2 * A
        assert abs(Fraction(float(comp[r])) - S) <= bound, (r, x, comp[r], float(S))
        viol_plain += abs(Fraction(float(plain[r])) - S) > bound
        rounded_ok += comp[r] == np.float32(float(S))
    assert viol_plain > 0           # the bound does separate plain from compensated sums
    assert rounded_ok >= 0.99 * len(rows)  # short rows: almost always the rounded exact sum


def test_bf16_add_bound_vs_exact_sums(rng):
    rows = random_rows(rng, 2000)
    rows = [O.float_to_bf16(x) for x in rows]
    v = np.concatenate(rows)
    rs = np.cumsum([0] + [len(r) for r in rows]).astype(np.int32)
    ref = O.OracleCsr(rs, np.arange(len(v), dtype=np.int32), Ncols=len(v))
    got = O.bf16_to_float(ref.gather(v.copy(), 1, O.ADD, O.TRANS, type=O.BF16))
    for r, x in enumerate(rows):
        xs = O.bf16_to_float(x)
        S = sum(Fraction(float(t)) for t in xs)
        A = sum(abs(Fraction(float(t))) for t in xs)
        # float Sum2 bound, then one rounding to bf16 (u_bf16 = 2^-8 relative)
        f_bound = U * abs(S) + Fraction(gamma(len(xs) - 1)) ** 2 * A
        bound = f_bound + Fraction(2.0 ** -8) * (abs(S) + f_bound)
        assert abs(Fraction(float(got[r])) - S) <= bound
    assert O.bf16_to_float(row_sums(O.float_to_bf16(np.array([1.0, 2.0 ** -12, -1.0], np.float32)),
                                    O.BF16))[0] == np.float32(2.0 ** -12)


# --------------------------------------------------------------------------
# Min/Max exact; operator relations on a mesh

def test_bf16_minmax_exact(rng):
    rows = [rng.integers(0, 2 ** 16, int(rng.integers(1, 9)), dtype=np.uint64).astype(np.uint16) for _ in range(500)]
    rows = [r[~np.isnan(O.bf16_to_float(r))] for r in rows]
    rows = [r for r in rows if len(r)]
    v = np.concatenate(rows)
    rs = np.cumsum([0] + [len(r) for r in rows]).astype(np.int32)
    ref = O.OracleCsr(rs, np.arange(len(v), dtype=np.int32), Ncols=len(v))
    mx = O.bf16_to_float(ref.gather(v.copy(), 1, O.MAX, O.TRANS, type=O.BF16))
    mn = O.bf16_to_float(ref.gather(v.copy(), 1, O.MIN, O.TRANS, type=O.BF16))
    for r, x in enumerate(rows):
        xs = O.bf16_to_float(x)
        # compare-select from -max keeps the first of +-0 ties: compare values
        assert mx[r] == max(np.float32(-3.38953139e38), xs.max())
        assert mn[r] == min(np.float32(3.38953139e38), xs.min())
    # a NaN member is skipped by the compare-select
    nanrow = np.array([0x7FC0, 0x3F80], np.uint16)
    g = O.OracleCsr(np.array([0, 2], np.int32), np.array([0, 1], np.int32), Ncols=2)
    assert O.bf16_to_float(g.gather(nanrow.copy(), 1, O.MAX, O.TRANS, type=O.BF16))[0] == 1.0


def test_bf16_empty_rows_give_init():
    # Gather NoTrans over N rows with no positive member gives the op's init
    rsT = np.array([0, 2, 3], np.int32)
    ciT = np.array([0, 1, 2], np.int32)
    rsN = np.array([0, 0, 1], np.int32)
    ciN = np.array([2], np.int32)
    ref = O.OracleCsr(rsT, ciT, rsN, ciN, NrowsN=2, Ncols=3)
    v = O.float_to_bf16(np.array([1.0, 2.0, 3.0], np.float32))
    assert list(ref.gather(v.copy(), 1, O.MAX, O.NOTRANS, type=O.BF16)) == [0xFF7F, v[2]]
    assert list(ref.gather(v.copy(), 1, O.MIN, O.NOTRANS, type=O.BF16)) == [0x7F7F, v[2]]
    assert list(ref.gather(v.copy(), 1, O.ADD, O.NOTRANS, type=O.BF16)) == [0x0000, v[2]]
    fv = np.array([1.0, 2.0, 3.0], np.float32)
    assert list(ref.gather(fv.copy(), 1, O.MAX, O.NOTRANS, type=O.FLOAT_C)) == [np.float32(-3.4028235e38), 3.0]


@pytest.mark.parametrize("ty", [O.BF16, O.FLOAT_C])
@pytest.mark.parametrize("op", [O.ADD, O.MAX, O.MIN])
def test_scatter_gather_is_gs_and_fields_interleave(rng, ty, op):
    ids = O.make_ids(4, 3, 5, 3)
    g = O.OracleOgs(ids, O.UNSIGNED)
    K = 3
    f = (10.0 ** rng.uniform(-3, 3, len(ids) * K) * rng.choice([-1.0, 1.0], len(ids) * K)).astype(np.float32)
    v = O.float_to_bf16(f) if ty == O.BF16 else f
    gs = g.gather_scatter(v.copy(), K, op, O.SYM, type=ty)
    gv = g.op_gather(v.copy(), K, op, O.TRANS, type=ty)  # operator level: an Unsigned ogs has no Gather
    sc = g.op_scatter(np.zeros_like(v), gv, K, O.NOTRANS, type=ty)
    assert sc.tobytes() == gs.tobytes()
    vk = v.reshape(-1, K)
    for k in range(K):
        one = g.gather_scatter(np.ascontiguousarray(vk[:, k]), 1, op, O.SYM, type=ty)
        assert one.tobytes() == np.ascontiguousarray(gs.reshape(-1, K)[:, k]).tobytes()


def test_precision_types_reject_mul():
    ref = O.OracleCsr(np.array([0, 1], np.int32), np.array([0], np.int32), Ncols=1)
    for ty, dt in ((O.BF16, np.uint16), (O.FLOAT_C, np.float32)):
        with pytest.raises(O.OracleError) as e:
            ref.gather_scatter(np.ones(1, dt), 1, O.MUL, O.SYM, type=ty)
        assert e.value.code == -6
    ids = np.array([1, 2, 1], np.int32)
    q = np.arange(6, dtype=np.uint16)
    assert list(O.extract(ids, q, 2, type=O.BF16)) == [2, 3, 4, 5, 2, 3]
