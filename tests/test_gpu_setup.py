"""Setup (ogsSetup.cpp:69-776 at one rank) on ids that are not a structured
numbering: sparse 64-bit ids, extreme magnitudes, ragged multiplicities and
zeros.  The product's CSR and counts (built on the device) must equal the
oracle's for every kind."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def sparse_ids(seed, n=20000, groups=3000):
    rng = np.random.default_rng(seed)
    vals = rng.integers(1, 2 ** 62, groups, dtype=np.int64)
    vals[:3] = [2 ** 63 - 1, 1, 2 ** 40]          # extremes among the magnitudes
    w = rng.random(groups) ** 3                    # ragged multiplicities
    ids = rng.choice(vals, n, p=w / w.sum())
    ids[rng.random(n) < 0.07] = 0
    return ids


def csr_of(g):
    c = g.csr()
    return {k: c[k] for k in ("rowStartsT", "colIdsT", "rowStartsN", "colIdsN", "blockRowStartsT",
                              "blockRowStartsN")}


@pytest.mark.parametrize("seed", [0, 1])
@pytest.mark.parametrize("kind,unique", [(O.UNSIGNED, False), (O.SIGNED, False), (O.SIGNED, True)])
def test_sparse_ids_setup_matches_oracle(seed, kind, unique):
    import dawgs_amd as D
    ids = sparse_ids(seed)
    if kind == O.SIGNED:
        rng = np.random.default_rng(seed + 10)
        neg = rng.random(len(ids)) < 0.4
        ids[neg] = -ids[neg]
    o = O.OracleOgs(ids, kind, unique)
    dids = ids.copy()
    g = D.ogs_t()
    g.Setup(len(dids), dids, kind, D.Auto, unique, False)
    assert np.array_equal(dids, o.ids)  # unique: the same signs written back
    assert (g.N, g.Ngather, g.NgatherGlobal, g.NlocalP, g.NlocalT) == (o.N, o.Ngather, o.NgatherGlobal,
                                                                       o.NlocalP, o.NlocalT)
    assert g.gather_defined == o.gather_defined
    c = csr_of(g)
    for name, arr in c.items():
        assert np.array_equal(arr, getattr(o, name)), name


def test_extreme_magnitudes():
    """+-(2^63 - 1), the largest |id| an hlong holds, groups like any other
    (its key needs all 64 bits of the sort).  INT64_MIN is no valid id: the
    reference takes abs() of it (ogsSetup.cpp:115), which overflows."""
    import dawgs_amd as D
    big = 2 ** 63 - 1
    ids = np.array([-big, 5, big, 0, 5, 2 ** 62, -big, 1], dtype=np.int64)
    for kind in (O.UNSIGNED, O.SIGNED):
        o = O.OracleOgs(ids, kind)
        g = D.ogs_t()
        g.Setup(len(ids), ids.copy(), kind)
        c = csr_of(g)
        assert g.NlocalT == o.NlocalT == 4
        for name in ("rowStartsT", "colIdsT", "rowStartsN", "colIdsN"):
            assert np.array_equal(c[name], getattr(o, name)), (kind, name)


@pytest.mark.parametrize("kind,unique", [(O.UNSIGNED, False), (O.SIGNED, False), (O.SIGNED, True)])
def test_global_to_local_mapping(kind, unique):
    """SetupGlobalToLocalMapping (ogsSetup.cpp:867-886): a NoTrans Scatter of
    the row indices, so every member of row r maps to r and id 0 to -1."""
    import dawgs_amd as D
    rng = np.random.default_rng(3)
    ids = O.make_ids(5, 3, 4, 2).copy()
    ids[rng.random(len(ids)) < 0.1] = 0
    if kind == O.SIGNED:
        neg = rng.random(len(ids)) < 0.3
        ids[neg] = -ids[neg]
    o = O.OracleOgs(ids, kind, unique)
    g = D.ogs_t()
    g.Setup(len(ids), ids.copy(), kind, D.Auto, unique, False)
    want = np.full(len(ids), -1, dtype=np.int32)
    for r in range(o.NlocalT):
        want[o.colIdsT[o.rowStartsT[r]:o.rowStartsT[r + 1]]] = r
    got = g.SetupGlobalToLocalMapping()
    assert np.array_equal(got, want)
    assert np.all(got[ids == 0] == -1) and np.all(got[ids != 0] >= 0)
