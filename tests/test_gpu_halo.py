"""halo_t at one rank (include/ogs.hpp:347-397) through the C-ABI: the
Halo-kind setup's counts and id check against the oracle's Setup of the same
kind (ogsSetup.cpp:55-64, 286-288), SetupFromGather (ogsSetup.cpp:888-914),
and Exchange / Combine, which at one rank move nothing (no node is shared
with another rank, ogsHalo.cpp:39-359 with zero halo rows)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

E_GATHER_UNDEF, E_HALO_POSITIVE, E_ARG = -3, -5, -1


@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available()
    return torch


def halo_ids(seed, box=(5, 4, 3, 3), zeros=0.1):
    """Flagged ids with exactly one positive member per group (the halo
    convention), some ids zero."""
    rng = np.random.default_rng(seed)
    ids = O.make_ids(*box).copy()
    ids[rng.random(len(ids)) < zeros] = 0
    a = np.abs(ids)
    order = rng.permutation(len(ids))
    seen = set()
    out = -a
    for n in order:  # a random member of each group is the positive one
        if a[n] and a[n] not in seen:
            out[n] = a[n]
            seen.add(a[n])
    return out.astype(np.int64)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_halo_setup_counts_match_oracle(seed):
    import dawgs_amd as D
    ids = halo_ids(seed)
    o = O.OracleOgs(ids, O.HALO)
    h = D.halo_t()
    h.Setup(len(ids), ids.copy())
    assert (h.N, h.Ngather, h.NlocalT, h.NhaloT, h.NlocalP, h.NhaloP, h.NgatherGlobal) == \
           (o.N, o.Ngather, o.NlocalT, o.NhaloT, o.NlocalP, o.NhaloP, o.NgatherGlobal)
    assert h.Nhalo == o.NhaloT - o.NhaloP == 0
    assert not h.gathered_halo


def test_halo_setup_rejects_groups_without_one_positive():
    import dawgs_amd as D
    from dawgs_amd._lib import OgsError
    ids = halo_ids(3)
    a = np.abs(ids[ids != 0])
    vals, counts = np.unique(a, return_counts=True)
    g = vals[counts >= 2][0]  # a shared node (a group of several members)
    two = ids.copy()
    two[np.abs(two) == g] = g  # every member positive
    with pytest.raises(O.OracleError):
        O.OracleOgs(two, O.HALO)
    with pytest.raises(OgsError) as e:
        D.halo_t().Setup(len(two), two.copy())
    assert e.value.code == E_HALO_POSITIVE
    none = ids.copy()
    none[np.abs(none) == g] = -g  # no positive member
    with pytest.raises(OgsError) as e:
        D.halo_t().Setup(len(none), none.copy())
    assert e.value.code == E_HALO_POSITIVE


DTYPES = ["float32", "float64", "int32", "int64", "bfloat16"]


@pytest.mark.parametrize("K", [1, 3])
@pytest.mark.parametrize("dt", DTYPES)
def test_halo_exchange_and_combine_move_nothing(torch, dt, K):
    import dawgs_amd as D
    ids = halo_ids(4)
    h = D.halo_t()
    h.Setup(len(ids), ids.copy())
    v = torch.arange(len(ids) * K, device="cuda").to(getattr(torch, dt))
    before = v.clone()
    h.Exchange(v, K)
    h.Combine(v, K)
    torch.cuda.synchronize()
    assert torch.equal(v, before)
    if dt == "float32":
        h.Exchange(v, K, type=D.FloatC)
        torch.cuda.synchronize()
        assert torch.equal(v, before)


def test_halo_setup_from_gather(torch):
    import dawgs_amd as D
    from dawgs_amd._lib import OgsError
    ids = O.make_ids(4, 3, 3, 3)
    sogs = D.ogs_t()
    sogs.Setup(len(ids), ids.copy(), D.Signed, D.Auto, True, False)
    h = D.halo_t()
    h.SetupFromGather(sogs)
    assert h.gathered_halo and not h.gather_defined
    assert (h.N, h.Ngather, h.NlocalT, h.NhaloT, h.NlocalP, h.NhaloP, h.NgatherGlobal, h.Nhalo) == \
           (sogs.NlocalT + sogs.NhaloT, sogs.Ngather, sogs.NlocalT, sogs.NhaloT, sogs.NlocalP, sogs.NhaloP,
            sogs.NgatherGlobal, 0)
    gv = torch.arange(h.N, dtype=torch.float64, device="cuda")
    before = gv.clone()
    h.Exchange(gv, 1)
    h.Combine(gv, 1)
    torch.cuda.synchronize()
    assert torch.equal(gv, before)
    # an ogs without a well-defined gather cannot carry a gathered halo
    ogs = D.ogs_t()
    ogs.Setup(len(ids), ids.copy(), D.Unsigned)
    with pytest.raises(OgsError) as e:
        D.halo_t().SetupFromGather(ogs)
    assert e.value.code == E_GATHER_UNDEF


def test_halo_argument_errors(torch):
    import dawgs_amd as D
    from dawgs_amd._lib import OgsError
    ids = halo_ids(5)
    h = D.halo_t()
    h.Setup(len(ids), ids.copy())
    v = torch.zeros(len(ids), dtype=torch.float64, device="cuda")
    with pytest.raises(OgsError) as e:
        h.Exchange(v, 0)
    assert e.value.code == E_ARG
    with pytest.raises(OgsError) as e:
        h.Combine(v, -1)
    assert e.value.code == E_ARG
