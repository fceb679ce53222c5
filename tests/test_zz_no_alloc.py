"""GPU check of ogs_setup_global_to_local_noalloc (the no-allocation variant
of ogs_setup_global_to_local, include/ogs_hip.h) through the C-ABI: two
calls in a row, the second on a smaller mesh, each equal to the allocating
entry point's map; a mesh above the fixed scratch is refused."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _setup(D, box, kind, unique, zero, seed):
    rng = np.random.default_rng(seed)
    ids = O.make_ids(*box).copy()
    ids[rng.random(len(ids)) < zero] = 0
    g = D.ogs_t()
    g.Setup(len(ids), ids, kind, D.Auto, unique, False)
    return g


def _noalloc(g):
    from dawgs_amd._lib import check, lib
    out = np.full(g.N, 12345, dtype=np.int32)
    check(lib().ogs_setup_global_to_local_noalloc(g.h, out.ctypes.data_as(C.c_void_p)),
          "ogs_setup_global_to_local_noalloc")
    return out


def test_noalloc_matches_allocating_entry_point_twice():
    import torch
    assert torch.cuda.is_available()
    import dawgs_amd as D
    big = _setup(D, (12, 10, 8, 3), D.Signed, False, 0.05, 1)      # 61,440 nodes, some ignored
    small = _setup(D, (5, 4, 3, 2), D.Unsigned, False, 0.2, 2)     # 1,620 nodes
    for g in (big, small):
        want = g.SetupGlobalToLocalMapping()
        got = _noalloc(g)
        assert np.array_equal(got, want)
        assert (got == -1).sum() > 0 and got.max() == g.NlocalT - 1


def test_noalloc_refuses_above_its_scratch():
    import dawgs_amd as D
    from dawgs_amd._lib import lib
    n = (1 << 22) + 1
    ids = np.arange(1, n + 1, dtype=np.int64)
    g = D.ogs_t()
    g.Setup(n, ids, D.Unsigned, D.Auto, False, False)
    out = np.zeros(n, dtype=np.int32)
    assert lib().ogs_setup_global_to_local_noalloc(g.h, out.ctypes.data_as(C.c_void_p)) == -9
    assert not out.any()
