"""CPU-side checks of the C-ABI library: it loads, exports every function
include/ogs_hip.h declares, and rejects bad arguments before touching the
device (no compute calls: this container has no GPU)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ogs_hip.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b((?:ogs|halo)_\w+)\s*\(", src)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for n in ("ogs_hip_plan_create", "ogs_hip_gather_scatter", "ogs_hip_gather", "ogs_hip_scatter",
              "ogs_hip_extract", "ogs_setup", "ogs_gather_scatter"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from dawgs_amd import _lib
    L = _lib.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    # and the Python signature table covers exactly the declared set
    assert sorted(_lib.SIGNATURES) == declared_functions()


def test_error_strings():
    from dawgs_amd import _lib
    L = _lib.lib()
    assert L.ogs_hip_error_string(0) == b"success"
    assert b"too large" in L.ogs_hip_error_string(-2)
    assert b"not well-defined" in L.ogs_hip_error_string(-3)


def test_argument_checks_need_no_device():
    from dawgs_amd import _lib
    L = _lib.lib()
    assert L.ogs_hip_plan_create(None, 0, 0, 0, None, None, None, None, None) == -1
    h = C.c_void_p()
    assert L.ogs_hip_plan_create(C.byref(h), 10, 5, 4, None, None, None, None, None) == -1  # NrowsN > NrowsT
    assert L.ogs_hip_gather_scatter(None, 1, 0, 0, 1, None, None) == -1
    assert L.ogs_hip_gather(None, 1, 0, 0, 1, None, None, None) == -1
    assert L.ogs_hip_scatter(None, 1, 0, 1, None, None, None) == -1
    assert L.ogs_hip_extract(9, 1, 1, None, None, None, None) == -1
    assert L.ogs_hip_extract(1, 0, 0, None, None, None, None) == -1
    assert L.ogs_gather_scatter(None, 1, 0, 0, 1, None, None) == -1
    assert L.ogs_setup(None, 0, None, 0, 0, 0, 0) == -1
    assert L.halo_setup(None, 0, None, 0, 0) == -1
    assert L.halo_setup_from_gather(None, None) == -1
    assert L.halo_exchange(None, 1, 1, None, None) == -1
    assert L.halo_combine(None, 1, 1, None, None) == -1
    assert L.halo_free(None) == -1


def test_python_mirror_enums_follow_reference():
    import dawgs_amd as D
    # include/ogs.hpp:179-200
    assert (D.Float, D.Double, D.Int32, D.Int64) == (0, 1, 2, 3)
    assert (D.Add, D.Mul, D.Max, D.Min) == (0, 1, 2, 3)
    assert (D.Sym, D.NoTrans, D.Trans) == (0, 1, 2)
    assert (D.Auto, D.Pairwise, D.CrystalRouter, D.AllToAll) == (0, 1, 2, 3)
    assert (D.Unsigned, D.Signed, D.Halo) == (0, 1, 2)


def test_make_ids_matches_oracle():
    import numpy as np
    from dawgs_amd.mesh import make_ids, mesh_counts
    from oracle import oracle as O
    for box in [(1, 1, 1, 1), (4, 3, 2, 3), (5, 5, 5, 7)]:
        a = make_ids(*box)
        assert np.array_equal(a, O.make_ids(*box))
        assert a.max() == mesh_counts(*box)["NgatherGlobal"]


def test_short_buffers_refused_before_launch():
    """The Python mirror checks buffer lengths against the operator's sizes
    before calling into the library (a short buffer would be indexed past
    its end on the device)."""
    import torch
    import dawgs_amd as D
    from dawgs_amd.ogs import Plan
    p = Plan.__new__(Plan)
    p.h, p.owned, p._keep = None, False, None
    p.Ncols, p.NrowsN, p.NrowsT = 10, 3, 4
    with pytest.raises(ValueError, match="v holds 9"):
        p.GatherScatter(torch.zeros(9), 1, D.Add, D.Sym)
    with pytest.raises(ValueError, match="v holds 19"):
        p.GatherScatter(torch.zeros(19), 2, D.Add, D.Sym)
    with pytest.raises(ValueError, match="gv holds 3 values, the operator needs 4"):
        p.Gather(torch.zeros(3), torch.zeros(10), 1, D.Add, D.Trans)
    with pytest.raises(ValueError, match="gv holds 2 values, the operator needs 3"):
        p.Gather(torch.zeros(2), torch.zeros(10), 1, D.Add, D.NoTrans)
    with pytest.raises(ValueError, match="gv holds 5 values, the operator needs 6"):
        p.Scatter(torch.zeros(20), torch.zeros(5), 2, D.Trans)
    with pytest.raises(ValueError, match="gv holds 7 values, the operator needs 8"):
        p.Scatter(torch.zeros(20), torch.zeros(7), 2, D.NoTrans)
    g = D.ogs_t()
    with pytest.raises(RuntimeError, match="not set up"):
        g.GatherScatter(torch.zeros(4), 1, D.Add, D.Sym)
    g.h = C.c_void_p(1)  # stands in for a set-up handle; never passed to the library here
    g.N, g.NrowsN, g.NrowsT = 8, 2, 5
    try:
        with pytest.raises(ValueError, match="v holds 15 values, the operator needs 16"):
            g.GatherScatter(torch.zeros(15), 2, D.Add, D.Sym)
        with pytest.raises(ValueError, match="gv holds 4 values, the operator needs 5"):
            g.Gather(torch.zeros(4), torch.zeros(8), 1, D.Add, D.Trans)
        with pytest.raises(ValueError, match="gv holds 1 values, the operator needs 2"):
            g.Scatter(torch.zeros(8), torch.zeros(1), 1, D.Trans)
    finally:
        g.h = None
