"""GPU parity of a C++ consumer of libp::ogs::ogs_t (tests/cpp/ogs_consumer.cpp,
the API of the reference's include/ogs.hpp as a dawgs-style code calls it,
linked against libogs_hip.so): its numbering, GatherScatter (Add over
doubles, Max over ints), Signed-unique Gather Trans and Scatter NoTrans are
bit-exact against the oracle on the same inputs, and Gather on a non-unique
ogs throws GATHER_UNDEF through libp::exception."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "ogs_consumer")


@pytest.mark.parametrize("box", [(6, 5, 4, 3), (24, 24, 24, 3), (16, 16, 16, 7)])
def test_cpp_consumer_matches_oracle(tmp_path, box):
    assert os.path.exists(EXE), "build it with make -C tests/cpp"
    print(f"consumer {box}", flush=True)
    r = subprocess.run([EXE, *map(str, box), "7", str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "undef caught" in r.stdout, r.stdout
    rd = lambda name, dt: np.fromfile(os.path.join(tmp_path, name), dtype=dt)
    ids = rd("ids.i64", np.int64)
    assert np.array_equal(ids, O.make_ids(*box))
    N = len(ids)
    q = rd("q.f64", np.float64)
    o = O.OracleOgs(ids, O.UNSIGNED)
    assert f"N={N} Ngather={o.Ngather} NgatherGlobal={o.NgatherGlobal}" in r.stdout
    assert rd("gs.f64", np.float64).tobytes() == o.gather_scatter(q.copy(), 1, O.ADD, O.SYM).tobytes()
    idx = np.arange(N, dtype=np.int32)
    assert rd("gs_max.i32", np.int32).tobytes() == o.gather_scatter(idx, 1, O.MAX, O.SYM).tobytes()
    s = O.OracleOgs(ids, O.SIGNED, True)
    g = s.gather(q, 1, O.ADD, O.TRANS)
    assert rd("gather.f64", np.float64).tobytes() == g.tobytes()
    assert rd("scatter.f64", np.float64).tobytes() == s.scatter(np.zeros(N), g, 1, O.NOTRANS).tobytes()
