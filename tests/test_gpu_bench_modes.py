"""The harness's reference modes on the GPU (bench.py run as the driver runs
it): the -cc correctness check against gslib on a non-cubic box prints the
reference's "Error = 0" lines, and the headline line on a non-cubic box."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def bench(*args):
    print("bench", *args, flush=True)
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                       timeout=300, cwd=ROOT)
    return r, json.loads(r.stdout.strip().splitlines()[-1])


def test_cc_mode_non_cubic():
    r, out = bench("--cc", "--nx", "9", "--ny", "7", "--nz", "5", "--p", "4")
    assert r.returncode == 0, r.stderr
    assert out["pass"] and out["config"]["local_dofs"] == 9 * 7 * 5 * 125
    assert len(out["errors"]) == 6
    assert r.stderr.count("Error = 0\n") == 6


def test_headline_non_cubic():
    r, out = bench("--nx", "30", "--ny", "24", "--nz", "18", "--p", "3", "--steps", "5", "--warmup", "2",
                   "--no-cpu", "--no-extras")
    assert r.returncode == 0, r.stderr
    c = out["config"]
    assert (c["nx"], c["ny"], c["nz"], c["p"]) == (30, 24, 18, 3)
    assert c["local_dofs"] == 30 * 24 * 18 * 64 and c["global_dofs"] == 91 * 73 * 55
    assert out["value"] > 0 and out["roofline"]["traffic"] is None
