"""The C++ drop-in layer (include/shinobi/grid.hpp) used the way the
reference's engine drives its builder: its grid and hits equal the oracle's
for the same scene and rays."""
import json
import os
import subprocess

import numpy as np
import pytest

import shinobi_amd as sa

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "drop_in_example")


def _fnv(a):
    h = 1469598103934665603
    for x in np.asarray(a, np.uint32).tolist():
        h = ((h ^ x) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def test_cpp_drop_in_matches_oracle(oracle):
    assert os.path.exists(EXE), "build() compiles tests/cpp/drop_in_example.cpp"
    n, nr = 20000, 50000
    out = subprocess.run([EXE, str(n), str(nr)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = json.loads(out.stdout.strip().splitlines()[-1])
    soa = sa.gen_soup(17, n, 0.02)
    rays = sa.gen_random_rays(18, nr, 0.0, 1.0)
    og = oracle.build_twolevel(soa, lambda_leaf=2.0)
    prim, _ = og.trace(rays)
    assert got["res"] == list(og.res)
    assert got["leavesCount"] == og.C2 and got["instancesCount"] == og.P2
    assert got["cells"] == _fnv(og.top_cells)
    assert got["leaves"] == _fnv(og.leaves)
    assert got["indices"] == _fnv(og.indices)
    assert got["prim"] == _fnv(prim)
    assert got["hits"] == int((prim != sa.MISS).sum())
