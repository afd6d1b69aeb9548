"""Golden fixtures (tests/golden/fixtures.json, written by tests/golden/make_fixtures.py):
the CPU oracle must reproduce them exactly, and so must the HIP library (GPU test)."""
import json
import os

import pytest

import sca2012_amd as sa
from oracle import orc

from golden.make_fixtures import CASES, digest, run_case

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "fixtures.json")))


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_fixture(name):
    rec = run_case(*CASES[name])
    assert rec == FIX[name]


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_reproduces_fixture(name):
    variant, build, flags, steps = CASES[name]
    p, a = build()
    fx = FIX[name]
    with sa.SPHContext(variant, p, 200000) as g:
        g.upload(a)
        for s in range(steps):
            g.calc_hash()
            g.sort()
            g.cell_ranges()
            g.density()
            g.force()
            g.integrate()
            g.energy()
            S = done = groups = 0
            if flags & sa.STEP_SPLIT:
                S, done = g.split()
            if flags & sa.STEP_MERGE:
                groups = g.merge()
                g.compact()
            want = fx["steps"][s]
            got = {"n": g.num_particles(), "split_requested": S, "split_done": done,
                   "merge_groups": groups, "sha256": digest(g.download())}
            assert got == want, "%s step %d" % (name, s)
