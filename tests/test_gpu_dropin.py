"""The C++ drop-in SPHSystem (include/sph_system_amd.h), driven like the reference's GLUT
application by sca2012_amd/_bin/sph_dropin_run, against the CPU oracle: live counts per
step and a sequential float64 checksum of the particle positions in slot order."""
import os
import subprocess

import numpy as np
import pytest

import sca2012_amd as sa
from sca2012_amd import scenes
from oracle import orc

pytestmark = pytest.mark.gpu

DEMO = os.path.join(os.path.dirname(sa.__file__), "_bin", "sph_dropin_run")


def checksum(b):
    cs = 0.0
    for x, y, z in b["pos"].astype(np.float64):
        cs += x + 2.0 * y + 3.0 * z
    return "%.9e" % cs


@pytest.mark.parametrize("variant,steps", [(sa.V3D2, 12), (sa.V3D4, 6)])
def test_dropin_matches_oracle(variant, steps):
    assert os.path.exists(DEMO), "build the library first (__graft_entry__.build)"
    r = subprocess.run([DEMO, str(variant), str(steps), "1", "1", "4"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout[-2000:])
    assert r.returncode == 0
    lines = [l.split() for l in r.stdout.splitlines() if l.startswith("step")]
    p, a = scenes.reference_scene(variant)
    assert r.stdout.splitlines()[0] == "init %d" % len(a)
    o = orc.Oracle(variant, p, 1000000)
    o.upload(a)
    for s in range(steps):
        flags = sa.STEP_SPLIT | sa.STEP_MERGE | (sa.STEP_MC if (s + 1) % 4 == 0 else 0)
        assert o.step(flags) == 0
        n = o.num_particles()
        assert int(lines[s][3]) == n, "step %d" % s
        assert lines[s][5] == checksum(o.download()), "step %d" % s
        if flags & sa.STEP_MC:
            assert int(lines[s][7]) == o.lib.orc_count(o.h, orc.TRI) // 18
