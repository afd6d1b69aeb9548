"""Parity at the benchmark sizes: the 3D-2 dam break (262,144 particles on 64^3 cells,
capacity 1M, a marching-cubes step on the 257^3 grid) and the 3D-4 collision (4,194,304
particles on 512^3 cells, capacity 16M, a marching-cubes step on the 513^3 grid) stepped
on the GPU and by the CPU oracle from the same seeded state, bit for bit: live counts,
split/merge decisions, every particle, the density grid and the triangle soup."""
import numpy as np
import pytest

import sca2012_amd as sa
from sca2012_amd import _abi, scenes
from oracle import orc

from helpers import assert_particles_equal, float_fields_equal

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("workload,steps,mc_step", [("adaptive_3d2_262k", 4, 3), ("adaptive_3d4_4m", 2, 1)])
def test_benchmark_workload_bitexact(workload, steps, mc_step):
    w = scenes.WORKLOADS[workload]
    p, a = w["build"]()
    print("workload", workload, len(a), "particles", flush=True)
    g = sa.SPHContext(w["variant"], p, w["capacity"])
    g.upload(a)
    o = orc.Oracle(w["variant"], p, w["capacity"])
    o.upload(a)
    for s in range(steps):
        f = w["flags"] | (sa.STEP_MC if s == mc_step else 0)
        rc, n = g.step_raw(f)
        assert rc == 0 and o.step(f) == 0
        assert n == o.num_particles(), "step %d: N %d vs %d" % (s, n, o.num_particles())
        c = g.counters()
        print("step", s, "N", n, "split", c["s_done"], "merge groups", c["groups"], flush=True)
    assert_particles_equal(g.download(), o.download(), what=workload)
    if mc_step is not None:
        assert float_fields_equal(g.get_f32(_abi.Q_DGRID), o.get_f32(orc.DGRID)), "density grid"
        tg = g.get_f32(_abi.Q_TRI)
        to = o.get_f32(orc.TRI)
        assert len(tg) == len(to) > 0
        assert np.array_equal(tg.view(np.uint32), to.view(np.uint32))
