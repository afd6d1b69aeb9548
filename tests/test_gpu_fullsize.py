"""Parity at the benchmark sizes, bit for bit against the CPU oracle from the same seeded
state: live counts, split/merge decisions, every particle, the density grid and the
triangle soup.

* 3D-2 dam break: 262,144 particles on 64^3 cells, capacity 1M, stepped with the
  benchmark's own flag sequence (split and merge every step, marching cubes on the 257^3
  grid every 10th step) and the pair-list autotuner running as in the benchmark, for 30
  steps. This crosses the step where ~190k parents split (262k -> 465k -> 511k live), the
  first merges (from step 12) and three marching-cubes steps, i.e. the regime that the
  benchmark's timed window covers.
* 3D-4 collision: 4,194,304 particles on 512^3 cells, capacity 16M, a marching-cubes step
  on the 513^3 grid (the first step merges ~940k groups of four).
* 256 batched tanks (2k-64k particles each, 4.86M in total), every tank against its own
  oracle.
"""
import numpy as np
import pytest

import sca2012_amd as sa
from sca2012_amd import _abi, scenes
from oracle import orc

from helpers import assert_particles_equal, float_fields_equal

pytestmark = pytest.mark.gpu


def _compare_grids(g, o):
    assert float_fields_equal(g.get_f32(_abi.Q_DGRID), o.get_f32(orc.DGRID)), "density grid"
    tg = g.get_f32(_abi.Q_TRI)
    to = o.get_f32(orc.TRI)
    assert len(tg) == len(to) > 0
    assert np.array_equal(tg.view(np.uint32), to.view(np.uint32)), "triangle soup"


def test_dam_break_3d2_through_timed_regime():
    workload = "adaptive_3d2_262k"
    w = scenes.WORKLOADS[workload]
    p, a = w["build"]()
    print("workload", workload, len(a), "particles", flush=True)
    g = sa.SPHContext(w["variant"], p, w["capacity"])
    g.upload(a)
    o = orc.Oracle(w["variant"], p, w["capacity"])
    o.upload(a)
    steps, split_total, groups_total = 30, 0, 0
    checkpoints = (5, 13, 19)  # after the big split, after the first merges, an MC step
    for s in range(steps):
        f = scenes.step_flags(workload, s)
        rc, n = g.step_raw(f)
        assert rc == 0 and o.step(f) == 0, "step %d" % s
        assert n == o.num_particles(), "step %d: N %d vs %d" % (s, n, o.num_particles())
        c = g.counters()
        split_total += c["s_done"]
        groups_total += c["groups"]
        print("step", s, "N", n, "split", c["s_done"], "merge groups", c["groups"],
              "mc" if f & sa.STEP_MC else "", flush=True)
        if s in checkpoints or s == steps - 1:
            assert_particles_equal(g.download(), o.download(), what="%s step %d" % (workload, s))
            if f & sa.STEP_MC:
                _compare_grids(g, o)
    print("splits", split_total, "merge groups", groups_total, "tuned lists", g.tuned_lists()[:2],
          flush=True)
    assert split_total > 200000 and groups_total > 0


def test_collision_3d4_4m_bitexact():
    workload = "adaptive_3d4_4m"
    w = scenes.WORKLOADS[workload]
    p, a = w["build"]()
    print("workload", workload, len(a), "particles", flush=True)
    g = sa.SPHContext(w["variant"], p, w["capacity"])
    g.upload(a)
    o = orc.Oracle(w["variant"], p, w["capacity"])
    o.upload(a)
    for s in range(2):
        f = w["flags"] | (sa.STEP_MC if s == 1 else 0)
        rc, n = g.step_raw(f)
        assert rc == 0 and o.step(f) == 0
        assert n == o.num_particles(), "step %d: N %d vs %d" % (s, n, o.num_particles())
        c = g.counters()
        print("step", s, "N", n, "split", c["s_done"], "merge groups", c["groups"], flush=True)
    assert_particles_equal(g.download(), o.download(), what=workload)
    _compare_grids(g, o)


def test_batched_256_tanks_bitexact():
    """The batched workload itself (256 tanks, split and merge on, capacity 4x each tank):
    per-tank live counts after every step, and every tank's particles after the last step,
    against one oracle per tank."""
    w = scenes.WORKLOADS["batched_3d2_256"]
    sizes, tanks = w["build"]()
    caps = [int(w["capacity_factor"] * n) for n in sizes]
    print("batched_3d2_256:", len(tanks), "tanks,", int(sizes.sum()), "particles", flush=True)
    b = sa.SPHBatch(w["variant"], [p for p, _ in tanks], caps)
    oracles = []
    for s, (p, a) in enumerate(tanks):
        b.upload(s, a)
        o = orc.Oracle(w["variant"], p, caps[s])
        o.upload(a)
        oracles.append(o)
    flags = w["flags"]
    for k in range(4):
        rc, n_tot = b.step_raw(flags)
        assert rc == 0, "step %d: batch status %d" % (k, rc)
        assert [o.step(flags) for o in oracles] == [0] * len(oracles)
        n_b = b.num_particles()
        n_o = np.array([o.num_particles() for o in oracles])
        bad = np.nonzero(n_b != n_o)[0]
        assert len(bad) == 0, "step %d: tanks %s differ in N (batch %s, oracle %s)" % (
            k, bad[:8], n_b[bad[:8]], n_o[bad[:8]])
        assert n_tot == int(n_o.sum())
        c = b.scene_counts()
        print("step", k, "N", n_tot, "split", int(c[:, 1].sum()), "merge groups",
              int(c[:, 2].sum()), flush=True)
        assert np.all(c[:, 3] == 0)
    for s, o in enumerate(oracles):
        assert_particles_equal(b.download(s), o.download(), what="tank %d" % s)
