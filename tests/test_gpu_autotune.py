"""The pair-list autotuner (sph_autotune / sph_set_lists and the batch's pair): every list
length gives bit-identical steps, tuning steps (a different candidate per step) change no
bits, and the tuner ends with a candidate length and a mean time for every candidate."""
import numpy as np
import pytest

import sca2012_amd as sa
from sca2012_amd import scenes

from helpers import assert_particles_equal

pytestmark = pytest.mark.gpu

SM = sa.STEP_SPLIT | sa.STEP_MERGE
FORCE_LISTS, ENERGY_LISTS = (24, 32, 48), (16, 24, 32)


def _run(v, p, a, flags, setup):
    with sa.SPHContext(v, p, 400000) as g:
        setup(g)
        g.upload(a)
        n = [g.step(f) for f in flags]
        return n, g.download(), g.tuned_lists()


@pytest.mark.parametrize("name", ["3d2_block_32k", "2d_ref", "3d4_ref"])
def test_every_list_length_and_tuning_give_the_same_bits(name):
    build = {
        "3d2_block_32k": lambda: (sa.V3D2,) + scenes.adaptive_3d2(n_side=32, cells=32),
        "2d_ref": lambda: (sa.V2D,) + scenes.reference_scene(sa.V2D),
        "3d4_ref": lambda: (sa.V3D4,) + scenes.reference_scene(sa.V3D4),
    }[name]
    v, p, a = build()
    flags = [SM] * 4 + [SM | (sa.STEP_MC if v != sa.V2D else 0)] + [SM] * 9
    n0, d0, t0 = _run(v, p, a, flags, lambda g: g.autotune(0))
    assert t0[:2] == (32, 16)  # untuned: the defaults
    for fl, el in zip(FORCE_LISTS, ENERGY_LISTS):
        n1, d1, t1 = _run(v, p, a, flags, lambda g: g.set_lists(fl, el))
        assert t1[:2] == (fl, el)
        assert n1 == n0
        assert_particles_equal(d1, d0, what="lists %d/%d" % (fl, el))
    n2, d2, t2 = _run(v, p, a, flags, lambda g: g.autotune(1, after=1))  # tunes in steps 2-7
    assert n2 == n0
    assert_particles_equal(d2, d0, what="autotuned")
    fl, el, best = t2
    print("tuned", fl, el, best, flush=True)
    assert fl in FORCE_LISTS and el in ENERGY_LISTS
    assert all(x > 0 for x in best["force"].values()) and all(x > 0 for x in best["energy"].values())
    # the fastest length, or the default when nothing beats it by more than 5%
    for k, chosen, default in (("force", fl, 32), ("energy", el, 16)):
        b = best[k]
        if chosen == default:
            assert min(b.values()) >= 0.95 * b[default]
        else:
            assert b[chosen] == min(b.values()) and b[chosen] < 0.95 * b[default]


def test_set_lists_rejects_other_lengths():
    p, a = scenes.reference_scene(sa.V3D2)
    with sa.SPHContext(sa.V3D2, p, 10000) as g:
        for fl, el in ((16, 16), (32, 48), (0, 0)):
            with pytest.raises(sa.SphError) as e:
                g.set_lists(fl, el)
            assert e.value.code == sa._abi.EINVAL


def test_batch_tuning_gives_the_same_bits():
    """The list tuner of stream steps (the graph steps' launch tuner is checked with them in
    tests/test_gpu_batch.py)."""
    sizes = [300, 1200, 77, 2500]
    tanks = [scenes.tank_3d2(int(n), scenes.SEED_BASE + 40, s) for s, n in enumerate(sizes)]
    caps = [4 * n + 8 for n in sizes]
    outs = []
    # the third run tunes in two rounds (steps 2-7, and 14-25: 12 steps after the first began)
    for setup in (lambda b: b.autotune(0), lambda b: b.set_lists(24, 32),
                  lambda b: b.autotune(2, after=1, every=12)):
        b = sa.SPHBatch(sa.V3D2, [p for p, _ in tanks], caps)
        b.set_graph(False)
        setup(b)
        for s, (_, a) in enumerate(tanks):
            b.upload(s, a)
        for k in range(26):
            b.step(SM)
        outs.append(([b.download(s) for s in range(len(sizes))], b.num_particles(), b.tuned_lists()))
        b.close()
    for d, n, t in outs[1:]:
        assert np.array_equal(n, outs[0][1])
        for s in range(len(sizes)):
            assert_particles_equal(d[s], outs[0][0][s], what="batch scene %d, lists %s" % (s, t[:2]))
    fl, el, best = outs[2][2]
    assert fl in FORCE_LISTS and el in ENERGY_LISTS
    assert all(x > 0 for x in best["force"].values())
