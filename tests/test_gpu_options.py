"""GPU parity of the interaction options inside the step, bit-exact against the CPU oracle:
the sphere obstacles (use_cylinder, SPH-3D-2/sph_system.cu:242-282), the keyboard push
(force_waiting, SPH-3D-2/sph_system.cu:228-235; force_waiting_left/right,
SPH-3D-4/sph_system.cu:229-242) and the 2D mouse force (SPH-2D/sph_system.cu:781-789).
Pushes and the mouse force last one step, as the reference clears them after each
animation() (SPH-3D-2/sph_system.cpp:403, SPH-3D-4/sph_system.cpp:451-452,
SPH-2D/sph_system.cpp:263)."""
import numpy as np
import pytest

import sca2012_amd as sa
from sca2012_amd import scenes
from oracle import orc

from helpers import assert_particles_equal

pytestmark = pytest.mark.gpu


def _cylinder_key(variant):
    return "flag1" if variant == sa.V3D2 else "flag2"


def _events_3d(variant, step):
    """Which push flags to raise before `step`."""
    keys = []
    if step % 4 == 1:
        keys.append("flag0")
    if variant == sa.V3D4 and step % 5 == 2:
        keys.append("flag1")
    return keys


@pytest.mark.parametrize("variant,steps", [(sa.V3D2, 30), (sa.V3D4, 16)])
def test_obstacles_and_push_bitexact(variant, steps):
    p, a = scenes.reference_scene(variant)
    p[_cylinder_key(variant)] = 1
    g = sa.SPHContext(variant, p, 400000)
    o = orc.Oracle(variant, p, 400000)
    g.upload(a)
    o.upload(a)
    flags = sa.STEP_SPLIT | sa.STEP_MERGE
    print("variant", variant, "steps", steps, flush=True)
    for s in range(steps):
        keys = _events_3d(variant, s)
        if keys:
            q = p.copy()
            for k in keys:
                q[k] = 1
            g.set_params(q)
            o.set_params(q)
        rc_g, n_g = g.step_raw(flags)
        assert rc_g == o.step(flags) == 0
        assert n_g == o.num_particles(), "step %d: N %d vs %d" % (s, n_g, o.num_particles())
        if keys:
            # the step consumed the push: the context's parameter block has it cleared
            back = g.get_params()
            assert all(back[k][0] == 0 for k in keys)
            assert back[_cylinder_key(variant)][0] == 1
        if s % 5 == 4 or s == steps - 1:
            assert_particles_equal(g.download(), o.download(), what="step %d" % s)


def test_mouse_force_2d_bitexact():
    p, a = scenes.dam_break_2d()
    g = sa.SPHContext(sa.V2D, p, 8192)
    o = orc.Oracle(sa.V2D, p, 8192)
    g.upload(a)
    o.upload(a)
    for s in range(30):
        if s % 3 == 0:
            q = p.copy()
            sa.add_extern_force(q, (0.2 + 0.01 * s, 0.15), (40.0, 25.0))
            q["extern_len"] = 0.1
            g.set_params(q)
            o.set_params(q)
        rc_g, n_g = g.step_raw(0)
        assert rc_g == o.step(0) == 0
        if s % 3 == 0:
            assert g.get_params()["force_waiting"][0] == 0
        if s % 5 == 4:
            assert_particles_equal(g.download(), o.download(), what="step %d" % s)
    # the force moved the fluid: compare with an undisturbed run
    h = sa.SPHContext(sa.V2D, p, 8192)
    h.upload(a)
    for s in range(30):
        h.step(0)
    assert not np.array_equal(h.download()["pos"], g.download()["pos"])
