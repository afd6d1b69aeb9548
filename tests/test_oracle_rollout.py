"""Pins for the oracle's differentiable-rollout model: the learnable field (trilinear
sample of the node values, added to the force in integrate) and the density sample
(particle-radius poly6 splat at the sample points), against closed forms in float64."""
import numpy as np
import pytest

import sca2012_amd as sa
from sca2012_amd import scenes
from oracle import orc


def _params(box=(12, 10, 8)):
    p, _ = scenes.rollout_block(n_side=2, box=box)
    return p


def test_field_reproduces_a_linear_function():
    """Trilinear interpolation of nodes sampled from an affine field is exact: the velocity
    update of an isolated particle is (A p + b + g) dt / dens."""
    p = _params()
    W = np.array([p["world_width"][0], p["world_height"][0], p["world_length"][0]], dtype=np.float64)
    nf = 5
    A = np.array([[0.3, -0.2, 0.1], [0.05, 0.4, -0.3], [-0.1, 0.2, 0.25]])
    b = np.array([0.5, -1.0, 0.25])
    g = np.arange(nf) / (nf - 1.0)
    theta = np.zeros((nf, nf, nf, 3))  # [z, y, x, c]
    for iz in range(nf):
        for iy in range(nf):
            for ix in range(nf):
                x = np.array([g[ix], g[iy], g[iz]]) * W
                theta[iz, iy, ix] = A @ x + b
    pts = np.array([[3.3, 4.1, 2.2], [6.0, 5.0, 4.0], [W[0] - 1.5, 1.25, W[2] - 1.75]], dtype=np.float32)
    a = np.zeros(len(pts), dtype=sa.PARTICLE3D_DTYPE)
    a["pos"] = pts
    a["level"] = 1
    a["dens"] = 2.0
    o = orc.Oracle(sa.V3D2, p, 16, fp64=True)
    o.upload(a)
    o.set_field(theta.astype(np.float32), nf)
    o.integrate()
    st = o.state_f64()
    dt = float(np.float32(p["time_step"][0]))
    grav = np.array(p["gravity"][0], dtype=np.float64)
    theta32 = theta.astype(np.float32).astype(np.float64)
    for k, x in enumerate(pts.astype(np.float64)):
        f = A @ x + b
        np.testing.assert_allclose(st[k, 3:6], (f + grav) * dt / 2.0, rtol=2e-6, atol=1e-7)
    # a node value moves exactly the particles in that node's eight cells
    o2 = orc.Oracle(sa.V3D2, p, 16, fp64=True)
    o2.upload(a)
    t2 = theta.astype(np.float32).copy()
    t2[0, 0, 0] += 1.0
    o2.set_field(t2, nf)
    o2.integrate()
    moved = np.any(o2.state_f64()[:, 3:6] != st[:, 3:6], axis=1)
    assert list(moved) == [False, False, False]  # none of them lies in the origin's cell
    del theta32


def test_density_sample_single_particle():
    p = _params(box=(8, 8, 8))
    a = np.zeros(1, dtype=sa.PARTICLE3D_DTYPE)
    a["pos"][0] = (3.3, 4.05, 2.9)
    a["level"] = 1
    o = orc.Oracle(sa.V3D2, p, 4, fp64=True)
    o.upload(a)
    ns = 16
    s = o.density_sample(ns)
    W = 8.0
    c = (np.arange(ns) + 0.5) * (W / ns)
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    d2 = (X - 3.3) ** 2 + (Y - np.float32(4.05)) ** 2 + (Z - np.float32(2.9)) ** 2
    r2 = float(p["large_radius"][0]) ** 2
    m, k = float(p["large_mass"][0]), float(p["large_poly6_radius"][0])
    want = np.where(d2 < r2, m * k * (r2 - d2) ** 3, 0.0).ravel()
    np.testing.assert_allclose(s, want, rtol=1e-6, atol=1e-9)
    assert np.count_nonzero(s) == np.count_nonzero(want) > 0


def test_rollout_field_changes_only_the_velocity_update():
    """With the field on, one step differs from the plain step only through integrate: the
    density and force of the step are those of the plain step (fp32, bitwise)."""
    p, a = scenes.rollout_block(n_side=6, box=(10, 10, 10), gravity=0.0)
    nf = 4
    th = scenes.field_theta(nf, 3)
    runs = []
    for field in (False, True):
        o = orc.Oracle(sa.V3D2, p, len(a))
        o.upload(a)
        if field:
            o.set_field(th, nf)
        o.step(0)
        runs.append(o.download())
    for f in ("dens", "press", "acc", "surface"):
        assert np.array_equal(runs[0][f], runs[1][f]), f
    assert not np.array_equal(runs[0]["vel"], runs[1]["vel"])


# ---- the derivative builds and the topology replay (the references of the adjoint) ----
SM = sa.STEP_SPLIT | sa.STEP_MERGE


def _adaptive(steps):
    p, a = scenes.rollout_block(n_side=8, box=(12, 10, 10), time_step=0.03)
    return p, a, scenes.field_theta(4, 11), steps


def _run(p, a, th, steps, mode=False, topo=None, dth=None, dpos=None):
    o = orc.Oracle(sa.V3D2, p, 2 * len(a), fp64=mode)
    o.upload(a)
    if mode in ("dual", "dual32"):
        o.set_field_tangent(np.asarray(th, dtype=np.float64), dth, 4)
        if dpos is not None:
            o.set_state_tangent(dpos=dpos)
    else:
        o.set_field(th, 4)
    if topo is not None:
        o.topology_from(topo)
        o.topology("replay")
    else:
        o.topology("record")
    for _ in range(steps):
        assert o.step(SM) == 0
    return o


def test_derivative_builds_reproduce_their_values():
    """The float32- and float64-valued derivative builds step exactly as the float32 and
    float64 oracles (state, topology, density sample), whatever direction they carry."""
    p, a, th, steps = _adaptive(70)
    d = np.random.default_rng(0).standard_normal(th.size)
    for val_mode, dual_mode in ((False, "dual32"), (True, "dual")):
        o = _run(p, a, th if val_mode is False else th.astype(np.float64), steps, val_mode)
        od = _run(p, a, th, steps, dual_mode, dth=d)
        assert o.topology_stats() == od.topology_stats()
        assert np.array_equal(o.state_f64(), od.state_f64())
        v, t = od.density_sample_tangent(12)
        assert np.array_equal(o.density_sample(12), v) and np.any(t != 0)
    st = o.topology_stats()
    assert st["splits"] > 100 and st["merge_groups"] > 10, st


def test_topology_replay_reproduces_the_recorded_run():
    """Replaying a recorded topology in the same build gives the recorded run bit for bit;
    a run with a perturbed field under replay keeps the recorded particle count and moves
    its particles only slightly."""
    p, a, th, steps = _adaptive(70)
    rec = _run(p, a, th, steps)
    rep = _run(p, a, th, steps, topo=rec)
    assert np.array_equal(rec.state_f64(), rep.state_f64())
    d = np.random.default_rng(1).standard_normal(th.size).astype(np.float32)
    held = _run(p, a, th + np.float32(1e-3) * d, steps, topo=rec)
    assert held.num_particles() == rec.num_particles()
    dx = np.abs(held.state_f64()[:, 0:3] - rec.state_f64()[:, 0:3]).max()
    assert 0 < dx < 0.1, dx


@pytest.mark.parametrize("steps", [20, 70])
def test_float64_derivative_matches_central_differences(steps):
    """The float64 derivative build against central differences of the float64 oracle
    under the replayed topology: over 20 steps (336 splits) at eps 1e-4, and over 70 steps
    (464 splits, 43 merge groups, where larger steps cross kinks of the viscosity kernel) at
    eps 1e-6."""
    p, a, th, _ = _adaptive(steps)
    rec = _run(p, a, th, steps)
    ns = 12
    target = rec.density_sample(ns) * 0.9
    d = np.random.default_rng(3).standard_normal(th.size)
    v, t = _run(p, a, th, steps, "dual", topo=rec, dth=d).density_sample_tangent(ns)
    dl = float(np.sum(2.0 * (v - target) * t))
    eps = 1e-4 if steps == 20 else 1e-6
    th64 = th.astype(np.float64)

    def loss(x):
        return float(np.sum((_run(p, a, x, steps, True, topo=rec).density_sample(ns) - target) ** 2))

    fd = (loss(th64 + eps * d) - loss(th64 - eps * d)) / (2 * eps)
    assert abs(fd - dl) <= 1e-6 * abs(dl), (fd, dl)
