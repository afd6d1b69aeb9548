"""GPU differentiable rollout (sph_adj_*), with and without split and merge: the forward is
bit-identical to the fp32 oracle stepping the same model (state, topology, density sample);
dL/dtheta and dL/d(initial positions) from the hand-written adjoint agree with the oracle's
forward-mode derivative along the same fp32 trajectory (fp32 values, float64 derivatives)
and with central differences of the float64 oracle replaying the fp32 run's topology (split
and merge decisions, wall clamps, surface-tension threshold) -- the model whose derivative
the adjoint takes; checkpointed and taped backward passes give the same bits."""
import numpy as np
import pytest

import sca2012_amd as sa
from sca2012_amd import scenes
from oracle import orc

from helpers import assert_particles_equal

pytestmark = pytest.mark.gpu

SM = sa.STEP_SPLIT | sa.STEP_MERGE
NF, NS = 4, 12


def fixed_setup(steps=10, n_side=6, box=(10, 10, 10), seed=11):
    """A fixed particle set away from the walls (no split, no merge)."""
    p, a = scenes.rollout_block(n_side=n_side, box=box, gravity=0.0)
    a["pos"] += np.float32(1.5)  # away from the walls: no clamp in the first steps
    a["vel"][:, 0] = np.float32(0.3)
    return p, a, 0, steps, scenes.field_theta(NF, seed), scenes.field_theta(NF, seed + 1)


def adaptive_setup(steps=70, seed=11):
    """An adaptive block falling into a small tank: 512 particles, 70 steps of dt 0.03 with
    464 splits and 43 merge groups (split parents merge again after settling)."""
    p, a = scenes.rollout_block(n_side=8, box=(12, 10, 10), time_step=0.03)
    return p, a, SM, steps, scenes.field_theta(NF, seed), scenes.field_theta(NF, seed + 1)


def oracle_run(p, a, flags, steps, theta, mode=False, topo=None, tangent=None):
    """The oracle stepping the rollout model. mode: False fp32, True fp64, "dual32" /
    "dual" with tangent = (dtheta, dpos) along which derivatives are carried; topo: an
    oracle whose recorded topology is replayed (else the run records its own)."""
    o = orc.Oracle(sa.V3D2, p, 2 * len(a), fp64=mode)
    o.upload(a)
    if mode in ("dual", "dual32"):
        dth, dpos = tangent
        o.set_field_tangent(np.asarray(theta, dtype=np.float64), dth, NF)
        o.set_state_tangent(dpos=dpos)
    else:
        o.set_field(theta, NF)
    if topo is not None:
        o.topology_from(topo)
        o.topology("replay")
    else:
        o.topology("record")
    for _ in range(steps):
        assert o.step(flags) == 0
    return o


def make_target(p, a, flags, steps, theta_t):
    return oracle_run(p, a, flags, steps, theta_t).density_sample(NS).astype(np.float32)


def gpu_adjoint(p, a, flags, steps, theta, target, ck=0):
    adj = sa.SPHAdjoint(p, len(a), NF, NS, checkpoint_every=ck, flags=flags)
    adj.upload(a)
    adj.set_target(target)
    loss = adj.forward(theta, steps)
    g = np.zeros_like(theta)
    adj.backward(g)
    return adj, loss, g


def loss_of(o, target):
    s = o.density_sample(NS)
    return float(np.sum((s - target.astype(np.float64)) ** 2))


def tangent_loss(o, target):
    v, d = o.density_sample_tangent(NS)
    return float(np.sum(2.0 * (v - target.astype(np.float64)) * d))


@pytest.mark.parametrize("setup", [fixed_setup, adaptive_setup])
def test_forward_bitexact_with_oracle(setup):
    p, a, flags, steps, theta, theta_t = setup()
    target = make_target(p, a, flags, steps, theta_t)
    adj = sa.SPHAdjoint(p, len(a), NF, NS, flags=flags)
    adj.upload(a)
    adj.set_target(target)
    sample = np.zeros(NS ** 3, dtype=np.float32)
    loss = adj.forward(theta, steps, sample)
    o = oracle_run(p, a, flags, steps, theta)
    topo = adj.topology(steps)
    st = o.topology_stats()
    print("%s: %d -> %d particles, %d splits, %d merge groups" % (
        setup.__name__, len(a), adj.num_particles(), topo[:, 1].sum(), topo[:, 2].sum()), flush=True)
    assert adj.num_particles() == o.num_particles()
    assert (int(topo[:, 1].sum()), int(topo[:, 2].sum())) == (st["splits"], st["merge_groups"])
    if flags:
        assert st["splits"] > 100 and st["merge_groups"] > 10
    assert_particles_equal(adj.download(), o.download(), fields=["pos", "vel", "ev", "level"],
                           what="rollout state")
    s = o.density_sample(NS)
    assert np.array_equal(sample, s.astype(np.float32))
    ref = float(np.sum((s - target.astype(np.float64)) ** 2))
    assert loss == pytest.approx(ref, rel=1e-5)  # fp32 sum in a fixed tree order


@pytest.mark.parametrize("setup", [fixed_setup, adaptive_setup])
def test_gradient_matches_float32_trajectory_derivative(setup):
    """The adjoint's directional derivatives against the oracle's forward-mode derivative
    along the fp32 trajectory (the GPU forward's own states and decisions): a random
    direction and the three largest gradient components of the field, and one particle's
    initial position along each axis."""
    p, a, flags, steps, theta, theta_t = setup()
    target = make_target(p, a, flags, steps, theta_t)
    adj, loss, g = gpu_adjoint(p, a, flags, steps, theta, target)
    assert np.all(np.isfinite(g)) and np.abs(g).max() > 0
    rng = np.random.default_rng(5)
    dirs = [rng.standard_normal(theta.shape)]
    for k in np.argsort(-np.abs(g))[:3]:
        e = np.zeros(theta.shape)
        e[k] = 1.0
        dirs.append(e)
    zero_pos = np.zeros((len(a), 3))
    errs = []
    for d in dirs:
        o = oracle_run(p, a, flags, steps, theta, mode="dual32", tangent=(d, zero_pos))
        ref = tangent_loss(o, target)
        ad = float(np.dot(g.astype(np.float64), d))
        errs.append(abs(ad - ref) / abs(ref))
        print("d loss / d theta: adjoint %.7e  fp32-trajectory derivative %.7e  rel %.2e" % (
            ad, ref, errs[-1]), flush=True)
    dpos, dvel, dev = adj.state_grad()
    k = int(np.argmax(np.abs(dpos).sum(axis=1)))
    for ax in range(3):
        e = np.zeros((len(a), 3))
        e[k, ax] = 1.0
        o = oracle_run(p, a, flags, steps, theta, mode="dual32", tangent=(np.zeros(theta.size), e))
        ref = tangent_loss(o, target)
        err = abs(dpos[k, ax] - ref) / max(abs(ref), 1e-3 * np.abs(dpos).max())
        errs.append(err)
        print("d loss / d pos[%d,%d]: adjoint %.7e  derivative %.7e  rel %.2e" % (
            k, ax, dpos[k, ax], ref, err), flush=True)
    assert max(errs) <= 1e-3, errs


def test_adaptive_gradient_matches_float64_central_differences():
    """The adaptive model of a 16^3 block in a 24 x 16 x 16 tank, 200 steps of dt 0.01
    (3,863 splits, 201 merge groups): the adjoint against central differences (eps 1e-4)
    of the float64 oracle replaying the fp32 run's topology, and the float64 forward-mode
    derivative between them (the differences' own error, measured)."""
    p, a = scenes.rollout_block(n_side=16, box=(24, 16, 16), time_step=0.01)
    steps, theta, theta_t = 200, scenes.field_theta(NF, 31), scenes.field_theta(NF, 32)
    target = make_target(p, a, SM, steps, theta_t)
    adj, loss, g = gpu_adjoint(p, a, SM, steps, theta, target, ck=20)
    rec = oracle_run(p, a, SM, steps, theta)  # the fp32 topology, the GPU's
    st = rec.topology_stats()
    print("16^3 adaptive rollout: %d splits, %d merge groups" % (st["splits"], st["merge_groups"]),
          flush=True)
    assert st["splits"] > 1000 and st["merge_groups"] > 50
    d = np.random.default_rng(7).standard_normal(theta.shape)
    th64 = theta.astype(np.float64)
    eps = 1e-4
    lp = loss_of(oracle_run(p, a, SM, steps, th64 + eps * d, mode=True, topo=rec), target)
    lm = loss_of(oracle_run(p, a, SM, steps, th64 - eps * d, mode=True, topo=rec), target)
    fd = (lp - lm) / (2 * eps)
    o = oracle_run(p, a, SM, steps, theta, mode="dual", topo=rec, tangent=(d, np.zeros((len(a), 3))))
    tan64 = tangent_loss(o, target)
    ad = float(np.dot(g.astype(np.float64), d))
    rel = abs(ad - fd) / abs(fd)
    print("adjoint %.7e  fp64 central differences %.7e (rel %.2e)  fp64 derivative %.7e "
          "(differences vs derivative %.2e)" % (ad, fd, rel, tan64, abs(fd - tan64) / abs(tan64)),
          flush=True)
    assert abs(fd - tan64) <= 1e-4 * abs(tan64)
    assert rel <= 1e-3


@pytest.mark.parametrize("setup", [fixed_setup, adaptive_setup])
def test_checkpointed_backward_equals_tape_bit_for_bit(setup):
    p, a, flags, steps, theta, theta_t = setup()
    target = make_target(p, a, flags, steps, theta_t)
    out = []
    for ck in (0, 3, 0):
        adj, loss, g = gpu_adjoint(p, a, flags, steps, theta, target, ck=ck)
        out.append((loss, g, adj.download(), adj.state_grad()[0]))
    for loss, g, state, dpos in out[1:]:
        assert loss == out[0][0]
        assert np.array_equal(g.view(np.uint32), out[0][1].view(np.uint32))
        assert np.array_equal(dpos.view(np.uint32), out[0][3].view(np.uint32))
        assert_particles_equal(state, out[0][2], what="state after backward")


def test_rollout_errors():
    p, a, flags, steps, theta, _ = adaptive_setup(steps=2)
    b = a.copy()
    b["level"][3] = 2
    adj = sa.SPHAdjoint(p, len(a), NF, NS, flags=SM)
    with pytest.raises(sa.SphError):
        adj.upload(b)  # a level-2 particle with split on: the 2n capacity bound
    with pytest.raises(sa.SphError):
        adj.backward(np.zeros_like(theta))  # no forward yet
    adj0 = sa.SPHAdjoint(p, len(a), NF, NS, flags=sa.STEP_MERGE)
    adj0.upload(b)  # without split any levels are fine
    with pytest.raises(sa.SphError):
        sa.SPHAdjoint(p, len(a), NF, NS, flags=sa.STEP_MC)
    q = p.copy()
    q["flag1"] = 1  # obstacle spheres
    with pytest.raises(sa.SphError):
        sa.SPHAdjoint(q, len(a), NF, NS)


def test_full_size_rollout_gradient():
    """The benchmark rollout (32^3 particles falling into a 48 x 32 x 32 tank, split and
    merge on, 8^3 x 3 field, 64^3 sample, 200 steps of dt 0.01, checkpoint every 20): the
    adjoint's directional derivative against the forward-mode derivative along the fp32
    trajectory and against central differences (eps 1e-4) of the float64 oracle replaying
    the fp32 topology."""
    w = scenes.WORKLOADS["rollout_3d2_32k"]
    r = w["rollout"]
    nf, ns, steps, flags = r["field_nodes"], r["sample_n"], w["steps"], w["flags"]
    p, a = w["build"]()
    theta_t = scenes.field_theta(nf, scenes.SEED_BASE + 51)
    theta = scenes.field_theta(nf, scenes.SEED_BASE + 50)
    adj = sa.SPHAdjoint(p, len(a), nf, ns, checkpoint_every=r["checkpoint_every"], flags=flags)
    adj.upload(a)
    target = np.zeros(ns ** 3, dtype=np.float32)
    adj.forward(theta_t, steps, target)
    adj.set_target(target)
    adj.forward(theta, steps)
    topo = adj.topology(steps)
    print("full-size rollout: %d -> %d particles, %d splits, %d merge groups" % (
        len(a), adj.num_particles(), topo[:, 1].sum(), topo[:, 2].sum()), flush=True)
    assert topo[:, 1].sum() > 10000 and topo[:, 2].sum() > 100
    g = np.zeros_like(theta)
    adj.backward(g)
    assert np.all(np.isfinite(g))
    d = np.random.default_rng(1).standard_normal(theta.shape)

    def run(th, mode, topo_=None, tangent=None):
        o = orc.Oracle(sa.V3D2, p, 2 * len(a), fp64=mode)
        o.upload(a)
        if tangent is not None:
            o.set_field_tangent(np.asarray(th, dtype=np.float64), tangent, nf)
        else:
            o.set_field(th, nf)
        if topo_ is not None:
            o.topology_from(topo_)
            o.topology("replay")
        else:
            o.topology("record")
        for _ in range(steps):
            assert o.step(flags) == 0
        return o

    rec = run(theta, False)
    v, t = run(theta, "dual32", tangent=d).density_sample_tangent(ns)
    ref32 = float(np.sum(2.0 * (v - target.astype(np.float64)) * t))
    eps = 1e-4
    th64 = theta.astype(np.float64)

    def l64(o):
        return float(np.sum((o.density_sample(ns) - target.astype(np.float64)) ** 2))

    fd = (l64(run(th64 + eps * d, True, rec)) - l64(run(th64 - eps * d, True, rec))) / (2 * eps)
    ad = float(np.dot(g.astype(np.float64), d))
    print("full-size rollout: adjoint %.7e  fp32-trajectory derivative %.7e (rel %.2e)  "
          "fp64 central differences %.7e (rel %.2e)" % (
              ad, ref32, abs(ad - ref32) / abs(ref32), fd, abs(ad - fd) / abs(fd)), flush=True)
    assert abs(ad - ref32) <= 1e-3 * abs(ref32)
    assert abs(ad - fd) <= 1e-3 * abs(fd)
