"""GPU differentiable rollout (sph_adj_*): the forward is bit-identical to the fp32 oracle
stepping the same model; dL/dtheta and dL/d(initial positions) from the hand-written
adjoint agree with central finite differences of the float64 oracle; checkpointed and
taped backward passes agree."""
import numpy as np
import pytest

import sca2012_amd as sa
from sca2012_amd import scenes
from oracle import orc

from helpers import assert_particles_equal

pytestmark = pytest.mark.gpu

NF, NS, STEPS = 4, 12, 10


def setup(n_side=6, box=(10, 10, 10), gravity=0.0, seed=11):
    p, a = scenes.rollout_block(n_side=n_side, box=box, gravity=gravity)
    a["pos"] += np.float32(1.5)  # away from the walls: no clamp in the first steps
    a["vel"][:, 0] = np.float32(0.3)
    theta = scenes.field_theta(NF, seed)
    target_theta = scenes.field_theta(NF, seed + 1)
    o = orc.Oracle(sa.V3D2, p, len(a))
    o.upload(a)
    o.set_field(target_theta, NF)
    for _ in range(STEPS):
        assert o.step(0) == 0
    target = o.density_sample(NS).astype(np.float32)
    return p, a, theta, target


def oracle_loss(p, a, theta, target, fp64=True, steps=STEPS):
    o = orc.Oracle(sa.V3D2, p, len(a), fp64=fp64)
    o.upload(a)
    o.set_field(theta, NF)
    for _ in range(steps):
        assert o.step(0) == 0
    s = o.density_sample(NS)
    return float(np.sum((s - target.astype(np.float64)) ** 2)), o, s


def test_forward_bitexact_with_oracle():
    p, a, theta, target = setup()
    adj = sa.SPHAdjoint(p, len(a), NF, NS)
    adj.upload(a)
    adj.set_target(target)
    sample = np.zeros(NS ** 3, dtype=np.float32)
    loss = adj.forward(theta, STEPS, sample)
    ref_loss, o, s = oracle_loss(p, a, theta, target, fp64=False)
    assert_particles_equal(adj.download(), o.download(), fields=["pos", "vel", "ev"], what="rollout")
    assert np.array_equal(sample, s.astype(np.float32))
    assert loss == pytest.approx(float(np.sum((sample.astype(np.float64) - target) ** 2)), rel=1e-12)
    assert loss == pytest.approx(ref_loss, rel=1e-12)


def _fd(p, a, theta, target, d, eps):
    lp, _, _ = oracle_loss(p, a, theta + eps * d, target)
    lm, _, _ = oracle_loss(p, a, theta - eps * d, target)
    return (lp - lm) / (2 * eps)


def test_field_gradient_matches_finite_differences():
    p, a, theta, target = setup()
    adj = sa.SPHAdjoint(p, len(a), NF, NS)
    adj.upload(a)
    adj.set_target(target)
    adj.forward(theta, STEPS)
    g = np.zeros_like(theta)
    adj.backward(g)
    assert np.all(np.isfinite(g)) and np.abs(g).max() > 0
    rng = np.random.default_rng(5)
    errs = []
    dirs = [rng.standard_normal(theta.shape)]
    for k in np.argsort(-np.abs(g))[:3]:
        e = np.zeros(theta.shape)
        e[k] = 1.0
        dirs.append(e)
    theta64 = theta.astype(np.float64)  # the float64 oracle keeps the perturbed values exactly
    for d in dirs:
        fd = _fd(p, a, theta64, target, d, 1e-4)
        ad = float(np.dot(g.astype(np.float64), d))
        errs.append(abs(ad - fd) / max(abs(fd), 1e-12))
        print("directional derivative: adjoint %.6e  fd %.6e  rel %.2e" % (ad, fd, errs[-1]), flush=True)
    assert max(errs) < 1e-2, errs


def test_state_gradient_matches_finite_differences():
    p, a, theta, target = setup()
    adj = sa.SPHAdjoint(p, len(a), NF, NS)
    adj.upload(a)
    adj.set_target(target)
    adj.forward(theta, STEPS)
    adj.backward(np.zeros_like(theta))
    dpos, dvel, dev = adj.state_grad()
    k = int(np.argmax(np.abs(dpos).sum(axis=1)))
    for ax in range(3):
        eps = 1e-3
        ap, am = a.copy(), a.copy()
        ap["pos"][k, ax] += np.float32(eps)
        am["pos"][k, ax] -= np.float32(eps)
        h = float(ap["pos"][k, ax]) - float(am["pos"][k, ax])
        lp, _, _ = oracle_loss(p, ap, theta, target)
        lm, _, _ = oracle_loss(p, am, theta, target)
        fd = (lp - lm) / h
        print("d loss / d pos[%d,%d]: adjoint %.6e fd %.6e" % (k, ax, dpos[k, ax], fd), flush=True)
        assert abs(dpos[k, ax] - fd) <= 2e-2 * max(abs(fd), 1e-3 * np.abs(dpos).max())


def test_checkpointed_backward_equals_tape():
    p, a, theta, target = setup()
    g = []
    for ck in (0, 3):
        adj = sa.SPHAdjoint(p, len(a), NF, NS, checkpoint_every=ck)
        adj.upload(a)
        adj.set_target(target)
        loss = adj.forward(theta, STEPS)
        d = np.zeros_like(theta)
        adj.backward(d)
        g.append((loss, d, adj.download()))
    assert g[0][0] == g[1][0]
    np.testing.assert_allclose(g[0][1], g[1][1], rtol=1e-6, atol=1e-7 * np.abs(g[0][1]).max())
    assert_particles_equal(g[0][2], g[1][2], fields=["pos", "vel", "ev"], what="state after backward")


def test_rollout_errors():
    p, a, theta, target = setup()
    b = a.copy()
    b["level"][3] = 2
    adj = sa.SPHAdjoint(p, len(a), NF, NS)
    with pytest.raises(sa.SphError):
        adj.upload(b)  # level-2 particle
    with pytest.raises(sa.SphError):
        adj.backward(np.zeros_like(theta))  # no forward yet
    q = p.copy()
    q["flag1"] = 1  # obstacle spheres
    with pytest.raises(sa.SphError):
        sa.SPHAdjoint(q, len(a), NF, NS)


def test_full_size_rollout_gradient_matches_finite_differences():
    """The benchmark rollout (32^3 particles against the floor, 8^3 x 3 field, 64^3 sample,
    200 steps of dt 0.005, checkpoint every 20): the adjoint's directional derivative
    against central differences of the float64 oracle (eps 1e-5; between eps 1e-5 and 1e-6
    the float64 differences agree to ~3e-4)."""
    w = scenes.WORKLOADS["rollout_3d2_32k"]
    r = w["rollout"]
    nf, ns, steps = r["field_nodes"], r["sample_n"], w["steps"]
    p, a = w["build"]()
    adj = sa.SPHAdjoint(p, len(a), nf, ns, checkpoint_every=r["checkpoint_every"])
    adj.upload(a)
    target = np.zeros(ns ** 3, dtype=np.float32)
    adj.forward(scenes.field_theta(nf, scenes.SEED_BASE + 51), steps, target)
    adj.set_target(target)
    theta = scenes.field_theta(nf, scenes.SEED_BASE + 50)
    adj.forward(theta, steps)
    g = np.zeros_like(theta)
    adj.backward(g)
    assert np.all(np.isfinite(g))
    d = np.random.default_rng(1).standard_normal(theta.shape)

    def loss64(th):
        o = orc.Oracle(sa.V3D2, p, len(a), fp64=True)
        o.upload(a)
        o.set_field(th, nf)
        for _ in range(steps):
            assert o.step(0) == 0
        return float(np.sum((o.density_sample(ns) - target.astype(np.float64)) ** 2))

    eps = 1e-5
    th64 = theta.astype(np.float64)
    fd = (loss64(th64 + eps * d) - loss64(th64 - eps * d)) / (2 * eps)
    ad = float(np.dot(g.astype(np.float64), d))
    rel = abs(ad - fd) / abs(fd)
    print("full-size rollout: adjoint %.6e fd %.6e rel %.2e" % (ad, fd, rel), flush=True)
    assert rel < 1e-2
