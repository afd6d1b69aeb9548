"""Run by tests/test_gpu_torch_rollout.py in a fresh process: torch is imported before
the HIP library, so both share torch's HIP runtime. Prints one result line; exits 0 on
success."""
import os
import sys

import numpy as np
import torch

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import sca2012_amd as sa  # noqa: E402
from sca2012_amd import scenes  # noqa: E402
from sca2012_amd.torch_rollout import rollout_loss  # noqa: E402


def main():
    if not torch.cuda.is_available():
        print("torch sees no GPU")
        return 2
    nf, ns, steps = 4, 12, 8
    p, a = scenes.rollout_block(n_side=6, box=(10, 10, 10), gravity=0.0)
    a["pos"] += np.float32(1.5)
    adj = sa.SPHAdjoint(p, len(a), nf, ns)
    adj.upload(a)
    target = np.zeros(ns ** 3, dtype=np.float32)
    adj.forward(scenes.field_theta(nf, 21), steps, target)
    adj.set_target(torch.tensor(target, device="cuda"))
    theta0 = scenes.field_theta(nf, 22)
    ref_loss = adj.forward(theta0, steps)
    ref_grad = np.zeros_like(theta0)
    adj.backward(ref_grad)

    theta = torch.tensor(theta0.reshape(nf, nf, nf, 3), device="cuda", requires_grad=True)
    loss = rollout_loss(theta, adj, steps)
    if loss.item() != ref_loss:
        print("loss %r != %r" % (loss.item(), ref_loss))
        return 1
    loss.backward()
    g = theta.grad.detach().cpu().numpy().reshape(-1)
    err = float(np.abs(g - ref_grad).max() / np.abs(ref_grad).max())
    if err > 1e-6:
        print("gradient differs: %g" % err)
        return 1
    with torch.no_grad():
        step = 0.05 / float(theta.grad.abs().max())
        theta2 = (theta - step * theta.grad).detach().requires_grad_(True)
    l2 = rollout_loss(theta2, adj, steps).item()
    if not l2 < ref_loss:
        print("gradient step did not lower the loss: %g -> %g" % (ref_loss, l2))
        return 1
    print("OK loss %.6g -> %.6g, gradient max rel diff %.2g" % (ref_loss, l2, err))
    return 0


if __name__ == "__main__":
    sys.exit(main())
