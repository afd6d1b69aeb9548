"""The torch.autograd.Function over the rollout C ABI (sca2012_amd/torch_rollout.py):
theta.grad after loss.backward() is the adjoint's dL/dtheta, and a gradient step lowers
the loss. Runs tests/torch_rollout_check.py in a fresh process, where torch is imported
before the HIP library so that both use one HIP runtime."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_autograd_function_matches_adjoint():
    r = subprocess.run([sys.executable, os.path.join(HERE, "torch_rollout_check.py")],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=280)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-3000:]
