"""torch.autograd.Function over the differentiable-rollout C ABI (sph_adj_*).

    adj = SPHAdjoint(params, n, field_nodes, sample_n, checkpoint_every)
    adj.upload(particles); adj.set_target(target)
    loss = rollout_loss(theta, adj, steps)     # theta: (nf, nf, nf, 3) float32 on the GPU
    loss.backward()                            # theta.grad = dL/dtheta from the HIP adjoint

The forward and backward passes run on the library's own HIP stream; torch's current
stream is synchronised before each call so that theta is complete, and the library
synchronises its stream before returning.
"""
import torch


class SPHRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, adj, steps):
        theta = theta.detach().contiguous()
        if theta.dtype != torch.float32 or not theta.is_cuda:
            raise TypeError("theta must be a float32 tensor on the GPU")
        torch.cuda.current_stream().synchronize()
        loss = adj.forward(theta, steps)
        ctx.adj = adj
        ctx.theta_shape = theta.shape
        ctx.device = theta.device
        return torch.tensor(loss, dtype=torch.float64, device=theta.device)

    @staticmethod
    def backward(ctx, grad_loss):
        d = torch.empty(ctx.theta_shape, dtype=torch.float32, device=ctx.device)
        torch.cuda.current_stream().synchronize()
        ctx.adj.backward(d)
        return d * grad_loss.to(torch.float32), None, None


def rollout_loss(theta, adj, steps):
    """Loss of a `steps`-step rollout with field theta, differentiable in theta."""
    return SPHRollout.apply(theta, adj, steps)
