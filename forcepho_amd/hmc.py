"""Device-side posterior and HMC leapfrog, capturable in a HIP graph.

``DevicePosterior`` is ``FastPosterior.evaluate`` (forcepho/model.py:302-368,
sampling=True, no prior) on the device: z -> q = lower + sigmoid(z)(upper -
lower) (``BoundedTransform``, forcepho/model.py:596-636) -> lnlike and
gradient from the HIP kernels -> lnp = lnlike + ln|dq/dz|,
grad_z = grad_q dq/dz + (1 - 2 sigmoid(z)).  Patches are independent, so
lnp is returned per patch.

``HMC`` runs fixed-length trajectories (identity mass matrix, per-patch
Metropolis acceptance) entirely on the device; ``capture()`` records one
trajectory in a ``torch.cuda.CUDAGraph`` so that a sampler replays it with
one launch.  This is the loop that littlemcmc drives from Python in the
reference (forcepho/fitting.py:255-342).
"""

import numpy as np
import torch

__all__ = ["DevicePosterior", "HMC", "segment_sum_index"]


def segment_sum_index(theta_off):
    """[n_patches, max_len] gather index and mask for deterministic per-patch
    sums over contiguous parameter blocks."""
    theta_off = np.asarray(theta_off, dtype=np.int64)
    counts = np.diff(theta_off)
    width = int(counts.max())
    idx = theta_off[:-1, None] + np.arange(width)[None, :]
    mask = np.arange(width)[None, :] < counts[:, None]
    idx = np.where(mask, idx, 0)
    return idx, mask


class DevicePosterior:
    """lnp per patch and gradient in the unconstrained space z."""

    def __init__(self, batch, splines, lower, upper):
        dev = batch.device
        self.batch = batch
        self.splines = splines
        self.lower = torch.as_tensor(np.asarray(lower, dtype=np.float64), device=dev)
        self.upper = torch.as_tensor(np.asarray(upper, dtype=np.float64), device=dev)
        self.range = self.upper - self.lower
        idx, mask = segment_sum_index(batch.theta_off)
        self.seg_idx = torch.as_tensor(idx, device=dev)
        self.seg_mask = torch.as_tensor(mask, dtype=torch.float64, device=dev)
        self.param_patch = batch.param_patch

    def segment_sum(self, v):
        return (v[self.seg_idx] * self.seg_mask).sum(dim=1)

    def transform(self, z):
        return self.lower + torch.sigmoid(z) * self.range

    def inverse_transform(self, q):
        y = (torch.as_tensor(q, dtype=torch.float64, device=self.lower.device) - self.lower) \
            / self.range
        return -torch.log(1.0 / y - 1.0)

    def __call__(self, z):
        y = torch.sigmoid(z)
        q = self.lower + y * self.range
        jac = self.range * y * (1.0 - y)
        ll, gq = self.batch.lnlike_and_grad(q, self.splines)
        lnp = ll + self.segment_sum(torch.log(jac))
        grad = gq * jac + (1.0 - 2.0 * y)
        return lnp, grad


class HMC:
    """Fixed-length HMC with identity mass and per-patch acceptance."""

    def __init__(self, posterior, z0, step_size=1e-3, n_steps=10, seed=0):
        self.post = posterior
        dev = posterior.lower.device
        self.eps = float(step_size)
        self.n_steps = int(n_steps)
        self.z = torch.as_tensor(z0, dtype=torch.float64, device=dev).clone()
        lnp, grad = posterior(self.z)
        self.lnp = lnp.clone()
        self.grad = grad.clone()
        self.n_accept = torch.zeros_like(self.lnp)
        # the default device generator is the one torch graphs replay safely
        torch.cuda.manual_seed(int(seed))
        self.graph = None

    def leapfrog(self, z, p, grad):
        """n_steps leapfrog steps from (z, p) with gradient ``grad`` at z."""
        eps = self.eps
        p = p + 0.5 * eps * grad
        lnp = None
        for i in range(self.n_steps):
            z = z + eps * p
            lnp, grad = self.post(z)
            p = p + (eps if i < self.n_steps - 1 else 0.5 * eps) * grad
        return z, p, lnp, grad

    def trajectory(self):
        """One trajectory and per-patch Metropolis step; updates the state
        tensors in place (graph-capturable)."""
        post = self.post
        p0 = torch.randn(self.z.shape, dtype=torch.float64, device=self.z.device)
        h0 = -self.lnp + 0.5 * post.segment_sum(p0 * p0)
        z1, p1, lnp1, grad1 = self.leapfrog(self.z, p0, self.grad)
        h1 = -lnp1 + 0.5 * post.segment_sum(p1 * p1)
        u = torch.rand(self.lnp.shape, dtype=torch.float64, device=self.z.device)
        acc = torch.log(u) < (h0 - h1)
        acc_p = acc[post.param_patch]
        self.z.copy_(torch.where(acc_p, z1, self.z))
        self.grad.copy_(torch.where(acc_p, grad1, self.grad))
        self.lnp.copy_(torch.where(acc, lnp1, self.lnp))
        self.n_accept.add_(acc.to(torch.float64))

    def capture(self, warmup=2):
        """Record one trajectory in a CUDA/HIP graph (after eager warm-up on
        a side stream, as torch.cuda.graphs requires)."""
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(warmup):
                self.trajectory()
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.trajectory()
        return self.graph

    def run(self, n_trajectories):
        for _ in range(int(n_trajectories)):
            if self.graph is not None:
                self.graph.replay()
            else:
                self.trajectory()
