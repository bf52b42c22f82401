"""Device-side posterior and HMC leapfrog, capturable in a HIP graph.

``DevicePosterior`` is ``FastPosterior.evaluate`` (forcepho/model.py:302-368,
sampling=True, no prior) on the device: z -> q = lower + sigmoid(z)(upper -
lower) (``BoundedTransform``, forcepho/model.py:596-636) -> lnlike and
gradient from the HIP kernels -> lnp = lnlike + ln|dq/dz|,
grad_z = grad_q dq/dz + (1 - 2 sigmoid(z)).  Patches are independent, so
lnp is returned per patch.

``HMC`` runs fixed-length trajectories (diagonal mass matrix, a step size
per patch, per-patch Metropolis acceptance) entirely on the device;
``capture()`` records one trajectory in a ``torch.cuda.CUDAGraph`` so that a
sampler replays it with one launch.  ``adapt()`` is the warm-up the
reference gets from littlemcmc's NUTS (forcepho/fitting.py:255-342: step
size tuned to a target acceptance, diagonal mass matrix adapted from the
warm-up draws): Nesterov dual averaging of each patch's step size (Hoffman &
Gelman 2014, Algorithm 5) and a regularised diagonal inverse mass matrix from
the draws of a middle window, after which the step size adaptation restarts.
Patches are independent chains, so every patch adapts its own step size.

The mass matrix is diagonal by default (``inv_mass`` per parameter) or,
with ``dense=True``, a full matrix per patch (forcepho/fitting.py:370-412,
``get_pot(full=True)``): the inverse mass blocks are kept padded to the
largest patch, [n_patches, m, m], momenta are drawn through the Cholesky
factor of each block's inverse, and the warm-up windows estimate each
patch's covariance (regularised toward the identity as Stan's dense
adaptation does).
"""

import numpy as np
import torch

__all__ = ["DevicePosterior", "HMC", "DenseMass", "segment_sum_index"]


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
        self.log_range = torch.log(self.range)
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
        # log |dq/dz| = log(range) + log(y (1 - y)), the latter written as
        # -|z| - 2 log(1 + exp(-|z|)) so that it stays finite where y (1 - y)
        # underflows (a parameter driven far against a bound)
        a = torch.abs(z)
        lnjac = self.log_range - a - 2.0 * torch.log1p(torch.exp(-a))
        lnp = ll + self.segment_sum(lnjac)
        grad = gq * jac + (1.0 - 2.0 * y)
        return lnp, grad


class DenseMass:
    """Per-patch dense inverse mass matrices over the flat parameter vector
    of a posterior (``seg_idx``/``seg_mask`` from :func:`segment_sum_index`)."""

    def __init__(self, post, inv_mass_blocks=None):
        self.idx = post.seg_idx
        self.mask = post.seg_mask
        P, m = self.idx.shape
        dev = self.mask.device
        eye = torch.eye(m, dtype=torch.float64, device=dev).expand(P, m, m)
        self.minv = eye.clone() if inv_mass_blocks is None else \
            torch.as_tensor(inv_mass_blocks, dtype=torch.float64, device=dev).clone()
        self.chol = torch.empty_like(self.minv)
        # flat parameter j sits at position pos[j] of the [P*m] padded blocks
        idx = self.idx.cpu().numpy()
        msk = self.mask.cpu().numpy() > 0
        pos = np.zeros(int(idx[msk].max()) + 1, dtype=np.int64)
        pp, kk = np.nonzero(msk)
        pos[idx[pp, kk]] = pp * m + kk
        self.pos = torch.as_tensor(pos, device=dev)
        self._pad_eye = torch.diag_embed(1.0 - self.mask)
        self.update(self.minv)

    def update(self, minv):
        """Set the inverse mass blocks; padded rows and columns become the
        identity, and the momentum factor chol(M) is recomputed."""
        mm = self.mask[:, :, None] * self.mask[:, None, :]
        self.minv.copy_(minv * mm + self._pad_eye)
        self.chol.copy_(torch.linalg.cholesky(torch.linalg.inv(self.minv)))

    def blocks(self, v):
        return v[self.idx] * self.mask

    def flat(self, b):
        return b.reshape(-1)[self.pos]

    def draw(self, like):
        e = torch.randn(self.idx.shape, dtype=torch.float64, device=like.device)
        return self.flat(torch.matmul(self.chol, e[..., None])[..., 0])

    def velocity(self, p):
        return self.flat(torch.matmul(self.minv, self.blocks(p)[..., None])[..., 0])


class HMC:
    """Fixed-length HMC with a diagonal (or, with ``dense``, a full
    per-patch) mass matrix, a step size per patch and per-patch acceptance."""

    def __init__(self, posterior, z0, step_size=1e-3, n_steps=10, seed=0, inv_mass=None,
                 jitter=0.2, dense=False):
        self.post = posterior
        # each trajectory draws its step sizes uniformly in eps (1 +- jitter),
        # so a fixed-length trajectory cannot lock onto a period of the target
        self.jitter = float(jitter)
        dev = posterior.lower.device
        self.n_steps = int(n_steps)
        self.z = torch.as_tensor(z0, dtype=torch.float64, device=dev).clone()
        lnp, grad = posterior(self.z)
        self.lnp = lnp.clone()
        self.grad = grad.clone()
        self.n_accept = torch.zeros_like(self.lnp)
        self.accept_prob = torch.zeros_like(self.lnp)   # of the last trajectory
        self.eps_patch = torch.full_like(self.lnp, float(step_size))
        self.inv_mass = (torch.ones_like(self.z) if inv_mass is None or dense else
                         torch.as_tensor(inv_mass, dtype=torch.float64, device=dev).clone())
        self.dense = DenseMass(posterior, inv_mass) if dense else None
        # the default device generator is the one torch graphs replay safely
        if dev.type == "cuda":
            torch.cuda.manual_seed(int(seed))
        else:
            torch.manual_seed(int(seed))
        self.graph = None

    # -- mass matrix ------------------------------------------------------
    def _draw_p(self):
        if self.dense is not None:
            return self.dense.draw(self.z)
        p = torch.randn(self.z.shape, dtype=torch.float64, device=self.z.device)
        return p * torch.rsqrt(self.inv_mass)

    def _velocity(self, p):
        return self.dense.velocity(p) if self.dense is not None else self.inv_mass * p

    def _kinetic(self, p):
        return 0.5 * self.post.segment_sum(p * self._velocity(p))

    @property
    def eps(self):
        """Step size of every parameter (its patch's)."""
        return self.eps_patch[self.post.param_patch]

    def leapfrog(self, z, p, grad, eps=None):
        """n_steps leapfrog steps from (z, p) with gradient ``grad`` at z."""
        if eps is None:
            eps = self.eps
        p = p + 0.5 * eps * grad
        lnp = None
        for i in range(self.n_steps):
            z = z + eps * self._velocity(p)
            lnp, grad = self.post(z)
            p = p + (eps if i < self.n_steps - 1 else 0.5 * eps) * grad
        return z, p, lnp, grad

    def trajectory(self):
        """One trajectory and per-patch Metropolis step; updates the state
        tensors in place (graph-capturable)."""
        post = self.post
        p0 = self._draw_p()
        h0 = -self.lnp + self._kinetic(p0)
        eps = self.eps_patch
        if self.jitter > 0.0:
            u = torch.rand(eps.shape, dtype=torch.float64, device=eps.device)
            eps = eps * (1.0 + self.jitter * (2.0 * u - 1.0))
        z1, p1, lnp1, grad1 = self.leapfrog(self.z, p0, self.grad, eps[post.param_patch])
        h1 = -lnp1 + self._kinetic(p1)
        dh = h0 - h1
        # a diverging trajectory (non-finite energy) is rejected
        dh = torch.where(torch.isfinite(dh), dh, torch.full_like(dh, -torch.inf))
        self.accept_prob.copy_(torch.exp(torch.clamp(dh, max=0.0)))
        u = torch.rand(self.lnp.shape, dtype=torch.float64, device=self.z.device)
        acc = torch.log(u) < dh
        acc_p = acc[post.param_patch]
        self.z.copy_(torch.where(acc_p, z1, self.z))
        self.grad.copy_(torch.where(acc_p, grad1, self.grad))
        self.lnp.copy_(torch.where(acc, lnp1, self.lnp))
        self.n_accept.add_(acc.to(torch.float64))

    def adapt(self, n_warmup, target_accept=0.8, gamma=0.05, t0=10.0, kappa=0.75,
              fast=(0.15, 0.1), first_window=25):
        """Warm-up in Stan's shape: dual-averaging step sizes (per patch)
        toward ``target_accept`` throughout; between an initial and a final
        fast phase (fractions ``fast`` of the warm-up), windows of doubling
        length (from ``first_window``) each end by setting the inverse mass
        matrix to the per-parameter variance of the window's draws (with a
        dense mass, each patch's covariance), shrunk toward 1e-3 of the
        patch's mean variance, and restarting the step size adaptation.  Ends with each patch's averaged step size.
        Runs eagerly; capture afterwards."""
        n_warmup = int(n_warmup)
        start = int(fast[0] * n_warmup)
        stop = n_warmup - int(fast[1] * n_warmup)
        ends, w, b = set(), int(first_window), start
        while b + w < stop:
            # a window that would leave less than twice its length is stretched
            if b + 3 * w > stop:
                break
            ends.add(b + w - 1)
            b += w
            w *= 2
        if stop > b:
            ends.add(stop - 1)

        def restart():
            mu = torch.log(10.0 * self.eps_patch)
            return mu, torch.zeros_like(mu), torch.zeros_like(mu), 0

        mu, hbar, log_eps_bar, m = restart()
        mean = torch.zeros_like(self.z)
        m2 = torch.zeros_like(self.z)
        if self.dense is not None:
            n_blocks, width = self.dense.idx.shape
            cmean = torch.zeros((n_blocks, width), dtype=torch.float64, device=self.z.device)
            cm2 = torch.zeros((n_blocks, width, width), dtype=torch.float64,
                              device=self.z.device)
        n_mass = 0
        for it in range(n_warmup):
            self.trajectory()
            m += 1
            a = self.accept_prob
            hbar = (1.0 - 1.0 / (m + t0)) * hbar + (target_accept - a) / (m + t0)
            log_eps = mu - (m ** 0.5) / gamma * hbar
            eta = m ** (-kappa)
            log_eps_bar = eta * log_eps + (1.0 - eta) * log_eps_bar
            self.eps_patch.copy_(torch.exp(torch.clamp(log_eps, -20.0, 3.0)))
            if start <= it < stop:
                n_mass += 1
                delta = self.z - mean
                mean += delta / n_mass
                m2 += delta * (self.z - mean)
                if self.dense is not None:
                    zb = self.dense.blocks(self.z)
                    db = zb - cmean
                    cmean += db / n_mass
                    cm2 += db[:, :, None] * (zb - cmean)[:, None, :]
                if it in ends and n_mass > 2:
                    var = m2 / (n_mass - 1)
                    scale = self.post.segment_sum(var) / self.post.segment_sum(
                        torch.ones_like(var))
                    w = n_mass / (n_mass + 5.0)
                    reg = 1e-3 * (5.0 / (n_mass + 5.0))
                    self.inv_mass.copy_(w * var + reg * scale[self.post.param_patch])
                    if self.dense is not None:
                        cov = cm2 / (n_mass - 1)
                        eye = torch.eye(cov.shape[-1], dtype=cov.dtype, device=cov.device)
                        self.dense.update(w * cov + reg * scale[:, None, None] * eye)
                        cmean.zero_()
                        cm2.zero_()
                    mu, hbar, log_eps_bar, m = restart()
                    mean.zero_()
                    m2.zero_()
                    n_mass = 0
        if m > 0:
            self.eps_patch.copy_(torch.exp(torch.clamp(log_eps_bar, -20.0, 3.0)))
        self.n_accept.zero_()

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
