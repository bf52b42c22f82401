"""NUTS driven by the HIP likelihood against the same NUTS driven by the FP64
oracle: four uneven patches as four chains, the same device generator seed,
step sizes and mass matrix (from a warm-up on the HIP path).  Every
transition of every chain must build the same tree (depth and leaf count,
the same multinomial and U-turn decisions) and reach the same draw, so the
HIP likelihood is accurate enough for every decision the sampler makes over
these transitions (forcepho/fitting.py:255-367, run_lmc, is the reference's
sampler caller)."""
import numpy as np
import pytest
import torch

import oracle
from forcepho_amd import synth
from forcepho_amd.batch import DeviceSplines, PatchBatch
from forcepho_amd.hmc import DevicePosterior
from forcepho_amd.nuts import NUTS
from helpers import make_small_patches

pytestmark = pytest.mark.gpu


class OracleBatch:
    """The lnlike_and_grad interface of a PatchBatch, computed for the same
    patches by the FP64 oracle on the host and returned as device tensors."""

    def __init__(self, batch, patches):
        self.device = batch.device
        self.theta_off = batch.theta_off
        self.param_patch = batch.param_patch
        self.patches = patches

    def lnlike_and_grad(self, q, splines=None):
        qh = q.detach().cpu().numpy()
        ll = np.empty(len(self.patches))
        g = np.empty_like(qh)
        for i, p in enumerate(self.patches):
            a, b = self.theta_off[i], self.theta_off[i + 1]
            ll[i], g[a:b] = oracle.lnlike_theta(p, qh[a:b])
        return (torch.as_tensor(ll, device=self.device),
                torch.as_tensor(g, device=self.device))


def test_nuts_transitions_equal_oracle_driven_nuts(hip):
    items = make_small_patches()
    patches = [p for p, _, _ in items]
    for p in patches:
        p.send_to_device()
    batch = PatchBatch(patches)
    splines = DeviceSplines()
    try:
        theta = np.concatenate([t for _, _, t in items])
        bounds = [synth.theta_bounds(t, len(p.band_N)) for p, _, t in items]
        lo = np.concatenate([b[0] for b in bounds])
        hi = np.concatenate([b[1] for b in bounds])
        post = DevicePosterior(batch, splines, lo, hi)
        post_ref = DevicePosterior(OracleBatch(batch, patches), None, lo, hi)
        z0 = post.inverse_transform(theta)

        # step sizes and a diagonal mass matrix from a warm-up on the HIP path
        warm = NUTS(post, z0, step_size=1e-4, max_treedepth=6, seed=21)
        warm.adapt(40)
        eps = warm.eps_patch.clone()
        inv_mass = warm.inv_mass.clone()
        start = warm.z.clone()

        runs = []
        for pst in (post, post_ref):
            s = NUTS(pst, start, max_treedepth=6, seed=5, inv_mass=inv_mass)
            s.eps_patch.copy_(eps)
            trace = []
            for _ in range(3):
                s.trajectory()
                trace.append((s.z.clone(), s.tree_depth.clone(), s.n_leapfrog.clone(),
                              s.lnp.clone()))
            runs.append(trace)
        torch.cuda.synchronize()

        for k, ((z, d, n, lp), (z_r, d_r, n_r, lp_r)) in enumerate(zip(*runs)):
            print(f"transition {k}: depth {d.cpu().numpy()} leaves {n.cpu().numpy()} "
                  f"max |dz| {float((z - z_r).abs().max()):.3e} "
                  f"max |dlnp| {float((lp - lp_r).abs().max()):.3e}", flush=True)
            assert torch.equal(d, d_r) and torch.equal(n, n_r), k
            assert float((z - z_r).abs().max()) <= 1e-6 * (1.0 + float(z_r.abs().max())), k
            assert float((lp - lp_r).abs().max()) <= 1e-6 * float(lp_r.abs().max()), k
        # the chains moved and the trees were built by U-turns, not the cap
        d = runs[0][-1][1]
        assert bool((d >= 1).all()) and bool((runs[0][-1][0] != start).any())
    finally:
        splines.free()
        batch.free()
        for p in patches:
            p.free()
