"""No-U-Turn sampling of many independent patches in lockstep.

The reference samples each patch with littlemcmc's NUTS
(forcepho/fitting.py:255-367, ``run_lmc``; max tree depth 10, diagonal or
dense mass matrix from ``get_pot``).  Here every patch of a batch is its own
chain and all chains advance together, one batched posterior evaluation
(all patches in one kernel launch) per leapfrog step:

* the tree of each chain is built by doubling (Hoffman & Gelman 2014,
  Algorithm 3, iteratively): at depth j a random direction is drawn per
  chain and a subtree of 2^j leapfrog steps is built from that end;
* samples are drawn multinomially (Betancourt 2017, "A Conceptual
  Introduction to Hamiltonian Monte Carlo", appendix A): uniform
  progressive sampling inside a subtree, biased progressive sampling when a
  subtree is merged into the tree;
* the generalised U-turn criterion is checked on the whole tree after
  every merge and, inside a subtree, on every sub-subtree that ends at the
  newest leaf, with the momentum checkpoints of the iterative formulation
  (Phan, Pradhan & Jankowiak 2019, "Composable effects for flexible and
  accelerated probabilistic programming in NumPyro", section 2.2);
* a leaf whose energy error exceeds ``max_delta_h`` marks a divergence,
  which stops that chain's tree without merging the subtree.

A chain whose tree has stopped keeps its state while the others continue;
the batch stops doubling when every chain has stopped or at
``max_treedepth``.  The step-size and mass-matrix warm-up is
:meth:`forcepho_amd.hmc.HMC.adapt`, driven by the mean acceptance statistic
of each tree's leaves.  The mass matrix is diagonal or, with
``dense=True``, a full matrix per patch (``full=True`` of ``run_lmc``).
"""

import numpy as np
import torch

from .hmc import HMC

__all__ = ["NUTS", "run_nuts"]


def _ckpt_range(n):
    """Checkpoint slots of the sub-subtrees that end at leaf n (odd n):
    (first, last) -- last = number of set bits of n >> 1, and as many
    subtrees end at n as n has trailing one bits."""
    idx_max = bin(n >> 1).count("1")
    trailing = (~n & (n + 1)).bit_length() - 1
    return idx_max - trailing + 1, idx_max


class NUTS(HMC):
    """Multinomial NUTS for the chains of a :class:`DevicePosterior` (one
    chain per patch), with a diagonal (``inv_mass``) or, with ``dense``, a
    full per-patch inverse mass matrix.

    ``trajectory()`` draws one sample per chain; ``tree_depth`` and
    ``n_leapfrog`` hold the last tree's depth and leaf count per chain,
    ``divergent`` whether it diverged, ``accept_prob`` the mean acceptance
    statistic of its leaves, and ``n_accept`` counts the transitions that
    moved the chain.  Runs eagerly (the tree length depends on the
    draws), so :meth:`capture` is not available.
    """

    def __init__(self, posterior, z0, step_size=1e-3, max_treedepth=10, seed=0,
                 inv_mass=None, jitter=0.0, max_delta_h=1000.0, dense=False):
        super().__init__(posterior, z0, step_size=step_size, n_steps=1, seed=seed,
                         inv_mass=inv_mass, jitter=jitter, dense=dense)
        self.max_treedepth = int(max_treedepth)
        self.max_delta_h = float(max_delta_h)
        self.tree_depth = torch.zeros_like(self.lnp)
        self.n_leapfrog = torch.zeros_like(self.lnp)
        self.divergent = torch.zeros(self.lnp.shape, dtype=torch.bool, device=self.lnp.device)

    def capture(self, warmup=2):
        raise NotImplementedError("NUTS trees have data-dependent length; run eagerly")

    # -- helpers --------------------------------------------------------
    def _turning(self, p_start, p_end, rho):
        """Generalised U-turn criterion per chain with the end momenta
        counted half (trapezoidal rule): turning when the momentum sum,
        mapped by M^-1, points against either end's momentum."""
        r = rho - 0.5 * (p_start + p_end)
        seg = self.post.segment_sum
        return (seg(self._velocity(p_start) * r) <= 0) | (seg(self._velocity(p_end) * r) <= 0)

    def _leapfrog1(self, z, p, grad, eps):
        p = p + 0.5 * eps * grad
        z = z + eps * self._velocity(p)
        lnp, grad = self.post(z)
        p = p + 0.5 * eps * grad
        return z, p, lnp, grad

    def trajectory(self):
        """One NUTS transition of every chain (in place)."""
        post = self.post
        dev = self.z.device
        pp = post.param_patch
        f64 = torch.float64
        nchain = self.lnp.shape[0]

        p0 = self._draw_p()
        h0 = -self.lnp + self._kinetic(p0)
        eps_c = self.eps_patch
        if self.jitter > 0.0:
            u = torch.rand(eps_c.shape, dtype=f64, device=dev)
            eps_c = eps_c * (1.0 + self.jitter * (2.0 * u - 1.0))

        # the tree: its two ends, its sample, log total weight, momentum sum
        zl, pl, gl = self.z.clone(), p0.clone(), self.grad.clone()
        zr, pr, gr = self.z.clone(), p0.clone(), self.grad.clone()
        zs, gs, lps = self.z.clone(), self.grad.clone(), self.lnp.clone()
        log_w = torch.zeros(nchain, dtype=f64, device=dev)
        rho = p0.clone()
        done = torch.zeros(nchain, dtype=torch.bool, device=dev)
        divergent = torch.zeros_like(done)
        depth = torch.zeros(nchain, dtype=f64, device=dev)
        acc_sum = torch.zeros(nchain, dtype=f64, device=dev)
        n_leaf = torch.zeros(nchain, dtype=f64, device=dev)
        ninf = torch.full((nchain,), -torch.inf, dtype=f64, device=dev)

        for j in range(self.max_treedepth):
            active = ~done
            if not bool(active.any()):
                break
            fwd = torch.rand(nchain, dtype=f64, device=dev) < 0.5
            fwd_p = fwd[pp]
            eps = torch.where(fwd, eps_c, -eps_c)[pp]
            # subtree grows from the tree's end in the chosen direction
            ze = torch.where(fwd_p, zr, zl)
            pe = torch.where(fwd_p, pr, pl)
            ge = torch.where(fwd_p, gr, gl)
            sub_z, sub_g, sub_lp = ze.clone(), ge.clone(), self.lnp.clone()
            sub_log_w = ninf.clone()
            sub_rho = torch.zeros_like(p0)
            sub_stop = torch.zeros_like(done)      # turned or diverged
            sub_div = torch.zeros_like(done)
            ck_p = torch.zeros((j + 1,) + p0.shape, dtype=f64, device=dev)
            ck_rho = torch.zeros_like(ck_p)
            for n in range(2 ** j):
                live = active & ~sub_stop
                live_p = live[pp]
                z1, p1, lp1, g1 = self._leapfrog1(ze, pe, ge, eps)
                ze = torch.where(live_p, z1, ze)
                pe = torch.where(live_p, p1, pe)
                ge = torch.where(live_p, g1, ge)
                dh = (-lp1 + self._kinetic(p1)) - h0
                bad = ~torch.isfinite(dh) | (dh > self.max_delta_h)
                dh = torch.where(torch.isfinite(dh), dh, torch.full_like(dh, torch.inf))
                acc_sum += torch.where(live, torch.exp(torch.clamp(-dh, max=0.0)),
                                       torch.zeros_like(dh))
                n_leaf += live.to(f64)
                newdiv = live & bad
                sub_div |= newdiv
                sub_stop |= newdiv
                ok = live & ~bad
                # uniform progressive sampling inside the subtree
                lw = -dh
                new_log_w = torch.logaddexp(sub_log_w, torch.where(ok, lw, ninf))
                u = torch.rand(nchain, dtype=f64, device=dev)
                take = ok & (torch.log(u) < lw - new_log_w)
                take_p = take[pp]
                sub_z = torch.where(take_p, z1, sub_z)
                sub_g = torch.where(take_p, g1, sub_g)
                sub_lp = torch.where(take, lp1, sub_lp)
                sub_log_w = torch.where(ok, new_log_w, sub_log_w)
                sub_rho = torch.where(ok[pp], sub_rho + p1, sub_rho)
                # U-turns of the sub-subtrees ending at this leaf
                if n % 2 == 0:
                    k = bin(n >> 1).count("1")
                    ck_p[k] = torch.where(ok[pp], p1, ck_p[k])
                    ck_rho[k] = torch.where(ok[pp], sub_rho, ck_rho[k])
                else:
                    lo, hi = _ckpt_range(n)
                    for k in range(lo, hi + 1):
                        r = sub_rho - ck_rho[k] + ck_p[k]
                        sub_stop |= ok & self._turning(ck_p[k], p1, r)
            # merge: a subtree that turned or diverged ends the chain's tree
            merge = active & ~sub_stop
            u = torch.rand(nchain, dtype=f64, device=dev)
            take = merge & (torch.log(u) < sub_log_w - log_w)
            take_p = take[pp]
            zs = torch.where(take_p, sub_z, zs)
            gs = torch.where(take_p, sub_g, gs)
            lps = torch.where(take, sub_lp, lps)
            log_w = torch.where(merge, torch.logaddexp(log_w, sub_log_w), log_w)
            m_r = (merge & fwd)[pp]
            m_l = (merge & ~fwd)[pp]
            zr, pr, gr = (torch.where(m_r, ze, zr), torch.where(m_r, pe, pr),
                          torch.where(m_r, ge, gr))
            zl, pl, gl = (torch.where(m_l, ze, zl), torch.where(m_l, pe, pl),
                          torch.where(m_l, ge, gl))
            rho = torch.where(merge[pp], rho + sub_rho, rho)
            depth += merge.to(f64)
            divergent |= active & sub_div
            done |= ~merge | self._turning(pl, pr, rho)

        moved = post.segment_sum((zs != self.z).to(f64)) > 0
        self.z.copy_(zs)
        self.grad.copy_(gs)
        self.lnp.copy_(lps)
        self.accept_prob.copy_(acc_sum / torch.clamp(n_leaf, min=1.0))
        self.tree_depth.copy_(depth)
        self.n_leapfrog.copy_(n_leaf)
        self.divergent.copy_(divergent)
        self.n_accept.add_(moved.to(f64))


def run_nuts(posterior, z0, n_draws, n_warmup=500, step_size=1e-3, max_treedepth=10,
             target_accept=0.8, seed=0, inv_mass=None, dense=False):
    """Warm up and draw ``n_draws`` samples of every chain (the role of
    ``run_lmc``, forcepho/fitting.py:255-367).  Returns (draws
    [n_draws, n_params] in the unconstrained space, the sampler)."""
    s = NUTS(posterior, z0, step_size=step_size, max_treedepth=max_treedepth, seed=seed,
             inv_mass=inv_mass, dense=dense)
    if n_warmup > 0:
        s.adapt(n_warmup, target_accept=target_accept)
    out = torch.empty((int(n_draws),) + tuple(s.z.shape), dtype=torch.float64,
                      device=s.z.device)
    for i in range(int(n_draws)):
        s.trajectory()
        out[i] = s.z
    return out, s
