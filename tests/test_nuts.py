"""Lockstep multinomial NUTS on Gaussian targets (host, CPU tensors)."""
import numpy as np
import torch

from forcepho_amd.hmc import segment_sum_index
from forcepho_amd.nuts import NUTS, _ckpt_range, run_nuts


class GaussianPatches:
    """lnp per patch = -1/2 z_k^T P_k z_k over the patch's parameters, with
    P_k block-diagonal per patch (the attributes the samplers read from a
    DevicePosterior)."""

    def __init__(self, precisions):
        self.blocks = [torch.as_tensor(p, dtype=torch.float64) for p in precisions]
        sizes = [b.shape[0] for b in self.blocks]
        theta_off = np.concatenate([[0], np.cumsum(sizes)])
        self.lower = torch.zeros(int(theta_off[-1]), dtype=torch.float64)
        idx, mask = segment_sum_index(theta_off)
        self.seg_idx = torch.as_tensor(idx)
        self.seg_mask = torch.as_tensor(mask, dtype=torch.float64)
        self.param_patch = torch.as_tensor(np.repeat(np.arange(len(sizes)), sizes))
        self.P = torch.block_diag(*self.blocks)
        self.n_evals = 0

    def segment_sum(self, v):
        return (v[self.seg_idx] * self.seg_mask).sum(dim=1)

    def __call__(self, z):
        self.n_evals += 1
        g = -(self.P @ z)
        return 0.5 * self.segment_sum(z * g), g


def test_checkpoint_slots_match_aligned_subtrees():
    """The sub-subtrees ending at leaf n are the aligned blocks of 2^k leaves
    (k >= 1) with (n + 1) % 2^k == 0; each was checkpointed at its first
    leaf f (even), in slot popcount(f >> 1)."""
    for n in range(1, 1024, 2):
        slots = set()
        k = 1
        while (n + 1) % (2 ** k) == 0:
            f = n + 1 - 2 ** k
            slots.add(bin(f >> 1).count("1"))
            k += 1
        lo, hi = _ckpt_range(n)
        assert slots == set(range(lo, hi + 1)), n


def test_nuts_draws_match_uneven_diagonal_targets():
    rng = np.random.default_rng(0)
    sig = [10.0 ** rng.uniform(-2, 0, size=m) for m in (3, 4, 2)]
    post = GaussianPatches([np.diag(1.0 / s ** 2) for s in sig])
    sigma = np.concatenate(sig)
    z0 = torch.as_tensor(rng.normal(size=9) * sigma * 0.1)
    draws, s = run_nuts(post, z0, n_draws=600, n_warmup=300, step_size=1e-3, seed=1)
    d = draws.numpy()
    ratio = d.std(axis=0) / sigma
    assert np.all((ratio > 0.85) & (ratio < 1.15)), ratio
    assert np.all(np.abs(d.mean(axis=0) / sigma) < 0.2)
    # adapted: acceptance statistic near the target, trees of a few levels
    assert np.all((s.accept_prob.numpy() > 0.3))
    assert 1.0 <= float(s.tree_depth.mean()) <= 6.0
    assert not bool(s.divergent.any())


def test_nuts_correlated_target_covariance():
    # one patch with correlation 0.95, one isotropic, unit mass matrix
    c = np.array([[1.0, 0.95], [0.95, 1.0]]) * np.array([[4.0, 2.0], [2.0, 1.0]])
    post = GaussianPatches([np.linalg.inv(c), np.eye(2)])
    z0 = torch.zeros(4, dtype=torch.float64)
    s = NUTS(post, z0, step_size=0.25, seed=5, max_treedepth=8)
    draws = []
    for _ in range(2000):
        s.trajectory()
        draws.append(s.z.clone())
    d = torch.stack(draws).numpy()[100:]
    cov = np.cov(d[:, :2].T)
    assert np.allclose(cov, c, rtol=0.15, atol=0.08), cov
    assert np.allclose(np.cov(d[:, 2:].T), np.eye(2), atol=0.12)
    # the correlated chain needs longer trees than the isotropic one
    assert float(s.tree_depth[0]) >= 1


def test_nuts_many_chains_standard_normal_is_exact():
    """64 one-parameter chains of N(0, 1) with a large and a small step:
    the pooled draws have unit variance and normal kurtosis (a sampler
    that mis-handles sub-tree U-turns or weights drifts here)."""
    for eps in (1.1, 0.2):
        post = GaussianPatches([np.eye(1)] * 64)
        rng = np.random.default_rng(3)
        s = NUTS(post, torch.as_tensor(rng.normal(size=64)), step_size=eps, seed=11)
        draws = []
        for _ in range(150):
            s.trajectory()
            draws.append(s.z.clone())
        d = torch.stack(draws).numpy()[20:].reshape(-1)
        var = d.var()
        kurt = np.mean(d ** 4) / var ** 2
        assert abs(var - 1.0) < 0.08, (eps, var)
        assert abs(kurt - 3.0) < 0.35, (eps, kurt)
        assert abs(d.mean()) < 0.06, (eps, d.mean())


def test_nuts_stops_on_divergence():
    # a step far beyond the stability limit of the integrator diverges
    post = GaussianPatches([np.eye(2) * 1e4])
    s = NUTS(post, torch.full((2,), 0.01, dtype=torch.float64), step_size=5.0, seed=2)
    z_before = s.z.clone()
    s.trajectory()
    assert bool(s.divergent[0])
    assert torch.isfinite(s.lnp).all()
    assert float(s.tree_depth[0]) <= 1
    # a diverged first subtree leaves the chain where it was
    if float(s.tree_depth[0]) == 0:
        assert torch.equal(s.z, z_before)


def test_dense_mass_blocks_match_explicit_matrix():
    from forcepho_amd.hmc import DenseMass
    rng = np.random.default_rng(4)
    sizes = (3, 1, 2)
    blocks = []
    for m in sizes:
        a = rng.normal(size=(m, m))
        blocks.append(a @ a.T + m * np.eye(m))
    post = GaussianPatches([np.eye(m) for m in sizes])
    mmax = max(sizes)
    pad = np.stack([np.pad(b, ((0, mmax - b.shape[0]), (0, mmax - b.shape[0])))
                    for b in blocks])
    dm = DenseMass(post, pad)
    full = torch.block_diag(*[torch.as_tensor(b) for b in blocks])
    p = torch.as_tensor(rng.normal(size=6))
    assert torch.allclose(dm.velocity(p), full @ p)
    # momenta ~ N(0, M) with M the inverse of the blocks
    torch.manual_seed(0)
    draws = torch.stack([dm.draw(p) for _ in range(20000)])
    cov = np.cov(draws.numpy().T)
    m_full = np.linalg.inv(full.numpy())
    assert np.allclose(cov, m_full, atol=0.05 * np.abs(m_full).max())


def test_dense_adaptation_learns_correlation_and_shortens_trees():
    c = np.array([[4.0, 1.9], [1.9, 1.0]])          # correlation 0.95
    post = GaussianPatches([np.linalg.inv(c)])
    z0 = torch.zeros(2, dtype=torch.float64)
    depth = {}
    for dense in (False, True):
        draws, s = run_nuts(post, z0, n_draws=800, n_warmup=400, step_size=0.1, seed=7,
                            dense=dense)
        d = draws.numpy()
        assert np.allclose(np.cov(d.T), c, rtol=0.2, atol=0.1), (dense, np.cov(d.T))
        depth[dense] = []
        for _ in range(200):
            s.trajectory()
            depth[dense].append(float(s.tree_depth[0]))
        if dense:
            minv = s.dense.minv[0].numpy()
            assert np.allclose(minv, c, rtol=0.35, atol=0.2), minv
    # with the correlation in the mass matrix the target is isotropic
    assert np.mean(depth[True]) < np.mean(depth[False])
