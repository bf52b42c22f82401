"""HIP-graph capture of the device posterior and of HMC trajectories."""
import numpy as np
import pytest
import torch

from forcepho_amd import synth
from forcepho_amd.batch import DeviceSplines, PatchBatch
from forcepho_amd.hmc import HMC, DevicePosterior
from helpers import make_config_patch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def post(hip):
    patch, scene, theta = make_config_patch(1)
    patch.send_to_device()
    batch = PatchBatch([patch])
    splines = DeviceSplines()
    lo, hi = synth.theta_bounds(theta, 1)
    p = DevicePosterior(batch, splines, lo, hi)
    yield p, theta
    batch.free()
    patch.free()


def test_graph_replays_match_eager(post):
    p, theta = post
    z = p.inverse_transform(theta).contiguous()
    lnp_e, grad_e = p(z)
    lnp_e, grad_e = lnp_e.clone(), grad_e.clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        p(z)
    torch.cuda.current_stream().wait_stream(s)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        lnp_g, grad_g = p(z)
    for k in range(4):
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(lnp_g, lnp_e), k
        assert torch.equal(grad_g, grad_e), k
        # eager work and a synchronisation between replays
        l2, g2 = p(z + 0.01)
        torch.cuda.synchronize()
        assert not torch.equal(l2, lnp_e)


def test_hmc_1000_leapfrog_steps_in_graph(post):
    p, theta = post
    z0 = p.inverse_transform(theta)
    hmc = HMC(p, z0, step_size=1e-3, n_steps=10, seed=1)
    hmc.capture()
    hmc.run(100)          # 100 trajectories x 10 steps
    torch.cuda.synchronize()
    assert torch.isfinite(hmc.lnp).all() and torch.isfinite(hmc.z).all()
    assert hmc.n_accept.sum().item() > 0
    # the state stays near the truth in the constrained space
    q = p.transform(hmc.z).cpu().numpy()
    assert np.all(np.isfinite(q))


def test_adapt_then_capture(post):
    """Warm-up on the device posterior, then graph-captured sampling with the
    adapted per-patch step size and diagonal mass matrix."""
    p, theta = post
    z0 = p.inverse_transform(theta)
    hmc = HMC(p, z0, step_size=1e-5, n_steps=10, seed=2)
    hmc.adapt(150, target_accept=0.8)
    assert torch.isfinite(hmc.eps_patch).all() and (hmc.eps_patch > 1e-5).all()
    assert torch.isfinite(hmc.inv_mass).all() and (hmc.inv_mass > 0).all()
    hmc.capture(warmup=2)
    hmc.run(60)
    torch.cuda.synchronize()
    acc = (hmc.n_accept / 62.0).cpu().numpy()
    assert np.all((acc > 0.3) & (acc <= 1.0)), acc
    assert torch.isfinite(hmc.z).all()


def test_nuts_on_device_posterior(hip):
    """NUTS transitions of several patches in lockstep on the device:
    warm-up, then draws whose trees stop by U-turn (not at the depth cap)."""
    from forcepho_amd.nuts import run_nuts
    from helpers import make_small_patches
    items = make_small_patches()
    patches = [p for p, _, _ in items]
    for p in patches:
        p.send_to_device()
    batch = PatchBatch(patches)
    splines = DeviceSplines()
    theta = np.concatenate([t for _, _, t in items])
    lo, hi = [], []
    for p, _, t in items:
        a, b = synth.theta_bounds(t, len(p.band_N))
        lo.append(a)
        hi.append(b)
    post = DevicePosterior(batch, splines, np.concatenate(lo), np.concatenate(hi))
    z0 = post.inverse_transform(theta)
    try:
        draws, s = run_nuts(post, z0, n_draws=20, n_warmup=60, step_size=1e-4,
                            max_treedepth=8, seed=4)
        torch.cuda.synchronize()
        assert torch.isfinite(draws).all() and torch.isfinite(s.lnp).all()
        assert (s.tree_depth >= 1).all() and (s.tree_depth < 8).all()
        assert (s.n_accept > 0).all()
        assert not bool(s.divergent.any())
        q = post.transform(s.z).cpu().numpy()
        assert np.all(np.isfinite(q))
    finally:
        batch.free()
        for p in patches:
            p.free()


def test_dense_mass_adapt_then_capture(post):
    """Dense (per-patch full) mass matrix: warm-up, then graph-captured
    trajectories, with every replay's state finite and the chain moving."""
    p, theta = post
    z0 = p.inverse_transform(theta)
    hmc = HMC(p, z0, step_size=1e-5, n_steps=10, seed=3, dense=True)
    hmc.adapt(150, target_accept=0.8)
    minv = hmc.dense.minv
    assert torch.isfinite(minv).all()
    assert (torch.linalg.eigvalsh(minv) > 0).all()
    hmc.capture(warmup=2)
    hmc.run(40)
    torch.cuda.synchronize()
    acc = (hmc.n_accept / 42.0).cpu().numpy()
    assert np.all((acc > 0.3) & (acc <= 1.0)), acc
    assert torch.isfinite(hmc.z).all() and torch.isfinite(hmc.lnp).all()
