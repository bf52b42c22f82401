"""Configuration 3 at its stated scope: configuration 2's patch (8 bands x 4
exposures x 128^2 pixels, 20 sources) behind the torch.autograd.Function,
280 parameters.

Against the committed FP64 oracle fixtures (tests/golden/config3_theta.npz,
written by tests/golden/make_scale_fixtures.py with scipy splines, FP64
source parameters and the C restatement):
  * lnlike and the 280-component adjoint of all 50 seeded proposals;
  * FP64 central finite differences (step 1e-5 relative) of 28 components
    (two sources of every parameter kind) at 3 proposals;
  * a 10-step leapfrog trajectory (step 1e-3, unit mass) in the sampling
    space of the bounded transform from a fixed (z, p), against the same
    trajectory driven by the oracle's gradient in FP64
    (forcepho/model.py:225-257 and forcepho/fitting.py:303-327 are the
    reference's lnp/gradient and its sampler caller).
Then 1000 leapfrog steps (100 trajectories of 10) in a captured HIP graph on
the 280-parameter patch."""
import os

import numpy as np
import pytest
import torch

from forcepho_amd import synth
from forcepho_amd.autograd import LnLike
from forcepho_amd.batch import DeviceSplines, PatchBatch
from forcepho_amd.hmc import HMC, DevicePosterior
from helpers import CHI2_RTOL, Fixture, load_fixture, make_config_patch

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "config3_theta.npz")


@pytest.fixture(scope="module")
def c3(hip):
    fx = load_fixture(FIXTURE)
    patch, scene, theta = make_config_patch(2)
    assert Fixture.sha256(patch.data) == fx["data_sha256"], \
        "configuration 2's data differ from the fixture's; regenerate the fixtures"
    patch.send_to_device()
    batch = PatchBatch([patch])
    splines = DeviceSplines()
    assert batch.n_theta == 280
    yield patch, theta, batch, splines, fx
    batch.free()
    patch.free()


def _within(g, g_ref, scale):
    """Stated gradient tolerance: relative 1e-4 or absolute 1e-6 * scale."""
    err = np.abs(g - g_ref)
    return (err <= 1e-4 * np.abs(g_ref)) | (err <= 1e-6 * scale)


def _lnlike_and_grad(batch, splines, t):
    tt = torch.as_tensor(t, dtype=torch.float64, device="cuda").requires_grad_(True)
    ll = LnLike.apply(tt, batch, splines)
    ll.sum().backward()
    return ll.item(), tt.grad.cpu().numpy()


def test_adjoint_of_50_proposals_against_the_fp64_oracle(c3):
    patch, theta, batch, splines, fx = c3
    worst_ll, worst_g = 0.0, 0.0
    for k, t in enumerate(fx["theta"]):
        ll, g = _lnlike_and_grad(batch, splines, t)
        ll_ref, g_ref = fx["lnlike"][k], fx["grad"][k]
        scale = np.abs(g_ref).max()
        ok = _within(g, g_ref, scale)
        worst_ll = max(worst_ll, abs(ll - ll_ref) / abs(ll_ref))
        worst_g = max(worst_g, float(np.max(np.abs(g - g_ref) / scale)))
        assert abs(ll - ll_ref) <= CHI2_RTOL * abs(ll_ref), (k, ll, ll_ref)
        assert ok.all(), (k, np.flatnonzero(~ok)[:5])
    print(f"50 proposals: lnlike rel err <= {worst_ll:.2e}, "
          f"gradient err <= {worst_g:.2e} of max|grad|", flush=True)


def test_adjoint_against_fp64_central_differences(c3):
    patch, theta, batch, splines, fx = c3
    for k in range(len(fx["fd"])):
        _, g = _lnlike_and_grad(batch, splines, fx["theta"][k])
        comps, fd = fx["fd_components"][k], fx["fd"][k]
        scale = np.abs(fx["grad"][k]).max()
        ok = _within(g[comps], fd, scale)
        rel = np.abs(g[comps] - fd) / np.maximum(np.abs(fd), 1e-6 * scale)
        print(f"proposal {k}: 28 components vs central differences, max rel err "
              f"{rel.max():.2e}", flush=True)
        assert ok.all(), (k, comps[~ok], g[comps][~ok], fd[~ok])


def test_leapfrog_trajectory_against_fp64_leapfrog(c3):
    patch, theta, batch, splines, fx = c3
    lo, hi = synth.theta_bounds(theta, 8)
    post = DevicePosterior(batch, splines, lo, hi)
    z0 = torch.as_tensor(fx["lf_z0"], dtype=torch.float64, device="cuda")
    p0 = torch.as_tensor(fx["lf_p0"], dtype=torch.float64, device="cuda")
    eps = float(fx["lf_eps"])
    hmc = HMC(post, z0, step_size=eps, n_steps=int(fx["lf_steps"]), jitter=0.0)
    lnp0 = float(hmc.lnp[0])
    z, p, lnp, grad = hmc.leapfrog(z0, p0, hmc.grad, hmc.eps)
    z, p, lnp = z.cpu().numpy(), p.cpu().numpy(), float(lnp[0])
    dz_ref = fx["lf_z"] - fx["lf_z0"]
    dz = z - fx["lf_z0"]
    ez = np.abs(dz - dz_ref).max() / np.abs(dz_ref).max()
    ep = np.abs(p - fx["lf_p"]).max() / np.abs(fx["lf_p"]).max()
    print(f"leapfrog: lnp {lnp0:.6f} -> {lnp:.6f} (oracle {float(fx['lf_lnp']):.6f}); "
          f"displacement err {ez:.2e}, momentum err {ep:.2e} (of their max)", flush=True)
    assert abs(lnp0 - fx["lf_lnp0"]) <= CHI2_RTOL * abs(fx["lf_lnp0"])
    assert abs(lnp - fx["lf_lnp"]) <= CHI2_RTOL * abs(fx["lf_lnp"])
    # gradients off by their whole stated tolerance (1e-4 relative on every
    # component) move this trajectory by 1.2e-4 of its displacement (FP64
    # emulation on the CPU: the trajectory does not amplify errors), so a
    # gradient within tolerance keeps it within 2e-4
    assert ez <= 2e-4 and ep <= 2e-4


def test_1000_leapfrog_steps_in_a_graph_on_280_parameters(c3):
    patch, theta, batch, splines, fx = c3
    lo, hi = synth.theta_bounds(theta, 8)
    post = DevicePosterior(batch, splines, lo, hi)
    z0 = post.inverse_transform(np.clip(theta, lo + 1e-9, hi - 1e-9))
    hmc = HMC(post, z0, step_size=1e-3, n_steps=10, seed=3)
    hmc.capture(warmup=1)
    hmc.n_accept.zero_()
    hmc.run(100)
    torch.cuda.synchronize()
    acc = hmc.n_accept.item() / 100.0
    print(f"100 trajectories x 10 steps: acceptance {acc:.2f}", flush=True)
    assert torch.isfinite(hmc.z).all() and torch.isfinite(hmc.lnp).all()
    assert acc > 0.0
    # the chain's cached lnp and gradient are those of its position
    lnp, grad = post(hmc.z)
    assert torch.equal(lnp, hmc.lnp) and torch.equal(grad, hmc.grad)
