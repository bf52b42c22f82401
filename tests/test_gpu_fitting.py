"""The host-driven posterior over the HIP proposer (FastPosterior, as
forcepho's model.py) and the fitting drivers on it: lnp and its gradient in
the sampling space against the FP64 oracle, BFGS from a perturbed start, and
a short NUTS chain."""
import numpy as np
import pytest

import oracle
from forcepho_amd import synth
from forcepho_amd.fitting import run_lmc, run_opt
from forcepho_amd.model import FastPosterior
from helpers import CHI2_RTOL, make_config_patch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1(hip):
    patch, scene, theta = make_config_patch(1)
    patch.send_to_device()
    lo, hi = synth.theta_bounds(theta, 1)
    yield patch, scene, theta, lo, hi
    patch.free()


def _model(c1, sampling):
    patch, scene, theta, lo, hi = c1
    return FastPosterior(patch.get_proposer(), patch, scene=scene, lower=lo, upper=hi,
                         sampling=sampling)


def test_fastposterior_against_oracle(c1):
    patch, scene, theta, lo, hi = c1
    m = _model(c1, sampling=False)
    for k, t in enumerate(synth.make_proposals(theta, 1, 5, config=3, patch_id=7)):
        t = np.clip(t, lo + 1e-9 * (hi - lo), hi - 1e-9 * (hi - lo))
        z = m.transform.inverse_transform(t)
        lnp, g = m.lnprob_and_grad(z)
        ll_ref, gq_ref = oracle.lnlike_theta_np(patch, t)
        g_ref = gq_ref * m.transform.jacobian(z)
        print(f"proposal {k}: lnp {lnp:.6f} oracle {ll_ref:.6f}", flush=True)
        assert abs(lnp - ll_ref) <= CHI2_RTOL * abs(ll_ref)
        err = np.abs(g - g_ref)
        assert np.all((err <= 1e-4 * np.abs(g_ref)) | (err <= 1e-6 * np.abs(g_ref).max()))


def test_run_opt_climbs_to_the_truth(c1):
    patch, scene, theta, lo, hi = c1
    m = _model(c1, sampling=False)
    start = synth.make_proposals(theta, 1, 1, config=3, patch_id=8)[0]
    start = np.clip(start, lo + 1e-6 * (hi - lo), hi - 1e-6 * (hi - lo))
    lnp0 = m.lnprob(m.transform.inverse_transform(start))
    res, sci = run_opt(m, start, maxiter=200)
    lnp1 = m.lnprob(m.transform.inverse_transform(res.chain[0]))
    lnp_truth = m.lnprob(m.transform.inverse_transform(theta))
    print(f"lnp start {lnp0:.3f} optimum {lnp1:.3f} truth {lnp_truth:.3f} "
          f"({sci.nit} iterations)", flush=True)
    assert lnp1 > lnp0
    # data = model(truth) + noise: the optimum is at least as good as the truth
    assert lnp1 >= lnp_truth - 1.0


def test_run_lmc_short_chain(c1):
    patch, scene, theta, lo, hi = c1
    m = _model(c1, sampling=True)
    res, sampler, stats = run_lmc(m, theta, n_draws=10, warmup=[20], random_seed=5)
    assert res.chain.shape == (10, len(theta))
    assert np.all(np.isfinite(res.chain))
    assert np.all((res.chain >= lo) & (res.chain <= hi))
    assert len(stats) == 10 and res.n_call > 10
    assert np.ptp(res.chain, axis=0).max() > 0
