"""run_opt, run_opt_bounded and run_lmc on a posterior with a closed form:
a correlated Gaussian in a bounded parameter space, written as forcepho's
Posterior (sampling-space lnprob_and_grad with the transform's log
Jacobian).  The optimisers must find the mode; NUTS draws must reproduce the
mean and covariance (CPU, the samplers' torch code on the host)."""
import numpy as np
import pytest
import torch

from forcepho_amd.fitting import Result, run_lmc, run_opt, run_opt_bounded
from forcepho_amd.model import BoundedTransform

MU = np.array([0.3, -0.2, 1.0])
COV = np.array([[0.04, 0.01, 0.0], [0.01, 0.09, -0.02], [0.0, -0.02, 0.16]])
PREC = np.linalg.inv(COV)


class GaussModel:
    """lnp(q) Gaussian; sampled in z with q = lower + sigmoid(z) range."""

    def __init__(self, sampling=True):
        self.transform = BoundedTransform(MU - 6.0, MU + 6.0)
        self.sampling = sampling
        self.ncall = 0

    def lnprob_and_grad(self, z):
        self.ncall += 1
        q = self.transform.transform(z)
        d = q - MU
        lnp = -0.5 * d @ PREC @ d
        g = -(PREC @ d) * self.transform.jacobian(z)
        if self.sampling:
            lnp += self.transform.lndetjac(z)
            g += self.transform.lndetjac_grad(z)
        return lnp, g

    def nll(self, z):
        lnp, g = self.lnprob_and_grad(z)
        return -lnp, -g

    def nll_nograd(self, z):
        return -self.lnprob_and_grad(z)[0]


def test_result_namespace():
    r = Result(a=1, chain=np.zeros((2, 3)))
    assert r.a == 1 and r.chain.shape == (2, 3)


@pytest.mark.parametrize("runner", [run_opt, run_opt_bounded])
def test_optimisers_find_the_mode(runner):
    m = GaussModel(sampling=False)
    res, sci = runner(m, MU + 0.5)
    assert res.chain.shape == (1, 3)
    assert np.allclose(res.chain[0], MU, atol=1e-4), res.chain
    assert res.lnp == pytest.approx(-0.5 * sci.fun)


def test_optimiser_without_gradient():
    m = GaussModel(sampling=False)
    res, _ = run_opt(m, MU + 0.5, jac=False)
    assert np.allclose(res.chain[0], MU, atol=1e-3)


@pytest.mark.parametrize("full", [False, True])
def test_run_lmc_moments(full):
    torch.manual_seed(0)
    m = GaussModel()
    res, sampler, stats = run_lmc(m, MU, n_draws=1500, warmup=[300], full=full,
                                  random_seed=3)
    chain = res.chain
    assert chain.shape == (1500, 3) and len(stats) == 1500
    assert not any(s["diverging"] for s in stats)
    se = np.sqrt(np.diag(COV) / 1500) * 4   # generous: draws are correlated
    assert np.all(np.abs(chain.mean(axis=0) - MU) < 4 * se), chain.mean(axis=0)
    assert np.allclose(np.cov(chain.T), COV, atol=0.03), np.cov(chain.T)
    assert m.ncall > 1500


def test_initial_cov_choices():
    from forcepho_amd.fitting import initial_cov
    assert np.array_equal(initial_cov(3), np.eye(3))
    assert np.array_equal(initial_cov(3, init_cov=COV), COV)
    rng = np.random.default_rng(1)
    tr = rng.multivariate_normal(MU, COV, size=4000).T
    assert np.allclose(initial_cov(3, trace=tr), COV, atol=0.01)
    # a singular covariance is regularised until it has a Cholesky factor
    sing = np.ones((2, 2))
    c = initial_cov(2, init_cov=sing)
    np.linalg.cholesky(c)
    assert np.allclose(c, sing + 1e-3 * np.eye(2))


def test_run_lmc_from_a_trace():
    rng = np.random.default_rng(2)
    m = GaussModel()
    tr = m.transform.inverse_transform(rng.multivariate_normal(MU, COV, size=500)).T
    res, sampler, stats = run_lmc(m, MU, n_draws=200, warmup=[0], trace=tr, random_seed=4)
    assert res.chain.shape == (200, 3)
    assert np.allclose(sampler.inv_mass.cpu().numpy(), np.diag(np.cov(tr)))
