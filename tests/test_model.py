"""The posterior objects of forcepho_amd.model on the CPU, with an analytic
likelihood standing in for the proposer: composition of likelihood, prior
and transform; reuse of the last evaluation; the central-difference check;
argument errors; a log-Jacobian that stays finite far against a bound."""
import numpy as np
import pytest

from forcepho_amd.model import BoundedTransform, Posterior, Transform

MU = np.array([0.3, -1.2, 2.0])
PREC = np.array([[4.0, 1.0, 0.0], [1.0, 3.0, 0.5], [0.0, 0.5, 2.0]])


class Quadratic(Posterior):
    """lnlike(q) = -1/2 (q - mu)^T P (q - mu)."""

    def __init__(self, **kw):
        super().__init__(**kw)
        self.calls = 0

    def lnlike(self, q):
        self.calls += 1
        d = q - MU
        return -0.5 * d @ PREC @ d, -(PREC @ d)


def test_composition_against_closed_form():
    lo, hi = MU - 3.0, MU + 4.0
    m = Quadratic(lower=lo, upper=hi)
    m._lnpriorfn = lambda q: (-0.5 * float(q @ q), -q)
    z = np.array([0.4, -0.7, 1.1])
    y = 1.0 / (1.0 + np.exp(-z))
    q = lo + (hi - lo) * y
    d = q - MU
    jac = (hi - lo) * y * (1 - y)
    lnp_ref = -0.5 * d @ PREC @ d - 0.5 * q @ q + np.sum(np.log(jac))
    grad_ref = (-(PREC @ d) - q) * jac + (1 - 2 * y)
    lnp, grad = m.lnprob_and_grad(z)
    assert np.isclose(lnp, lnp_ref, rtol=1e-13)
    assert np.allclose(grad, grad_ref, rtol=1e-12)
    # nll divides by ndof and flips the sign; not sampling drops the Jacobian term
    m.ndof = 2.0
    v, g = m.nll(z)
    assert np.isclose(v, -lnp / 2.0) and np.allclose(g, -grad / 2.0)
    m2 = Quadratic(lower=lo, upper=hi)
    m2.sampling = False
    assert np.isclose(m2.lnprob(z), -0.5 * d @ PREC @ d, rtol=1e-13)


def test_last_evaluation_is_reused():
    m = Quadratic(transform=Transform(3))
    z = np.array([1.0, 2.0, 3.0])
    m.lnprob(z)
    m.lnprob_grad(z)
    m.nll(z)
    m.nll_nograd(z.copy())
    assert m.calls == 1 and m.ncall == 1
    m.lnprob(z + 1e-9)
    assert m.calls == 2
    assert np.array_equal(m._z, z + 1e-9) and m._lnp == m.lnprob(z + 1e-9)


def test_check_grad_matches_analytic():
    m = Quadratic(lower=MU - 2.0, upper=MU + 2.0)
    analytic, numeric = m.check_grad(np.array([0.2, -0.3, 0.9]), delta=1e-6)
    assert np.allclose(analytic, numeric, rtol=1e-6, atol=1e-8)


def test_argument_errors():
    with pytest.raises(ValueError):
        Quadratic()                                          # neither
    with pytest.raises(ValueError):
        Quadratic(transform=Transform(3), lower=MU, upper=MU + 1)   # both
    with pytest.raises(ValueError):
        Quadratic(transform=Transform(3), colour="blue")     # not understood
    with pytest.raises(ValueError):
        BoundedTransform(np.zeros(3), np.ones(2))


def test_bounded_transform_log_jacobian_far_from_the_centre():
    t = BoundedTransform(np.array([0.0, -1.0]), np.array([2.0, 5.0]))
    z = np.array([0.3, -0.8])
    direct = np.sum(np.log(t.jacobian(z)))
    assert np.isclose(t.lndetjac(z), direct, rtol=1e-13)
    # where y (1 - y) underflows the product form is -inf; the log form is not
    far = np.array([800.0, -900.0])
    assert np.isfinite(t.lndetjac(far))
    assert np.isclose(t.lndetjac(far), np.log(2.0) + np.log(6.0) - 1700.0, rtol=1e-12)
    q = t.transform(z)
    assert np.allclose(t.inverse_transform(q), z, rtol=1e-12)
